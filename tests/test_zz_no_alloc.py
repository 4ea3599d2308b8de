"""kbwh_exp_value_noalloc against kbwh_exp_value, through the C ABI.

The allocation-free variant keeps its block partials, per-string results and
string table in one fixed device array per GPU instead of the handle's
buffers.  It runs the same kernels in the same reduction order, so its value
must equal kbwh_exp_value's bit for bit.  Calls come in a row on one GPU, the
later ones with smaller inputs, so a value left in the shared array by an
earlier call would show up as a difference.
"""
import ctypes as C

import numpy as np
import pytest

from kbw_amd import DenseHip, Hamiltonian, _lib, circuits

pytestmark = pytest.mark.gpu


def _call(sim, fn, ham):
    paulis, qubits, offsets, coeffs = ham.flatten()
    out = C.c_double()
    rc = getattr(sim._lib, fn)(sim.handle, _lib.ptr(paulis), _lib.ptr(qubits), _lib.ptr(offsets),
                               _lib.ptr(coeffs), len(coeffs), C.byref(out))
    return rc, out.value


def _both(sim, ham):
    rc0, ref = _call(sim, "kbwh_exp_value", ham)
    rc1, got = _call(sim, "kbwh_exp_value_noalloc", ham)
    assert rc0 == 0 and rc1 == 0, (rc0, rc1, sim._lib.kbwh_last_error())
    return ref, got


def test_noalloc_exp_value_equals_exp_value_on_calls_in_a_row():
    rng = np.random.default_rng(4242)
    # 1: a 20-qubit flat state, 70 strings in grouped sweeps plus two strings
    # too wide for a tile (summed alone)
    n = 20
    sim = DenseHip(n)
    sim.run(circuits.random_circuit(rng, n, 150), mode=_lib.RUN_TILED)
    ham = circuits.random_hamiltonian(rng, n, 70, max_weight=4)
    strings = ham.pauli_strings + [[("X", q) for q in range(3, 15)], [("Y", 19), ("X", 4)] + [("X", q) for q in range(6, 16)]]
    big = Hamiltonian(strings, ham.coefficients + [0.3, -0.6])
    ref, got = _both(sim, big)
    assert got == ref, (got, ref)
    first = got
    # 2: a smaller input on the same GPU right after: 12 qubits, 5 strings
    small = DenseHip(12)
    small.run(circuits.random_circuit(rng, 12, 60))
    h5 = circuits.random_hamiltonian(rng, 12, 5, max_weight=3)
    ref2, got2 = _both(small, h5)
    assert got2 == ref2, (got2, ref2)
    # 3: smaller still, a Block state (one reduction per string, replayed swaps)
    blk = DenseHip(9, block=True, block_size=6)
    blk.run(circuits.random_circuit(rng, 9, 40))
    h3 = circuits.random_hamiltonian(rng, 9, 3, max_weight=3)
    ref3, got3 = _both(blk, h3)
    assert got3 == ref3, (got3, ref3)
    # the first state again: the same bits as before
    _, again = _call(sim, "kbwh_exp_value_noalloc", big)
    assert again == first
    for s in (sim, small, blk):
        s.close()


def test_noalloc_exp_value_limits():
    sim = DenseHip(11)
    too_many = Hamiltonian([[("Z", k % 11)] for k in range(4097)], [1.0] * 4097)
    rc, _ = _call(sim, "kbwh_exp_value_noalloc", too_many)
    assert rc == _lib.ERR_CAPACITY
    ok = Hamiltonian([[("Z", k % 11)] for k in range(4096)], [1.0] * 4096)
    ref, got = _both(sim, ok)
    assert got == ref == pytest.approx(4096.0)
    sim.close()
    q = DenseHip(5, queue=True)
    q.hadamard(0)
    rc, _ = _call(q, "kbwh_exp_value_noalloc", Hamiltonian([[("X", 0)]], [1.0]))
    assert rc == _lib.ERR_INVALID_ARGUMENT
    q.synchronize()
    rc, v = _call(q, "kbwh_exp_value_noalloc", Hamiltonian([[("X", 0)]], [1.0]))
    assert rc == 0 and v == pytest.approx(1.0, abs=1e-15)
    q.close()
