"""Batched adjoint gradients (kbwh_adjoint_batch): many parameter sets of one
circuit in one sweep, against the single-set adjoint gradient, the oracle's
parameter-shift rule and FP64 central finite differences.

The sets lie one after another as one state whose top qubits select the set,
so a wrong set offset, a wrong set count in a pass or a reduction mixing sets
shows as a set whose energy or gradient differs from its own single-set run.
Per-set values agree with the single-set path to rounding (coefficients come
from the device's sincos there, the host's libm here): 1e-11.
"""
import math

import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, Hamiltonian, KbwhError, circuits

pytestmark = pytest.mark.gpu

TOL = 1e-11


def mixed_circuit(rng, n, layers):
    """Parameterised RX, RY, RZ and P gates, some controlled, fixed gates
    between them, and a parameter shared by two gates."""
    ol = circuits.OpList()
    np_ = 0
    for _ in range(layers):
        for q in range(n):
            kind = (5, 6, 7, 4)[int(rng.integers(0, 4))]
            if rng.random() < 0.25:
                c = int(rng.integers(0, n - 1))
                c = c if c < q else c + 1
                ol.add(kind, q, [c], param=np_)
            else:
                ol.add(kind, q, param=np_)
            np_ += 1
        for q in range(0, n - 1, 2):
            ol.add(2, q + 1, [q])  # CZ
        ol.add(3, int(rng.integers(0, n)))  # H
        ol.add(0, int(rng.integers(0, n)))  # X
    ol.add(6, n - 1, param=0)  # a parameter used twice
    return ol.array(), np_


def wide_hamiltonian(rng, n):
    ham = circuits.random_hamiltonian(rng, n, 10, max_weight=4)
    strings = ham.pauli_strings + [[("X", q) for q in range(2, min(n, 14))], [("Y", 0), ("Y", n - 1)]]
    return Hamiltonian(strings, ham.coefficients + [0.3, -0.45])


@pytest.mark.parametrize("n,nsets", [(11, 1), (11, 5), (13, 8), (16, 3)])
def test_batch_equals_single_sets(n, nsets):
    print(f"adjoint batch n={n} sets={nsets}", flush=True)
    rng = np.random.default_rng(100 * n + nsets)
    ops, nparams = mixed_circuit(rng, n, 3)
    ham = wide_hamiltonian(rng, n)
    sets = rng.uniform(-math.pi, math.pi, size=(nsets, nparams))
    sim = DenseHip(n)
    energies, grads = sim.adjoint_batch(ops, sets, ham)
    for k in range(nsets):
        e, g = sim.adjoint_gradient(ops, sets[k], ham)
        assert abs(energies[k] - e) <= TOL, (k, energies[k], e)
        assert np.max(np.abs(grads[k] - g)) <= TOL, (k, np.max(np.abs(grads[k] - g)))
    sim.close()


def test_batch_256_sets_16_qubits_against_oracle_and_fd():
    """The 16-qubit, 12-layer RY/RZ + CZ-ring ansatz with an XXZ chain over
    256 parameter sets: every set's energy and eight sets' gradients against
    the single-set adjoint; set 37 against the oracle's shift rule and central
    differences (step 1e-6) on six parameters."""
    n, nsets = 16, 256
    ops, params = circuits.hea(n, 12, seed=9)
    ham = circuits.xxz(n, seed=10)
    rng = np.random.default_rng(11)
    sets = rng.uniform(-math.pi, math.pi, size=(nsets, len(params)))
    sim = DenseHip(n)
    energies, grads = sim.adjoint_batch(ops, sets, ham)
    print("batch done", flush=True)
    for k in range(nsets):
        e, g = (sim.adjoint_gradient(ops, sets[k], ham) if k % 32 == 5 else (None, None))
        if e is None:
            continue
        assert abs(energies[k] - e) <= TOL
        assert np.max(np.abs(grads[k] - g)) <= TOL
    single_e = [sim.adjoint_batch(ops, sets[k:k + 1], ham)[0][0] for k in range(0, nsets, 17)]
    assert np.max(np.abs(np.array(single_e) - energies[::17])) <= TOL
    k = 37
    idx = [0, 51, 130, 255, 300, 383]
    p = sets[k]

    def energy(pp):
        o = oracle.DenseV2(n)
        o.run(ops, pp)
        return o.exp_value(ham.pauli_strings, ham.coefficients)

    assert abs(energies[k] - energy(p)) <= 1e-12
    for j in idx:
        pp = p.copy()
        pp[j] += math.pi / 2
        ep = energy(pp)
        pp[j] -= math.pi
        em = energy(pp)
        shift = 0.5 * (ep - em)
        assert abs(grads[k, j] - shift) <= 1e-10, (j, grads[k, j], shift)
        h = 1e-6
        pp = p.copy()
        pp[j] += h
        fp = energy(pp)
        pp[j] -= 2 * h
        fm = energy(pp)
        fd = (fp - fm) / (2 * h)
        assert abs(grads[k, j] - fd) <= 1e-8, (j, grads[k, j], fd)
    sim.close()


def test_batch_small_states_and_errors():
    n = 8
    rng = np.random.default_rng(5)
    ops, nparams = mixed_circuit(rng, n, 2)
    ham = circuits.tfim(n, seed=2)
    sets = rng.uniform(-math.pi, math.pi, size=(3, nparams))
    sim = DenseHip(n)
    energies, grads = sim.adjoint_batch(ops, sets, ham)
    for k in range(3):
        e, g = sim.adjoint_gradient(ops, sets[k], ham)
        assert energies[k] == e and np.array_equal(grads[k], g)
    with pytest.raises(KbwhError):
        sim.adjoint_batch(ops, np.zeros((0, nparams)), ham)
    bad = ops.copy()
    bad[0]["param"] = nparams + 3
    with pytest.raises(KbwhError):
        DenseHip(12).adjoint_batch(bad, np.zeros((2, nparams)), ham)
    sim.close()


def test_batch_compiled_equals_generic():
    """Parameter sets on compiled forward and reverse kernels (per-set
    records) against the generic kernels, 16 qubits x 64 sets."""
    print("adjoint batch compiled", flush=True)
    n, nsets = 16, 64
    ops, params = circuits.hea(n, 6, seed=19)
    ham = wide_hamiltonian(np.random.default_rng(20), n)
    rng = np.random.default_rng(21)
    sets = rng.uniform(-math.pi, math.pi, size=(nsets, len(params)))
    sim = DenseHip(n)
    e0, g0 = sim.adjoint_batch(ops, sets, ham)
    sim.set_compiled(True)
    e1, g1 = sim.adjoint_batch(ops, sets, ham)
    assert np.max(np.abs(e1 - e0)) <= 1e-12
    assert np.max(np.abs(g1 - g0)) <= 1e-12
    sim.close()
