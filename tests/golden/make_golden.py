"""Writes the golden fixtures in tests/golden/ from the CPU oracle.

The reference (Rust) cannot be built or run here and ships no golden vectors,
so each fixture is produced by the oracle and cross-checked, before it is
written, against the independent Kronecker-product model in
tests/numpy_model.py (amplitudes, expectation values) and against closed
forms (Grover).  Run from the repository root:  python tests/golden/make_golden.py
"""
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from kbw_amd import circuits  # noqa: E402
from tests import numpy_model as nm  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def grover12():
    n = 12
    ops, target = circuits.grover(n, 50, seed=0)
    s = oracle.DenseV2(n)
    s.run(ops)
    amps = s.state()
    theta = math.asin(1 / math.sqrt(1 << n))
    assert abs(abs(amps[target]) - abs(math.sin(101 * theta))) < 1e-12
    u = np.random.default_rng(1).random(n)
    out = s.measure(list(range(n)), u)
    np.savez(os.path.join(HERE, "grover12.npz"), ops=ops.view(np.uint8), target=target, amps=amps,
             uniforms=u, measured=np.uint64(out), post=s.state())


def brickwork10():
    n = 10
    ops, ham = circuits.brickwork(n, 6, seed=3, nstrings=16)
    s = oracle.DenseV2(n)
    s.run(ops)
    amps = s.state()
    assert np.max(np.abs(amps - nm.run_ops(n, ops))) < 1e-12
    e = s.exp_value(ham.pauli_strings, ham.coefficients)
    assert abs(e - nm.expectation(amps, n, ham)) < 1e-12
    pa, qu, off, co = oracle.flatten_hamiltonian(ham.pauli_strings, ham.coefficients)
    np.savez(os.path.join(HERE, "brickwork10.npz"), ops=ops.view(np.uint8), amps=amps, paulis=pa,
             qubits=qu, offsets=off, coeffs=co, exp_value=e)


def sample9():
    n = 9
    rng = np.random.default_rng(17)
    ops = circuits.random_circuit(rng, n, 60)
    s = oracle.DenseV2(n)
    s.run(ops)
    qubits = [int(q) for q in rng.permutation(n)]
    u = rng.random(4000)
    out = s.sample_hier(qubits, len(u), u)
    seq = s.sample_sequential(qubits, len(u), u)
    assert np.mean(out == seq) > 0.999
    np.savez(os.path.join(HERE, "sample9.npz"), ops=ops.view(np.uint8), qubits=np.array(qubits, np.uint32),
             uniforms=u, outcomes=out)


def block8():
    n, L = 8, 5
    rng = np.random.default_rng(23)
    ops = circuits.random_circuit(rng, n, 80, max_controls=2)
    b = oracle.Block(n, L)
    for o in ops:
        m = int(o["ctrl_mask"])
        b.gate(int(o["kind"]), float(o["angle"]), int(o["target"]), [q for q in range(n) if m >> q & 1])
    flat = oracle.DenseV2(n)
    flat.run(ops)
    assert np.array_equal(b.logical_state(), flat.state())
    k, re, im = b.dump(list(range(n)))
    np.savez(os.path.join(HERE, "block8.npz"), ops=ops.view(np.uint8), block_size=L,
             physical=b.physical_state(), l2p=b.qubit_map(), swaps=b.swap_count(), dump_keys=k)


if __name__ == "__main__":
    grover12()
    brickwork10()
    sample9()
    block8()
    print("fixtures written to", HERE)
