"""Adjoint gradients (kbwh_adjoint_gradient) against parameter shift on the
oracle, central finite differences, the GPU plain path, and torch autograd."""
import math

import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, circuits
from kbw_amd import autograd as kag

pytestmark = pytest.mark.gpu


def oracle_energy(n, ops, params, ham):
    o = oracle.DenseV2(n)
    o.run(ops, params)
    return o.exp_value(ham.pauli_strings, ham.coefficients)


def oracle_shift(n, ops, params, ham):
    g = np.zeros(len(params))
    for j in range(len(params)):
        p = params.copy()
        p[j] += math.pi / 2
        ep = oracle_energy(n, ops, p, ham)
        p[j] -= math.pi
        em = oracle_energy(n, ops, p, ham)
        g[j] = 0.5 * (ep - em)
    return g


@pytest.mark.parametrize("n,layers,ham_kind", [(6, 2, "tfim"), (10, 3, "xxz"), (13, 2, "tfim")])
def test_adjoint_matches_oracle_parameter_shift(n, layers, ham_kind):
    print(f"adjoint n={n}", flush=True)
    ops, params = circuits.hea(n, layers, seed=n)
    ham = circuits.tfim(n, seed=1) if ham_kind == "tfim" else circuits.xxz(n, seed=2)
    sim = DenseHip(n)
    e, g = sim.adjoint_gradient(ops, params, ham)
    assert e == pytest.approx(oracle_energy(n, ops, params, ham), abs=1e-12)
    ref = oracle_shift(n, ops, params, ham)
    assert np.max(np.abs(g - ref)) < 1e-10
    # the handle's own state is untouched
    st = sim.state()
    assert st[0] == 1 and np.count_nonzero(st) == 1


@pytest.mark.parametrize("n,num_ops", [(9, 60), (11, 90), (12, 160)])
def test_adjoint_controlled_and_phase_vs_finite_differences(n, num_ops):
    """Controlled RX/RY/RZ/P and fixed gates; from 11 qubits the reverse sweep
    runs in tiled passes (controls inside and outside the tile, more
    parameterised gates than one pass takes)."""
    print(f"adjoint fd n={n}", flush=True)
    rng = np.random.default_rng(4 + n)
    ops = circuits.random_circuit(rng, n, num_ops, max_controls=2)
    kinds = ops["kind"]
    par = np.isin(kinds, [4, 5, 6, 7])
    ops["param"][par] = np.arange(par.sum())
    params = ops["angle"][par].copy()
    ham = circuits.random_hamiltonian(rng, n, 10)
    sim = DenseHip(n)
    _, g = sim.adjoint_gradient(ops, params, ham)
    h = 1e-5
    fd = np.zeros_like(params)
    for j in range(len(params)):
        p = params.copy()
        p[j] += h
        ep = oracle_energy(n, ops, p, ham)
        p[j] -= 2 * h
        em = oracle_energy(n, ops, p, ham)
        fd[j] = (ep - em) / (2 * h)
    assert np.max(np.abs(g - fd)) < 1e-7


def test_adjoint_grouped_and_wide_strings():
    """From 11 qubits lambda = H phi is built in grouped tile sweeps; a string
    with more than 8 X/Y qubits above qubit 2 takes a launch of its own."""
    from kbw_amd.dense import Hamiltonian

    n = 13
    print(f"adjoint strings n={n}", flush=True)
    ops, params = circuits.hea(n, 2, seed=5)
    strings = [
        [("X", q) for q in range(3, 13)] + [("Y", 0)],  # 10 X/Y qubits above qubit 2
        [("Y", 4)],
        [("Y", 2), ("Z", 7)],
        [("X", 1), ("Y", 11), ("Z", 12)],
        [("Z", 0), ("Z", 12)],
        [("Y", q) for q in range(5, 14 - 1)],  # 8 Y qubits above qubit 2: fits one tile
    ]
    ham = Hamiltonian(strings, [0.7, -0.4, 1.3, 0.25, -0.9, 0.55])
    sim = DenseHip(n)
    e, g = sim.adjoint_gradient(ops, params, ham)
    assert e == pytest.approx(oracle_energy(n, ops, params, ham), abs=1e-12)
    ref = oracle_shift(n, ops, params, ham)
    assert np.max(np.abs(g - ref)) < 1e-10


def test_adjoint_24q_subset_vs_plain_parameter_shift():
    """The 24-qubit, 12-layer, 576-parameter ansatz with the XXZ chain: the
    adjoint gradient against the GPU plain path's shift rule on 6 parameters."""
    n = 24
    ops, params = circuits.hea(n, 12, seed=7)
    ham = circuits.xxz(n, seed=8)
    sim = DenseHip(n)
    e, g = sim.adjoint_gradient(ops, params, ham)
    assert len(g) == 576
    idx = [0, 100, 287, 300, 451, 575]
    ps = kag.parameter_shift(sim, ops, params, ham, indices=idx)
    assert np.max(np.abs(g[idx] - ps)) < 1e-9
    e2 = kag.energy(sim, ops, params, ham)
    assert e == pytest.approx(e2, abs=1e-11)


def test_torch_autograd_function():
    import torch

    n = 8
    ops, params = circuits.hea(n, 2, seed=3)
    ham = circuits.tfim(n, seed=3)
    sim = DenseHip(n)
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    e = kag.expval(p, sim, ops, ham)
    (2.0 * e).backward()
    _, g = sim.adjoint_gradient(ops, params, ham)
    assert np.allclose(p.grad.numpy(), 2.0 * g, rtol=0, atol=1e-15)
    assert float(e) == pytest.approx(oracle_energy(n, ops, params, ham), abs=1e-12)


@pytest.mark.parametrize("compiled", [False, True])
@pytest.mark.parametrize("kind", [4, 5, 6, 7])
def test_adjoint_every_position_and_control(kind, compiled):
    """Parameterised P / RX / RY / RZ on each of qubits 0..10 of a 13-qubit
    state (every tile position of the reverse sweep) under every control
    placement, after H on every qubit and with X, Y, H between: gradients
    against central differences of the oracle's energy (step 1e-6)."""
    print(f"adjoint every position kind={kind} compiled={compiled}", flush=True)
    n = 13
    ops_k, params = circuits.every_position(kind, n, params=True)
    pre = circuits.OpList()
    for q in range(n):
        pre.add(3, q)
    mid = circuits.OpList()
    for q in range(0, 11, 2):
        mid.add((0, 1, 3)[q % 3], q, [(q + 5) % n])
    ops = np.concatenate([pre.array(), ops_k[:22], mid.array(), ops_k[22:]])
    rng = np.random.default_rng(60 + kind)
    ham = circuits.random_hamiltonian(rng, n, 12, max_weight=3)
    sim = DenseHip(n)
    sim.set_compiled(compiled)
    e, g = sim.adjoint_gradient(ops, params, ham)
    assert e == pytest.approx(oracle_energy(n, ops, params, ham), abs=1e-12)
    h = 1e-6
    for j in range(len(params)):
        p = params.copy()
        p[j] += h
        fp = oracle_energy(n, ops, p, ham)
        p[j] -= 2 * h
        fm = oracle_energy(n, ops, p, ham)
        assert abs(g[j] - (fp - fm) / (2 * h)) <= 2e-8, (j, g[j], (fp - fm) / (2 * h))
    sim.close()


@pytest.mark.parametrize("n", [13, 20])
def test_adjoint_compiled_equals_generic(n):
    """kbwh_set_compiled: the forward and reverse passes on compiled kernels
    give the generic kernels' energy and gradient to rounding, and a second
    call with other parameters compiles nothing."""
    from kbw_amd import compiled_stats

    print(f"adjoint compiled n={n}", flush=True)
    ops, params = circuits.hea(n, 4, seed=50 + n)
    ham = circuits.xxz(n, seed=51 + n)
    sim = DenseHip(n)
    e0, g0 = sim.adjoint_gradient(ops, params, ham)
    sim.set_compiled(True)
    e1, g1 = sim.adjoint_gradient(ops, params, ham)
    assert abs(e1 - e0) <= 1e-12 and np.max(np.abs(g1 - g0)) <= 1e-12
    before = compiled_stats()
    p2 = params + 0.1
    e2, g2 = sim.adjoint_gradient(ops, p2, ham)
    assert compiled_stats()["compiled"] == before["compiled"]
    sim.set_compiled(False)
    e3, g3 = sim.adjoint_gradient(ops, p2, ham)
    assert abs(e2 - e3) <= 1e-12 and np.max(np.abs(g2 - g3)) <= 1e-12
    sim.close()
