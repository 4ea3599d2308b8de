"""Dependency-ordered tiled passes (KBWH_RUN_REORDER) on the GPU.

Gates that share no qubit may trade places, so a reordered run applies the
circuit's unitary with another rounding order: every amplitude agrees with
the oracle's program-order run to 1e-12 (absolute), in no more passes than
KBWH_RUN_TILED.  A Block state keeps the reference's map and swap count; its
swaps are ordered with the gates."""
import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, RUN_PER_GATE, RUN_REORDER, RUN_TILED, circuits

pytestmark = pytest.mark.gpu

TOL = 1e-12


def passes_of(n, ops, mode, params=None):
    g = DenseHip(n)
    g.run(ops, params=params, mode=mode)
    _, p = g.run_stats()
    return g, p


@pytest.mark.parametrize("n,ngates,seed", [(11, 300, 0), (13, 300, 1), (16, 400, 2), (21, 300, 3)])
def test_reorder_random_circuits(n, ngates, seed):
    print(f"reorder random n={n}", flush=True)
    rng = np.random.default_rng(seed + 2000)
    ops = circuits.random_circuit(rng, n, ngates, max_controls=4)
    g, p = passes_of(n, ops, RUN_REORDER)
    _, pt = passes_of(n, ops, RUN_TILED)
    assert p <= pt
    o = oracle.DenseV2(n)
    o.run(ops)
    assert np.max(np.abs(g.state() - o.state())) < TOL


@pytest.mark.parametrize("n,seed", [(12, 0), (15, 1), (19, 2)])
def test_reorder_fused_diagonals(n, seed):
    """Chains of uncontrolled Z, P and RZ on each qubit, interleaved with CZ,
    controlled phases and CNOTs whose controls sit on those qubits (diagonal
    gates fuse across them), with H and RY (X-type on their target: fusion
    stops there) and pairs that cancel exactly (Z Z, P(a) P(-a) on the |1>
    half): the dependency-ordered run agrees with the oracle to 1e-12."""
    print(f"reorder fused diagonals n={n}", flush=True)
    rng = np.random.default_rng(seed + 3000)
    ops = circuits.OpList()
    for layer in range(12):
        for q in range(n):
            r = int(rng.integers(0, 6))
            a = float(rng.uniform(-np.pi, np.pi))
            if r == 0:
                ops.add(2, q)
                ops.add(2, q)  # cancels
            elif r == 1:
                ops.add(4, q, angle=a)
                ops.add(4, q, angle=-a)
            elif r == 2:
                ops.add(7, q, angle=a)
                ops.add(4, q, angle=0.5 * a)
            elif r == 3:
                ops.add(4, q, angle=a)
            elif r == 4:
                ops.add(7, q, angle=a)
            else:
                ops.add(2, q)
        for q in range(layer % 2, n - 1, 2):
            c = int(rng.integers(0, 3))
            if c == 0:
                ops.add(2, q + 1, [q])  # CZ
            elif c == 1:
                ops.add(4, q + 1, [q], angle=float(rng.uniform(-np.pi, np.pi)))  # controlled phase
            else:
                ops.add(0, q + 1, [q])  # CNOT: fusion continues on q, stops on q + 1
        if layer % 3 == 2:
            for q in range(0, n, 3):
                ops.add(3 if q % 2 else 6, q, angle=0.7)
    arr = ops.array()
    g, _ = passes_of(n, arr, RUN_REORDER)
    o = oracle.DenseV2(n)
    o.run(arr)
    assert np.max(np.abs(g.state() - o.state())) < TOL


@pytest.mark.parametrize("kind,n", [("brickwork", 20), ("hea", 18)])
def test_reorder_layered_circuits(kind, n):
    print(f"reorder {kind} n={n}", flush=True)
    if kind == "brickwork":
        ops, _ = circuits.brickwork(n, 12, seed=3)
        params = None
    else:
        ops, params = circuits.hea(n, 6, seed=4)
    g, p = passes_of(n, ops, RUN_REORDER, params)
    _, pt = passes_of(n, ops, RUN_TILED, params)
    assert p < pt
    o = oracle.DenseV2(n)
    o.run(ops, params)
    assert np.max(np.abs(g.state() - o.state())) < TOL


@pytest.mark.parametrize("n,L,kind,targets", [(14, 12, "random", 64), (15, 11, "random", 3), (16, 12, "qft", 64),
                                              (23, 18, "qft", 64), (23, 18, "random", 2)])
def test_reorder_block_states(n, L, kind, targets):
    """A Block state keeps the reference's map and swap count (decided in
    program order); its swaps are scheduled with the gates, so the physical
    amplitudes agree with the oracle's Block to rounding.  Cap 2 still fits a
    swap of two qubits above qubit 2."""
    print(f"reorder block n={n} L={L} {kind} cap {targets}", flush=True)
    if kind == "qft":
        ops = circuits.qft_plus_brickwork(n, 4, seed=2)
    else:
        ops = circuits.random_circuit(np.random.default_rng(9 + n), n, 300, max_controls=2)
    g = DenseHip(n, block=True, block_size=L)
    g.set_run_config(tile_targets=targets)
    g.run(ops, mode=RUN_REORDER)
    _, p = g.run_stats()
    t = DenseHip(n, block=True, block_size=L)
    t.set_run_config(tile_targets=targets)
    t.run(ops, mode=RUN_TILED)
    _, pt = t.run_stats()
    assert p <= pt
    o = oracle.Block(n, L)
    for op in ops:
        m = int(op["ctrl_mask"])
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), [q for q in range(n) if m >> q & 1])
    assert np.array_equal(g.qubit_map(), o.qubit_map())
    assert g.swap_count() == o.swap_count() > 0
    assert np.max(np.abs(g.state() - o.physical_state())) < TOL


def test_reorder_30q_against_per_gate():
    """30 qubits, the brickwork circuit's first 5 layers: the reordered run
    against the per-gate run (checked against the oracle at 30 qubits)."""
    n = 30
    ops, _ = circuits.brickwork(n, 20, seed=1)
    pre = ops[:230]
    a, p = passes_of(n, pre, RUN_REORDER)
    b = DenseHip(n)
    b.run(pre, mode=RUN_PER_GATE)
    print(f"30q reorder: {len(pre)} gates in {p} passes", flush=True)
    chunk = 1 << 26
    worst = 0.0
    for first in range(0, 1 << n, chunk):
        worst = max(worst, float(np.max(np.abs(a.state(first, chunk) - b.state(first, chunk)))))
    assert worst < TOL
