"""The tiled multi-gate pass (KBWH_RUN_TILED) against the oracle: same bits.

A tiled run applies each gate with the same per-pair arithmetic as the plain
path, in program order, so every amplitude must equal the oracle's exactly.
"""
import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, RUN_PER_GATE, RUN_TILED, circuits

pytestmark = pytest.mark.gpu


def assert_same(a, b):
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{bad.size} amplitudes differ, first {bad[:4]}: {a[bad[:4]]} vs {b[bad[:4]]}"


@pytest.mark.parametrize("n,ngates,seed", [(12, 300, 0), (13, 300, 1), (16, 400, 2), (21, 200, 3)])
def test_tiled_random_circuits_bit_exact(n, ngates, seed):
    print(f"tiled random n={n}", flush=True)
    rng = np.random.default_rng(seed + 1000)
    ops = circuits.random_circuit(rng, n, ngates, max_controls=4)
    g = DenseHip(n)
    g.run(ops, mode=RUN_TILED)
    gates, passes = g.run_stats()
    assert gates == ngates and 0 < passes < ngates
    o = oracle.DenseV2(n)
    o.run(ops)
    assert_same(g.state(), o.state())


@pytest.mark.parametrize("n,layers", [(12, 6), (17, 6), (22, 3)])
def test_tiled_brickwork_bit_exact(n, layers):
    print(f"tiled brickwork n={n}", flush=True)
    ops, _ = circuits.brickwork(n, layers, seed=5)
    g = DenseHip(n)
    g.run(ops, mode=RUN_TILED)
    o = oracle.DenseV2(n)
    o.run(ops)
    assert_same(g.state(), o.state())


@pytest.mark.parametrize("mode", ["tiled", "reorder"])
@pytest.mark.parametrize("kind", range(8))
def test_tiled_every_kind_at_every_tile_position(kind, mode):
    """One gate of `kind` at each tile position: lane bits 0..5, wave bits 6..7,
    register bits 8..10; uncontrolled, with a control on a lane bit, with one
    on a register bit and with one outside the tile, on a random state.  In a
    pass whose first-used targets are qubits 3..10 (each used once), the
    planner gives them positions 8, 9, 10, 3, 4, 5, 6, 7 in that order; qubits
    0..2 are always positions 0..2.  The other targets get X gates.
    Dependency-ordered runs (the contracted kernel's in-place forms) give the
    targets some bijection onto the same positions, so every position is
    covered too; they agree with the per-gate path to 1e-13."""
    from kbw_amd import RUN_REORDER

    print(f"tiled kind {kind} at every position, {mode}", flush=True)
    n = 13
    rng = np.random.default_rng(40 + kind)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    st /= np.linalg.norm(st)
    cases = []
    for test_q in range(11):
        lane_c = [0] if test_q != 0 else [1]
        reg_c = [3] if test_q != 3 else [4]
        # angles whose half-angle (or phase) has cos > 0 and |cos| > |sin|,
        # cos > 0 and |sin| > |cos|, cos < 0 and |sin| > |cos|, cos < 0 and
        # |cos| > |sin|: every shear and lane form of the in-place kernel
        for ctrl, ang in zip(([], lane_c, reg_c, [12]), (0.9, 2.5, 4.0, 5.9)):
            ops = circuits.OpList()
            if test_q < 3:
                ops.add(kind, test_q, ctrl, angle=ang)
            for t in range(3, 11):
                if t == test_q:
                    ops.add(kind, t, ctrl, angle=ang)
                else:
                    ops.add(0, t)
            cases.append((test_q, ctrl, ops.array()))
    for test_q, ctrl, ops in cases:
        a = DenseHip(n)
        a.set_state(st)
        a.run(ops, mode=RUN_TILED if mode == "tiled" else RUN_REORDER)
        assert a.run_stats()[1] == 1
        b = DenseHip(n)
        b.set_state(st)
        b.run(ops, mode=RUN_PER_GATE)
        if mode == "tiled":
            bad = np.nonzero(a.state() != b.state())[0]
            assert bad.size == 0, (kind, test_q, ctrl, bad.size)
        else:
            d = float(np.max(np.abs(a.state() - b.state())))
            assert d <= 1e-13, (kind, test_q, ctrl, d)
        a.close()
        b.close()


def test_tiled_diagonal_masks_in_one_pass():
    """Z, P and RZ with their target outside the tile, on register bits and on
    lane and wave bits, under control sets that mix lane, register, wave and
    outside bits, all in one pass (the gate records carry the thread and
    register-slot masks of target and controls; a Z or P target outside the
    tile joins the outside controls).  X gates on qubits 3..10 fix the tile
    positions as in the every-position test: 3, 4, 5 -> register bits 8, 9,
    10; 6, 7, 8 -> lane bits 3, 4, 5; 9, 10 -> wave bits 6, 7; qubits 11 and
    12 stay outside.  Tiled runs are bit-exact with the per-gate path;
    dependency-ordered runs agree to 1e-12."""
    from kbw_amd import RUN_REORDER

    print("tiled diagonal masks", flush=True)
    n = 13
    rng = np.random.default_rng(77)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    st /= np.linalg.norm(st)
    cases = [(11, []), (11, [0]), (11, [3]), (11, [7]), (11, [9]), (11, [12]), (11, [0, 3, 12]),
             (3, []), (3, [11]), (3, [0, 12]), (4, [1, 5]), (5, [8, 10]),
             (0, []), (0, [3, 4]), (1, [11, 12]), (7, [2, 9]), (9, []), (10, [3, 6, 11]), (12, [4, 10])]
    ops = circuits.OpList()
    for t in range(3, 11):
        ops.add(0, t)
    for kind in (2, 4, 7):
        for t, ctrl in cases:
            ops.add(kind, t, ctrl, angle=0.3 + 0.1 * t)
    arr = ops.array()
    ref = DenseHip(n)
    ref.set_state(st)
    ref.run(arr, mode=RUN_PER_GATE)
    want = ref.state()
    ref.close()
    a = DenseHip(n)
    a.set_state(st)
    a.run(arr, mode=RUN_TILED)
    assert a.run_stats()[1] == 1
    bad = np.nonzero(a.state() != want)[0]
    assert bad.size == 0, bad[:8]
    a.close()
    b = DenseHip(n)
    b.set_state(st)
    b.run(arr, mode=RUN_REORDER)
    assert np.max(np.abs(b.state() - want)) <= 1e-12
    b.close()


def test_tiled_small_state_uses_per_gate():
    n = 8
    ops = circuits.random_circuit(np.random.default_rng(3), n, 100)
    g = DenseHip(n)
    g.run(ops, mode=RUN_TILED)
    o = oracle.DenseV2(n)
    o.run(ops)
    assert_same(g.state(), o.state())


def test_tiled_block_mode_bit_exact():
    n, L = 14, 12
    rng = np.random.default_rng(9)
    ops = circuits.random_circuit(rng, n, 200, max_controls=2)
    g = DenseHip(n, block=True, block_size=L)
    g.run(ops, mode=RUN_TILED)
    o = oracle.Block(n, L)
    for op in ops:
        m = int(op["ctrl_mask"])
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), [q for q in range(n) if m >> q & 1])
    assert np.array_equal(g.qubit_map(), o.qubit_map())
    assert g.swap_count() == o.swap_count() > 0
    assert_same(g.state(), o.physical_state())


@pytest.mark.parametrize("targets", [1, 2, 3, 8])
def test_tiled_block_swaps_in_passes(targets):
    """Block swaps folded into tiled passes (the tile stored through the
    exchanged qubits): same map, swap count and physical amplitudes as the
    oracle's Block.  With one target position per pass a swap cannot share a
    tile and takes a swap pass of its own."""
    n, L = 15, 11
    print(f"block swaps targets={targets}", flush=True)
    rng = np.random.default_rng(17)
    ops = circuits.random_circuit(rng, n, 300, max_controls=3)
    g = DenseHip(n, block=True, block_size=L)
    g.set_run_config(tile_targets=targets)
    g.run(ops, mode=RUN_TILED)
    o = oracle.Block(n, L)
    for op in ops:
        m = int(op["ctrl_mask"])
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), [q for q in range(n) if m >> q & 1])
    assert np.array_equal(g.qubit_map(), o.qubit_map())
    assert g.swap_count() == o.swap_count() > 20
    assert_same(g.state(), o.physical_state())


def test_tiled_params_vector():
    n = 14
    ops, params = circuits.hea(n, 3, seed=4)
    g = DenseHip(n)
    g.run(ops, params=params * 0.5, mode=RUN_TILED)
    o = oracle.DenseV2(n)
    o.run(ops, params * 0.5)
    assert_same(g.state(), o.state())


def test_tiled_30q_equals_per_gate():
    """At 30 qubits: the tiled run of the brickwork circuit's first 5 layers
    equals the per-gate run bit for bit (the per-gate path is itself checked
    against the oracle at 30 qubits)."""
    n = 30
    ops, _ = circuits.brickwork(n, 20, seed=1)
    pre = ops[:230]
    a = DenseHip(n)
    a.run(pre, mode=RUN_TILED)
    gates, passes = a.run_stats()
    print(f"30q tiled: {gates} gates in {passes} passes", flush=True)
    b = DenseHip(n)
    b.run(pre, mode=RUN_PER_GATE)
    chunk = 1 << 26
    for first in range(0, 1 << n, chunk):
        assert_same(a.state(first, chunk), b.state(first, chunk))
