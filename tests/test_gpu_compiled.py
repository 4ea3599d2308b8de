"""Pass-specialised kernels (KBWH_RUN_COMPILED) against the per-gate path,
the oracle and the dependency-ordered generic kernel (KBWH_RUN_REORDER).

A compiled pass runs the same gate arithmetic as the contracted generic pass
kernel, with each gate's kind, tile position and masks fixed in its code, so
it agrees with the per-gate path to rounding (1e-13 on a unit-norm state) and
with KBWH_RUN_REORDER to rounding.  From 12 qubits on the compiled passes use
12-qubit tiles (16 amplitudes per thread) and a searched plan, so they may
take fewer passes than KBWH_RUN_REORDER; on 11 qubits they use the 11-qubit
tile.  Each case puts every gate of one kind into one pass, so a kind's every
tile position and control placement compiles once.
"""
import math

import numpy as np
import pytest

import oracle
from kbw_amd import (DenseHip, RUN_COMPILED, RUN_PER_GATE, RUN_REORDER, circuits, compiled_stats, order_tuning_stats,
                     set_order_tuning)

pytestmark = pytest.mark.gpu


def random_state(rng, n):
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return st / np.linalg.norm(st)


def run_modes(n, st, ops, modes):
    out = []
    for m in modes:
        g = DenseHip(n)
        g.set_state(st)
        g.run(ops, mode=m)
        out.append((g.state(), g.run_stats()[1]))
        g.close()
    return out


@pytest.mark.parametrize("kind", range(8))
@pytest.mark.parametrize("n,tile", [(13, 12), (11, 11)])
def test_compiled_every_position_and_control_placement(kind, n, tile):
    """One pass of gates of `kind` whose targets cover every tile qubit: on 13
    qubits the 12-qubit tile (lane bits 0..5, wave bits 6..7, register bits
    8..11 in some bijection), on 11 qubits the 11-qubit tile (register bits
    8..10); each target uncontrolled, with a control on a lane qubit, on a
    tile qubit 3.. and, on 13 qubits, outside the tile (qubit 12); angles
    covering every sign and magnitude form of the half-angle coefficients."""
    print(f"compiled kind {kind} n={n}", flush=True)
    rng = np.random.default_rng(140 + kind + n)
    st = random_state(rng, n)
    ops = circuits.OpList()
    for q in range(tile):
        lane_c = [0] if q != 0 else [1]
        tile_c = [3] if q != 3 else [4]
        outside = [n - 1] if n > tile else [5 if q != 5 else 6]
        for ctrl, ang in zip(([], lane_c, tile_c, outside), (0.9, 2.5, 4.0, 5.9)):
            ops.add(kind, q, ctrl, angle=ang + 0.01 * q)
    arr = ops.array()
    (ref, _), (comp, passes) = run_modes(n, st, arr, (RUN_PER_GATE, RUN_COMPILED))
    assert passes == 1
    d = float(np.max(np.abs(comp - ref)))
    assert d <= 1e-13, (kind, d)


def test_compiled_diagonal_masks_in_one_pass():
    """Z, P and RZ with targets outside the tile, on register, lane and wave
    bits, under control sets mixing lane, register, wave and outside bits."""
    print("compiled diagonal masks", flush=True)
    n = 13
    rng = np.random.default_rng(77)
    st = random_state(rng, n)
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
    (ref, _), (comp, passes) = run_modes(n, st, arr, (RUN_PER_GATE, RUN_COMPILED))
    assert passes == 1
    assert float(np.max(np.abs(comp - ref))) <= 1e-12


@pytest.mark.parametrize("n,layers,seed", [(16, 6, 3), (20, 4, 4)])
def test_compiled_brickwork_against_oracle_and_reorder(n, layers, seed):
    """The brickwork circuit (every gate kind, CZ/CNOT, multi-controlled X)
    from |0>: amplitudes against the oracle's program-order run to 1e-12, and
    against KBWH_RUN_REORDER's generic passes to 1e-13, in no more passes."""
    print(f"compiled brickwork n={n}", flush=True)
    ops = circuits.OpList()
    circuits.brickwork_layers(ops, np.random.default_rng(seed), n, layers)
    arr = ops.array()
    o = oracle.DenseV2(n)
    o.run(arr)
    want = o.state()
    g = DenseHip(n)
    g.run(arr, mode=RUN_COMPILED)
    comp, passes = g.state(), g.run_stats()[1]
    g.clear()
    g.run(arr, mode=RUN_REORDER)
    reo, passes_r = g.state(), g.run_stats()[1]
    g.close()
    assert passes <= passes_r
    assert float(np.max(np.abs(comp - want))) <= 1e-12
    assert float(np.max(np.abs(comp - reo))) <= 1e-13


def test_compiled_kernels_are_reused_for_new_angles():
    """A second run of the same circuit structure with other parameters
    compiles nothing and finds every pass's kernel; the energies of both
    parameter sets match the oracle."""
    print("compiled reuse", flush=True)
    n = 14
    ops, params = circuits.hea(n, 4, seed=21)
    ham = circuits.xxz(n, seed=22)
    g = DenseHip(n)
    g.run(ops, params, mode=RUN_COMPILED)
    e1 = g.exp_value(ham)
    passes = g.run_stats()[1]
    before = compiled_stats()
    p2 = np.random.default_rng(23).uniform(-math.pi, math.pi, size=len(params))
    g.clear()
    g.run(ops, p2, mode=RUN_COMPILED)
    e2 = g.exp_value(ham)
    after = compiled_stats()
    assert after["compiled"] == before["compiled"]
    assert after["cache_hits"] - before["cache_hits"] == passes
    for pp, e in ((params, e1), (p2, e2)):
        o = oracle.DenseV2(n)
        o.run(ops, pp)
        want = o.exp_value(ham.pauli_strings, ham.coefficients)
        assert abs(e - want) <= 1e-10 * max(1.0, abs(want)), (e, want)
    g.close()


def test_compiled_long_pass_runs_the_generic_kernel():
    """On 11 qubits a pass of more than 256 gates (RY gates on qubits 0..2
    under changing controls all fit one tile) runs the generic kernel: same
    result as KBWH_RUN_REORDER, one more pass counted as generic."""
    print("compiled long pass", flush=True)
    n = 11
    rng = np.random.default_rng(5)
    st = random_state(rng, n)
    ops = circuits.OpList()
    for i in range(300):
        ops.add(6, i % 3, [3 + (i % 8)], angle=0.01 * i)
    arr = ops.array()
    before = compiled_stats()
    (comp, passes), (reo, _) = run_modes(n, st, arr, (RUN_COMPILED, RUN_REORDER))
    after = compiled_stats()
    assert passes == 1
    assert after["generic_passes"] - before["generic_passes"] == 1
    assert np.array_equal(comp, reo)


def test_compiled_wide_long_pass_is_cut():
    """On 12 qubits and more a pass of more than 256 gates is cut into passes
    of at most 256, each compiled: 300 gates in one tile run as 2 compiled
    passes, equal to KBWH_RUN_REORDER's single generic pass to rounding."""
    print("compiled wide long pass", flush=True)
    n = 12
    rng = np.random.default_rng(6)
    st = random_state(rng, n)
    ops = circuits.OpList()
    for i in range(300):
        ops.add(6, i % 3, [3 + (i % 9)], angle=0.01 * i)
    arr = ops.array()
    before = compiled_stats()
    (comp, passes), (reo, passes_r) = run_modes(n, st, arr, (RUN_COMPILED, RUN_REORDER))
    after = compiled_stats()
    assert (passes, passes_r) == (2, 1)
    assert after["generic_passes"] == before["generic_passes"]
    assert float(np.max(np.abs(comp - reo))) <= 1e-13


@pytest.mark.parametrize("trials", [0, 64])
def test_compiled_plan_search_and_cache(trials):
    """The 30-qubit-shaped plan search on a 16-qubit brickwork: a searched
    plan never has more passes than the greedy one, a second run of the same
    structure with other angles reuses the plan and the kernels, and both
    runs match the oracle to 1e-12."""
    print(f"compiled plan search trials={trials}", flush=True)
    n = 16
    ops = circuits.OpList()
    circuits.brickwork_layers(ops, np.random.default_rng(9), n, 8)
    arr = ops.array()
    g = DenseHip(n)
    g.set_plan_search(0)
    g.run(arr, mode=RUN_COMPILED)
    greedy = g.run_stats()[1]
    g.clear()
    g.set_plan_search(trials)
    g.run(arr, mode=RUN_COMPILED)
    searched = g.run_stats()[1]
    o = oracle.DenseV2(n)
    o.run(arr)
    assert float(np.max(np.abs(g.state() - o.state()))) <= 1e-12
    assert searched <= greedy
    arr2 = arr.copy()
    rot = np.isin(arr2["kind"], [5, 6, 7])
    arr2["angle"][rot] += 0.37
    before = compiled_stats()
    g.clear()
    g.run(arr2, mode=RUN_COMPILED)
    after = compiled_stats()
    assert g.run_stats()[1] == searched
    assert after["compiled"] == before["compiled"]
    o2 = oracle.DenseV2(n)
    o2.run(arr2)
    assert float(np.max(np.abs(g.state() - o2.state()))) <= 1e-12
    g.close()


def test_compiled_tile_order_tuning_keeps_the_bits():
    """The tile-order tuner (24 qubits: 4096 tiles per pass): the first runs
    of each pass layout launch it with each candidate order, then the fastest
    is kept.  The order only permutes which workgroup runs which tile, so
    every run from the same state gives the same bits; after the candidate
    runs every layout of the circuit is tuned, and the state matches the
    oracle's program-order run to 1e-12.  With tuning off the bits are the
    same again."""
    print("compiled tile-order tuning 24q", flush=True)
    n = 24
    ops = circuits.OpList()
    circuits.brickwork_layers(ops, np.random.default_rng(21), n, 6)
    arr = ops.array()
    g = DenseHip(n)
    states = []
    for _ in range(6):
        g.clear()
        g.run(arr, mode=RUN_COMPILED)
        states.append(g.state())
    passes = g.run_stats()[1]
    st = order_tuning_stats()
    assert st["pending"] == 0 and st["tuned"] >= passes and st["trials"] >= 4 * passes, st
    for s in states[1:]:
        assert np.array_equal(s, states[0])
    set_order_tuning(False)
    try:
        g.clear()
        g.run(arr, mode=RUN_COMPILED)
        assert np.array_equal(g.state(), states[0])
    finally:
        set_order_tuning(True)
    o = oracle.DenseV2(n)
    o.run(arr)
    assert float(np.max(np.abs(states[0] - o.state()))) <= 1e-12
    g.close()


@pytest.mark.parametrize("n,L,kind", [(14, 12, "random"), (16, 12, "qft"), (23, 18, "qft"), (20, 16, "random")])
def test_compiled_block_states(n, L, kind):
    """A Block state in KBWH_RUN_COMPILED: the reference's map and swap count
    (decided in program order), passes with folded global/local swaps run on
    kernels compiled with a second store spec (the tile written through the
    exchanged positions after every load has returned), the physical
    amplitudes against the oracle's Block to 1e-12 and against
    KBWH_RUN_REORDER's generic passes to rounding; no kernel fails to
    compile."""
    print(f"compiled block n={n} L={L} {kind}", flush=True)
    if kind == "qft":
        ops = circuits.qft_plus_brickwork(n, 4, seed=2)
    else:
        ops = circuits.random_circuit(np.random.default_rng(31 + n), n, 300, max_controls=2)
    before = compiled_stats()
    g = DenseHip(n, block=True, block_size=L)
    g.run(ops, mode=RUN_COMPILED)
    after = compiled_stats()
    assert after["failures"] == before["failures"]
    assert after["compiled"] + after["cache_hits"] > before["compiled"] + before["cache_hits"]
    r = DenseHip(n, block=True, block_size=L)
    r.run(ops, mode=RUN_REORDER)
    o = oracle.Block(n, L)
    for op in ops:
        m = int(op["ctrl_mask"])
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), [q for q in range(n) if m >> q & 1])
    assert np.array_equal(g.qubit_map(), o.qubit_map())
    assert g.swap_count() == o.swap_count() > 0
    st = g.state()
    assert np.max(np.abs(st - o.physical_state())) < 1e-12
    assert np.max(np.abs(st - r.state())) < 1e-13
    g.close()
    r.close()
