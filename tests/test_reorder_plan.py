"""The dependency-ordered pass plan of KBWH_RUN_REORDER, on the host
(kbwh_reorder_plan needs no device).

The plan may trade the places of gates that share no qubit and nothing
else: on every qubit the gates touching it keep their program order, each
pass needs at most the tile's free target positions, and running the ops in
plan order gives the program-order state up to rounding (checked with the
oracle on small states)."""
import numpy as np
import pytest

import oracle
from kbw_amd import circuits
from kbw_amd.dense import reorder_plan

DIAG = {2, 4, 7}


def qubits(op):
    m = int(op["ctrl_mask"]) | (1 << int(op["target"]))
    return [q for q in range(64) if m >> q & 1]


def inorder_passes(ops, free=8):
    """Passes of the program-order planner (KBWH_RUN_TILED)."""
    passes, targets = 1, set()
    for op in ops:
        k, t = int(op["kind"]), int(op["target"])
        if k in DIAG or t < 3 or t in targets:
            continue
        if len(targets) < free:
            targets.add(t)
            continue
        passes += 1
        targets = {t}
    return passes


def check_plan(ops, n, cap=64):
    order, passes = reorder_plan(ops, n, cap)
    m = len(ops)
    assert sorted(order.tolist()) == list(range(m))
    # per qubit, program order is kept
    pos = np.empty(m, dtype=np.int64)
    pos[order] = np.arange(m)
    last = {}
    for i in range(m):
        for q in qubits(ops[i]):
            if q in last:
                assert pos[last[q]] < pos[i], (q, last[q], i)
            last[q] = i
    # every pass fits the tile: qubits 0..2 plus at most min(cap, 8) targets
    assert passes[0][0] == 0 and passes[-1][1] == m
    for (a, b), (c, _) in zip(passes, passes[1:] + [(m, m)]):
        assert a < b == c
        t = {int(ops[i]["target"]) for i in order[a:b] if int(ops[i]["kind"]) not in DIAG}
        assert len(t - {0, 1, 2}) <= min(cap, 8)
    return order, passes


def test_brickwork_30q_plan():
    ops, _ = circuits.brickwork(30, 20, seed=1)
    _, passes = check_plan(ops, 30)
    assert inorder_passes(ops) == 56
    assert len(passes) == 21


def test_hea_24q_plan():
    ops, _ = circuits.hea(24, 12, seed=7)
    _, passes = check_plan(ops, 24)
    assert inorder_passes(ops) == 32
    assert len(passes) == 17


def test_qft_33q_plan():
    ops = circuits.qft_plus_brickwork(33, 4, seed=2)
    _, passes = check_plan(ops, 33)
    assert len(passes) < inorder_passes(ops)


@pytest.mark.parametrize("cap", [1, 2, 4, 8])
def test_random_controlled_plan_caps(cap):
    rng = np.random.default_rng(cap)
    ops = circuits.random_circuit(rng, 16, 400, max_controls=4)
    _, passes = check_plan(ops, 16, cap)
    assert len(passes) <= inorder_passes(ops, min(cap, 8))


@pytest.mark.parametrize("n,seed", [(6, 0), (9, 1), (12, 2)])
def test_plan_order_gives_the_same_state(n, seed):
    """Running the ops in plan order equals program order up to rounding."""
    rng = np.random.default_rng(seed + 50)
    ops = circuits.random_circuit(rng, n, 250, max_controls=3)
    order, _ = check_plan(ops, n, 4)
    assert not np.array_equal(order, np.arange(len(ops)))
    a = oracle.DenseV2(n)
    a.run(ops)
    b = oracle.DenseV2(n)
    b.run(ops[order])
    assert np.max(np.abs(a.state() - b.state())) < 1e-12


def test_plan_rejects_bad_input():
    from kbw_amd import KbwhError
    ops = circuits.random_circuit(np.random.default_rng(0), 8, 10)
    with pytest.raises(KbwhError):
        reorder_plan(ops, 5)
    with pytest.raises(KbwhError):
        reorder_plan(ops, 8, 0)


@pytest.mark.parametrize("seed", range(12))
def test_random_plans_many_shapes(seed):
    """Random circuits of 3..40 qubits with up to 5 controls per gate and
    every target cap: the plan invariants hold (also run under the
    AddressSanitizer / UBSan build of the planners)."""
    rng = np.random.default_rng(900 + seed)
    n = int(rng.integers(3, 41))
    ops = circuits.random_circuit(rng, n, int(rng.integers(1, 400)), max_controls=min(5, n - 1))
    for cap in (1, 3, 8, 64):
        check_plan(ops, n, cap)
