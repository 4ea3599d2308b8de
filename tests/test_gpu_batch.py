"""Batched persistent launch (kbwh_batch_*) against the oracle, the per-gate
path, and itself under hipGraph replay."""
import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, circuits
from kbw_amd.batch import Batch

pytestmark = pytest.mark.gpu


def oracle_expval(n, ops, params, ham):
    o = oracle.DenseV2(n)
    o.run(ops, params)
    return o.exp_value(ham.pauli_strings, ham.coefficients), o.state()


def mixed_batch(seed):
    rng = np.random.default_rng(seed)
    circs, params = [], []
    for n in (3, 10, 11, 12, 14, 17, 9, 13):
        if n % 2:
            ops, p = circuits.hea(n, 3, seed=int(rng.integers(1 << 30)))
        else:
            ops = circuits.random_circuit(rng, n, 80, max_controls=2)
            rot = np.isin(ops["kind"], [4, 5, 6, 7])
            ops["param"][rot] = np.arange(rot.sum())
            p = ops["angle"][rot].copy()
        ham = circuits.random_hamiltonian(rng, n, 10, max_weight=3)
        circs.append((n, ops, ham, len(p)))
        params.append(p)
    return circs, np.concatenate(params) if params else np.zeros(0)


def test_batch_matches_oracle():
    circs, params = mixed_batch(1)
    b = Batch(circs)
    out = b.run(params)
    off = b.param_offsets
    for i, (n, ops, ham, npc) in enumerate(circs):
        e, st = oracle_expval(n, ops, params[off[i]:off[i + 1]], ham)
        assert abs(out[i] - e) <= 1e-12 * max(1.0, abs(e)), (i, n, out[i], e)
        got = b.state(i, 1 << n)
        assert np.max(np.abs(got - st)) < 1e-13


def test_batch_string_placement_matches_oracle():
    # strings evaluated in the last gate pass, in read-only passes, with Y
    # factors, on gate-free circuits and on circuits narrower than a tile
    rng = np.random.default_rng(21)
    circs, params = [], []
    empty = circuits.OpList().array()
    ham_y = circuits.Hamiltonian([[("Y", 0)], [("X", 1), ("Y", 2)], [("Z", 0), ("Z", 2)]], [0.7, -1.1, 0.4])
    circs.append((3, empty, ham_y, 0))
    params.append(np.zeros(0))
    for n in (20, 12, 19):
        ops, p = circuits.hea(n, 2, seed=int(rng.integers(1 << 30)))
        strings = [[("X", q)] for q in range(n)] + [[("Y", q), ("Y", (q + 7) % n)] for q in range(0, n, 3)]
        strings += [[("Z", q), ("Z", q + 1)] for q in range(n - 1)]
        ham = circuits.Hamiltonian(strings, rng.normal(0, 1, size=len(strings)))
        circs.append((n, ops, ham, len(p)))
        params.append(p)
    circs.append((14, empty, circuits.Hamiltonian([[("X", 13)], [("Z", 0)]], [1.0, 2.0]), 0))
    params.append(np.zeros(0))
    params = np.concatenate(params)
    b = Batch(circs)
    out = b.run(params)
    off = b.param_offsets
    for i, (n, ops, ham, npc) in enumerate(circs):
        e, st = oracle_expval(n, ops, params[off[i]:off[i + 1]], ham)
        assert abs(out[i] - e) <= 1e-12 * max(1.0, abs(e)), (i, n, out[i], e)
        assert np.max(np.abs(b.state(i, 1 << n) - st)) < 1e-13


def test_batch_graph_replays_match_eager():
    circs, params = mixed_batch(2)
    b = Batch(circs)
    rng = np.random.default_rng(3)
    sets = [params, params + rng.normal(0, 0.3, size=params.shape), params * 0.5]
    eager = [b.run(p) for p in sets]
    b.use_graph(True)
    for rep in range(3):
        for k, p in enumerate(sets):
            got = b.run(p)
            assert np.array_equal(got, eager[k]), (rep, k)
        b.use_graph(False)
        assert np.array_equal(b.run(sets[1]), eager[1])  # eager work between replays
        b.use_graph(True)


def test_batch_c3_workload_spot_check():
    circs, params = circuits.hea_batch(4096, seed=12)
    b = Batch(circs)
    out = b.run(params)
    assert np.all(np.isfinite(out))
    gates, passes, amps = b.stats()
    print(f"c3: {len(circs)} circuits, {gates} gates, {passes} passes", flush=True)
    off = b.param_offsets
    rng = np.random.default_rng(4)
    for i in rng.choice(len(circs), size=6, replace=False):
        n, ops, ham, _ = circs[i]
        s = DenseHip(n)
        s.run(ops, params=params[off[i]:off[i + 1]])
        e = s.exp_value(ham)
        s.close()
        assert abs(out[i] - e) <= 1e-12 * max(1.0, abs(e)), (i, n, out[i], e)


def group_batch(seed, specs):
    """HEA circuits: specs = [(n, layers, count)], each with its own angles and
    transverse-field weights (one structure per (n, layers)), interleaved
    with a few random-gate circuits that run in the persistent kernel."""
    rng = np.random.default_rng(seed)
    circs, params = [], []
    for n, layers, count in specs:
        for _ in range(count):
            ops, p = circuits.hea(n, layers, seed=int(rng.integers(1 << 30)))
            ham = circuits.tfim(n, seed=int(rng.integers(1 << 30)))
            circs.append((n, ops, ham, len(p)))
            params.append(p)
    for n in (12, 15):
        ops = circuits.random_circuit(rng, n, 60, max_controls=2)
        rot = np.isin(ops["kind"], [4, 5, 6, 7])
        ops["param"][rot] = np.arange(rot.sum())
        circs.append((n, ops, circuits.random_hamiltonian(rng, n, 8, max_weight=3), int(rot.sum())))
        params.append(ops["angle"][rot].copy())
    order = rng.permutation(len(circs))
    circs = [circs[i] for i in order]
    params = [params[i] for i in order]
    return circs, np.concatenate(params)


def check_against_oracle(b, circs, params, out, picks, tol=1e-11):
    off = b.param_offsets
    for i in picks:
        n, ops, ham, _ = circs[i]
        e, st = oracle_expval(n, ops, params[off[i]:off[i + 1]], ham)
        assert abs(out[i] - e) <= tol * max(1.0, abs(e)), (i, n, out[i], e)
        assert np.max(np.abs(b.state(i, 1 << n) - st)) < 1e-12, (i, n)


def test_batch_structure_groups_match_oracle():
    """Two groups (40 x 14 qubits = 320 tiles, 300 x 10 qubits on 11-qubit
    sets = 300 tiles), a group too small to split off (5 x 13 qubits) and two
    random circuits: energies and states against the oracle."""
    print("batch groups", flush=True)
    circs, params = group_batch(31, [(14, 3, 40), (10, 2, 300), (13, 2, 5)])
    b = Batch(circs)
    groups = b.groups()
    assert sorted((g["num_qubits"], g["num_sets"]) for g in groups) == [(11, 300), (14, 40)]
    out = b.run(params)
    rng = np.random.default_rng(7)
    nq = np.array([c[0] for c in circs])
    picks = list(rng.choice(np.nonzero(nq == 14)[0], 4, replace=False))
    picks += list(rng.choice(np.nonzero(nq == 10)[0], 4, replace=False))
    picks += list(np.nonzero((nq == 13) | (nq == 12) | (nq == 15))[0])
    check_against_oracle(b, circs, params, out, picks)
    b.close()


def test_batch_groups_configure_autotune_and_graph():
    """Target caps and compiled kernels per group agree with the generic run
    to rounding; the tuner keeps one configuration per group and reports its
    time; graph replays of the tuned batch equal its eager runs."""
    print("batch groups tuning", flush=True)
    circs, params = group_batch(32, [(16, 3, 12), (12, 3, 200)])
    b = Batch(circs)
    assert len(b.groups()) == 2
    base = b.run(params)
    b.configure(6, compiled=True)
    assert all(g["compiled"] == g["passes"] for g in b.groups())
    assert np.max(np.abs(b.run(params) - base)) <= 1e-12
    ms = b.autotune(params)
    assert ms.shape == (2, 4) and np.all(ms[:, :3] > 0)
    info = b.groups()
    assert all(g["tuned_ms"] > 0 and g["tile_targets"] in (8, 6, 4) for g in info)
    tuned = b.run(params)
    assert np.max(np.abs(tuned - base)) <= 1e-12
    rng = np.random.default_rng(8)
    check_against_oracle(b, circs, params, tuned, rng.choice(len(circs), 6, replace=False))
    p2 = params + rng.normal(0, 0.2, size=params.shape)
    eager = b.run(p2)
    b.use_graph(True)
    for rep in range(3):
        assert np.array_equal(b.run(p2), eager), rep
        b.use_graph(False)
        assert np.array_equal(b.run(params), tuned)
        b.use_graph(True)
    b.close()


@pytest.mark.parametrize("kind", range(8))
def test_batch_every_kind_every_position(kind):
    """One circuit of `kind` on each of qubits 0..10 of 13 under every
    control placement, after H on every qubit: alone (the persistent kernel)
    and as a group of 64 sets with their own parameters (one launch per pass
    over all sets), energies and states against the oracle."""
    print(f"batch every position kind={kind}", flush=True)
    n = 13
    ops_k, p0 = circuits.every_position(kind, n, params=True)
    pre = circuits.OpList()
    for q in range(n):
        pre.add(3, q)
    ops = np.concatenate([pre.array(), ops_k])
    rng = np.random.default_rng(80 + kind)
    strings = circuits.random_hamiltonian(rng, n, 10, max_weight=3).pauli_strings
    circs, params = [], []
    for c in range(65):
        ham = circuits.Hamiltonian(strings, rng.normal(0, 1, size=len(strings)))
        p = p0 + (rng.normal(0, 0.3, size=len(p0)) if c else 0.0)
        circs.append((n, ops, ham, len(p0)))
        params.append(p)
    # circuit 64 differs in structure (an extra gate): it runs in the persistent kernel
    extra = np.concatenate([ops, circuits.OpList().array()])
    lone = circuits.OpList()
    lone.add(0, 12)
    circs[64] = (n, np.concatenate([extra, lone.array()]), circs[64][2], len(p0))
    params = np.concatenate(params) if len(p0) else np.zeros(0)
    b = Batch(circs)
    assert [g["num_sets"] for g in b.groups()] == [64]
    out = b.run(params)
    check_against_oracle(b, circs, params, out, [0, 17, 63, 64])
    b.close()
