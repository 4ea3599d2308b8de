"""Fused k-qubit blocks on FP64 MFMA (KBWH_RUN_FUSED) and the autotuner.

Fused blocks accumulate products in another order than gate by gate, so the
bar is the amplitude tolerance, 1e-12 max-abs against the oracle, not bit
equality; the tiled and per-gate modes stay bit-exact."""
import numpy as np
import pytest

import oracle
from kbw_amd import RUN_AUTO, RUN_FUSED, RUN_PER_GATE, RUN_REORDER, RUN_TILED, DenseHip, circuits

pytestmark = pytest.mark.gpu
TOL = 1e-12


def fused_vs_oracle(n, ops, k, params=None):
    g = DenseHip(n)
    g.set_run_config(fusion_width=k)
    g.run(ops, params=params, mode=RUN_FUSED)
    o = oracle.DenseV2(n)
    o.run(ops, params)
    err = float(np.max(np.abs(g.state() - o.state())))
    gates, passes = g.run_stats()
    g.close()
    return err, passes


@pytest.mark.parametrize("k", [4, 5])
@pytest.mark.parametrize("n,ngates,seed", [(11, 200, 0), (13, 300, 1), (16, 300, 2), (20, 150, 3)])
def test_fused_random_circuits(k, n, ngates, seed):
    print(f"fused k={k} random n={n}", flush=True)
    rng = np.random.default_rng(seed + 500)
    ops = circuits.random_circuit(rng, n, ngates, max_controls=3)  # some gates wider than k
    err, passes = fused_vs_oracle(n, ops, k)
    assert err < TOL, err
    assert passes < ngates


@pytest.mark.parametrize("k", [4, 5])
def test_fused_brickwork_and_ansatz(k):
    print(f"fused k={k} brickwork + ansatz", flush=True)
    ops, _ = circuits.brickwork(18, 6, seed=7)
    err, _ = fused_vs_oracle(18, ops, k)
    assert err < TOL, err
    ops, params = circuits.hea(14, 4, seed=8)
    err, _ = fused_vs_oracle(14, ops, k, params * 0.7)
    assert err < TOL, err


def test_fused_small_state_runs_per_gate():
    ops = circuits.random_circuit(np.random.default_rng(9), 8, 60)
    err, passes = fused_vs_oracle(8, ops, 5)
    assert err < TOL and passes == 60


def test_fused_24q_against_tiled():
    """At 24 qubits the fused result stays within the tolerance of the exact
    tiled run (itself bit-equal to the oracle)."""
    n = 24
    ops, _ = circuits.brickwork(n, 4, seed=11)
    a = DenseHip(n)
    a.run(ops, mode=RUN_FUSED)
    b = DenseHip(n)
    b.run(ops, mode=RUN_TILED)
    chunk = 1 << 22
    err = max(float(np.max(np.abs(a.state(f, chunk) - b.state(f, chunk)))) for f in range(0, 1 << n, chunk))
    assert err < TOL, err


def test_fused_rejects_bad_config_and_block_states():
    g = DenseHip(12)
    with pytest.raises(Exception):
        g.set_run_config(fusion_width=3)
    with pytest.raises(Exception):
        g.set_run_config(tile_targets=65)
    b = DenseHip(14, block=True, block_size=12)
    ops = circuits.random_circuit(np.random.default_rng(1), 14, 10)
    with pytest.raises(Exception):
        b.run(ops, mode=RUN_FUSED)


@pytest.mark.parametrize("allow_inexact", [False, True])
def test_autotune_picks_a_measured_configuration(allow_inexact):
    n = 18
    ops, _ = circuits.brickwork(n, 6, seed=3)
    g = DenseHip(n)
    g.hadamard(0)
    before = g.state()
    mode, k, targets, ms = g.autotune(ops, allow_inexact=allow_inexact)
    print(f"autotune inexact={allow_inexact}: mode {mode} k {k} targets {targets} ms {ms}", flush=True)
    assert np.array_equal(g.state(), before)  # the state is not touched
    # per-gate, tiled (3 target caps); inexact: fused 4 and 5, reordered
    assert len(ms) == 7
    run = [m for i, m in enumerate(ms) if allow_inexact or i < 4]
    assert all(m > 0 for m in run)
    if not allow_inexact:
        assert mode in (RUN_PER_GATE, RUN_TILED) and ms[4] == ms[5] == ms[6] == 0
    # RUN_AUTO uses the choice; exact choices reproduce the oracle bit for bit
    g.clear()
    g.run(ops, mode=RUN_AUTO)
    o = oracle.DenseV2(n)
    o.run(ops)
    if mode in (RUN_FUSED, RUN_REORDER):
        assert float(np.max(np.abs(g.state() - o.state()))) < TOL
    else:
        assert np.array_equal(g.state(), o.state())


def test_tile_targets_cap_keeps_results_exact():
    n = 16
    ops = circuits.random_circuit(np.random.default_rng(4), n, 300, max_controls=2)
    o = oracle.DenseV2(n)
    o.run(ops)
    passes = []
    for cap in (64, 4, 1):
        g = DenseHip(n)
        g.set_run_config(tile_targets=cap)
        g.run(ops, mode=RUN_TILED)
        passes.append(g.run_stats()[1])
        assert np.array_equal(g.state(), o.state()), cap
    assert passes[0] < passes[1] < passes[2]
