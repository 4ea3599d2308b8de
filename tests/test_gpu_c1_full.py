"""The 30-qubit brickwork workload on the path that bench.py times, end to
end against the oracle.

bench.py's value is the 30-qubit, 20-layer brickwork circuit (seed 1, 922
gates) run in KBWH_RUN_REORDER: dependency-ordered passes in the contracted
pass kernel k_tile_fma, after uncontrolled diagonal gates are fused.  Here
all 922 gates run that way on the GPU, and the oracle runs them in program
order (dense_v2.rs:48-147, restated in oracle/dense_v2.c).  Bars:
  * amplitudes within 1e-12 max-abs over all 2^30 (compared in 2^26 chunks);
  * exp_value within 1e-10 relative, per string, for a seeded subset of the
    circuit's 64-string Hamiltonian (Y strings, multi-X strings) plus strings
    too wide for one tile (summed alone by k_pauli_partial), against the
    oracle's default exp_value (quantum_execution.rs:109-158) in its chunked
    form (the same terms; 2^20-amplitude chunks summed in order).
The oracle needs 16 GiB of host memory and runs the 922 gates in about
1.5 minutes on 16 host threads.
"""
import time

import numpy as np
import pytest

import oracle
from kbw_amd import RUN_REORDER, DenseHip, Hamiltonian, circuits

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[c1-full {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def _string_subset(ham, n):
    """Twelve strings of the circuit's Hamiltonian (every Y string and every
    multi-X string among them kept) plus two strings wider than a tile."""
    rng = np.random.default_rng(2024)
    idx = sorted(int(i) for i in rng.choice(len(ham.pauli_strings), size=12, replace=False))
    strings = [ham.pauli_strings[i] for i in idx]
    coeffs = [ham.coefficients[i] for i in idx]
    wide_x = [("X", q) for q in range(3, 15)]  # 12 X qubits above qubit 2
    wide_mixed = [("Y", 29), ("X", 27), ("Z", 5)] + [("X", q) for q in range(16, 25)] + [("Y", 1)]
    strings += [wide_x, wide_mixed]
    coeffs += [0.41, -0.73]
    letters = ["".join(p for p, _ in s) for s in strings]
    assert any("Y" in l for l in letters[:12]), "subset has a Y string"
    assert any(l.count("X") >= 2 for l in letters[:12]), "subset has a multi-X string"
    assert all(q < n for s in strings for _, q in s)
    return Hamiltonian(strings, coeffs)


@pytest.mark.timeout(900)
def test_brickwork_30q_reorder_all_922_gates_against_oracle():
    n = 30
    ops, ham = circuits.brickwork(n, 20, seed=1)
    assert len(ops) == 922
    _log("gpu: 922 gates in dependency-ordered passes")
    g = DenseHip(n)
    g.run(ops, mode=RUN_REORDER)
    g.synchronize()
    gates, passes = g.run_stats()
    assert gates == 922 and passes <= 21, (gates, passes)
    _log(f"gpu: done in {passes} passes; oracle: 922 gates in program order")
    o = oracle.DenseV2(n)
    t0 = time.time()
    for first in range(0, len(ops), 100):
        o.run(ops[first:first + 100])
        _log(f"oracle: {min(first + 100, len(ops))}/922 gates, {time.time() - t0:.0f} s")
    chunk = 1 << 26
    worst = 0.0
    for first in range(0, 1 << n, chunk):
        a = g.state(first, chunk)
        d = float(np.max(np.abs(a - o.state_view(first, chunk))))
        worst = max(worst, d)
        assert d <= 1e-12, f"chunk at {first}: max |diff| {d}"
    _log(f"amplitudes: max |gpu - oracle| = {worst:.3e} over 2^30")
    h = _string_subset(ham, n)
    eo_all = 0.0
    for s, c in zip(h.pauli_strings, h.coefficients):
        eg = g.exp_value(Hamiltonian([s], [c]))
        eo = o.exp_value_chunked([s], [c])
        eo_all += eo
        _log(f"string {''.join(p + str(q) for p, q in s)}: gpu {eg:.15e} oracle {eo:.15e}")
        assert abs(eg - eo) <= 1e-10 * abs(eo) + 1e-13, (s, eg, eo)
    # all strings at once: the grouped sweeps plus the wide strings alone
    eg_all = g.exp_value(h)
    assert abs(eg_all - eo_all) <= 1e-10 * abs(eo_all) + 1e-13, (eg_all, eo_all)
    _log(f"exp_value of {len(h.coefficients)} strings: gpu {eg_all:.15e} oracle {eo_all:.15e}")
    g.close()
