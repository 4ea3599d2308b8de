"""exp_value sweeps compiled per string group (kbwh_set_compiled) against the
generic sweep kernel and the oracle's default exp_value
(quantum_execution.rs:109-158, restated in oracle/).

A compiled sweep forms conj(a_{i^x}) a_i once per distinct x of its group
(the partner from a register, a lane move or the tile in LDS, chosen at
compile time) and adds each string's terms with compile-time signs, so it is
the generic sweep's sum in another order: equal to rounding (1e-12 relative
here), and the oracle's to 1e-10 relative.  The strings put X/Y factors on
every kind of tile position (lane bits, wave bits, register bits), in pairs
across kinds, with Y and Z factors, plus diagonal strings and the identity.
"""
import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, Hamiltonian, circuits, compiled_stats

pytestmark = pytest.mark.gpu


def crafted_strings(n, rng):
    s = []
    for q in range(n):
        s.append([("X", q)])
        s.append([("Y", q), ("Z", (q + 5) % n)])  # n > 5: a different qubit
    pairs = [(3, 4), (1, 5), (0, 3), (2, 6), (4, 7), (6, 7), (7, 9), (8, 10), (0, n - 1), (5, n - 2), (1, 2)]
    for a, b in pairs:
        if a < n and b < n and a != b:
            s.append([("X", a), ("Y", b)])
            z = next(q for q in range((a + b) % n, (a + b) % n + n) if q % n not in (a, b)) % n
            s.append([("Y", a), ("X", b), ("Z", z)])
    for _ in range(10):
        qs = rng.choice(n, size=3, replace=False)
        s.append([("Z", int(q)) for q in qs])
    s.append([])
    return s


@pytest.mark.parametrize("n,seed", [(11, 3), (14, 4), (20, 5)])
def test_compiled_sweeps_against_generic_and_oracle(n, seed):
    print(f"compiled sweeps n={n}", flush=True)
    rng = np.random.default_rng(seed + 1200)
    ops = circuits.random_circuit(rng, n, 150)
    strings = crafted_strings(n, rng)
    ham = circuits.random_hamiltonian(rng, n, 40, max_weight=4)
    strings += ham.pauli_strings
    coeffs = list(rng.normal(size=len(strings) - len(ham.pauli_strings))) + list(ham.coefficients)
    h = Hamiltonian(strings, coeffs)
    g = DenseHip(n)
    g.run(ops)
    o = oracle.DenseV2(n)
    o.run(ops)
    generic = g.exp_value(h)
    before = compiled_stats()
    g.set_compiled(True)
    comp = g.exp_value(h)
    after = compiled_stats()
    eo = o.exp_value(h.pauli_strings, h.coefficients)
    assert after["failures"] == before["failures"]
    assert after["compiled"] + after["cache_hits"] > before["compiled"] + before["cache_hits"]
    assert abs(comp - generic) <= 1e-12 * max(1.0, abs(generic)), (comp, generic)
    assert abs(comp - eo) <= 1e-10 * abs(eo) + 1e-13, (comp, eo)
    assert g.exp_value(h) == comp  # deterministic, kernels reused
    if n == 14:
        # string by string: a group of one string each
        for s, c in zip(strings[:60], coeffs[:60]):
            one = Hamiltonian([s], [c])
            ec = g.exp_value(one)
            g.set_compiled(False)
            eg = g.exp_value(one)
            g.set_compiled(True)
            e1 = o.exp_value([s], [c])
            assert abs(ec - eg) <= 1e-12 * max(1.0, abs(eg)), (s, ec, eg)
            assert abs(ec - e1) <= 1e-10 * abs(e1) + 1e-13, (s, ec, e1)
    g.close()
