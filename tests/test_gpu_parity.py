"""GPU parity: the HIP engine (through the C ABI) against the CPU oracle.

Bars: amplitudes bit-exact in value (the plain path reproduces the reference's
arithmetic; +0/-0 compare equal), measurement/sample outcomes and dump
entries bit-exact, measure_p1 within 1e-14 (the reduction order is the
engine's own; the reference's rayon order is schedule-dependent), exp_value
within 1e-10 relative.
"""
import math

import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, KbwhError, circuits
from kbw_amd import _lib

pytestmark = pytest.mark.gpu

X, Y, Z, H, P, RX, RY, RZ = range(8)


def ctrl_list(mask, n):
    return [q for q in range(n) if mask >> q & 1]


def random_state(rng, n):
    psi = rng.normal(size=2 ** n) + 1j * rng.normal(size=2 ** n)
    return psi / np.linalg.norm(psi)


def pair(n, psi=None):
    g, o = DenseHip(n), oracle.DenseV2(n)
    if psi is not None:
        g.set_state(psi)
        o.set_state(psi)
    return g, o


def assert_same(g, o):
    a, b = g.state(), o.state()
    bad = np.nonzero(a != b)[0]
    assert bad.size == 0, f"{bad.size} amplitudes differ, first {bad[:4]}: {a[bad[:4]]} vs {b[bad[:4]]}"


@pytest.mark.parametrize("n", [1, 2, 5, 9, 13])
@pytest.mark.parametrize("kind", range(8))
def test_single_gates_every_target_bit_exact(n, kind):
    print(f"gates n={n} kind={kind}", flush=True)
    rng = np.random.default_rng(n * 31 + kind)
    psi = random_state(rng, n)
    g, o = pair(n, psi)
    for t in range(n):
        ang = float(rng.uniform(-7, 7))
        nc = int(rng.integers(0, min(3, n - 1) + 1))
        others = [q for q in range(n) if q != t]
        ctrl = [int(c) for c in rng.choice(others, size=nc, replace=False)] if nc else []
        g.gate(kind, ang, t, ctrl)
        o.gate(kind, ang, t, ctrl)
        g.gate(kind, ang, t)
        o.gate(kind, ang, t)
    assert_same(g, o)


@pytest.mark.parametrize("n", [6, 7, 8, 12])
def test_low_target_gates_with_low_and_high_controls(n):
    """Targets 0..5 (one amplitude per lane, partner by cross-lane shuffle)
    under controls below bit 3 (tested per lane), at bits 3..5 (inserted into
    the index below the target or above it, which moves the partner's lane),
    at bit 6 and above and mixed; n=7 and 8 with every high bit a control
    leave one or two 64-amplitude segments per launch."""
    print(f"low-target gates n={n}", flush=True)
    rng = np.random.default_rng(500 + n)
    g, o = pair(n, random_state(rng, n))
    high = list(range(6, n))
    for kind in range(8):
        for t in range(6):
            low = [q for q in range(6) if q != t]
            mid = [q for q in low if q >= 3]
            sets = [[], [low[0]], low[1:3], high[:1], high, [low[2]] + high[-2:], mid, mid[:1] + [low[0]] + high[:1]]
            for ctrl in sets:
                ang = float(rng.uniform(-7, 7))
                g.gate(kind, ang, t, ctrl)
                o.gate(kind, ang, t, ctrl)
    assert_same(g, o)


def test_32_qubit_gates_over_more_than_2_31_amplitudes():
    """At 32 qubits an uncontrolled RZ spans 2^32 amplitudes and a low-target
    X 2^32 lanes: both run as grid-striding loops.  Two 4096-amplitude
    windows of the state are set (bits 0..11 vary; the second has bits 14, 17
    and 31 set) and checked bit-exactly against a 15-qubit oracle holding the
    same windows, with oracle qubits 12, 13, 14 standing for bits 14, 17, 31."""
    print("32-qubit gates", flush=True)
    n = 32
    rng = np.random.default_rng(32)
    wa, wb = random_state(rng, 12), random_state(rng, 12)
    first_b = (1 << 31) | (1 << 17) | (1 << 14)
    g = DenseHip(n)
    g.set_state(wa, first=0)
    g.set_state(wb, first=first_b)
    o = oracle.DenseV2(15)
    psi = np.zeros(1 << 15, dtype=complex)
    base_b = 0b111 << 12
    psi[:4096], psi[base_b:base_b + 4096] = wa, wb
    o.set_state(psi)
    gates = [(RZ, 0.7, 0, []), (RZ, -1.9, 31, []), (RZ, 2.3, 17, [3]),
             (X, 0.0, 2, []), (RY, 1.1, 5, [1, 14]), (H, 0.0, 0, [])]
    where = {31: 14, 17: 13, 14: 12}
    for kind, ang, t, ctrl in gates:
        g.gate(kind, ang, t, ctrl)
        o.gate(kind, ang, where.get(t, t), [where.get(c, c) for c in ctrl])
    ref = o.state()
    assert np.array_equal(g.state(0, 4096), ref[:4096])
    assert np.array_equal(g.state(first_b, 4096), ref[base_b:base_b + 4096])
    assert not np.any(g.state(1 << 20, 4096))
    g.close()


@pytest.mark.parametrize("n,ngates,seed", [(4, 300, 0), (10, 400, 1), (15, 300, 2), (20, 120, 3)])
def test_random_circuits_bit_exact(n, ngates, seed):
    print(f"random circuit n={n}", flush=True)
    rng = np.random.default_rng(seed)
    ops = circuits.random_circuit(rng, n, ngates, max_controls=4)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    assert_same(g, o)


def test_many_controls_and_control_on_target():
    n = 12
    rng = np.random.default_rng(77)
    g, o = pair(n, random_state(rng, n))
    for kind in range(8):
        ctrl = [q for q in range(n) if q != 5]
        g.gate(kind, 0.9, 5, ctrl)
        o.gate(kind, 0.9, 5, ctrl)
        g.gate(kind, 1.3, 2, [2, 7])  # control equal to the target is dropped
        o.gate(kind, 1.3, 2, [2, 7])
    assert_same(g, o)


def test_grover_12q_amplitudes_and_measurement():
    print("grover 12q", flush=True)
    n = 12
    ops, target = circuits.grover(n, 50, seed=0)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    assert_same(g, o)
    u = np.random.default_rng(1).random(n)
    qubits = list(range(n))
    assert g.measure(qubits, u) == o.measure(qubits, u)
    assert_same(g, o)


@pytest.mark.parametrize("n", [3, 11, 17])
def test_measure_p1_and_collapse(n):
    rng = np.random.default_rng(n)
    ops = circuits.random_circuit(rng, n, 60)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    for t in range(n):
        a, b = g.measure_p1(t), o.measure_p1(t)
        assert abs(a - b) <= 1e-14 * max(1.0, abs(b))
    p = 1 / math.sqrt(max(o.measure_p1(1), 1e-300))
    g.measure_collapse(1, True, p)
    o.measure_collapse(1, True, p)
    assert_same(g, o)


@pytest.mark.parametrize("n,seed", [(6, 0), (12, 1), (16, 2)])
def test_measure_outcomes_bit_exact(n, seed):
    rng = np.random.default_rng(seed + 500)
    ops = circuits.random_circuit(rng, n, 80)
    for trial in range(3):
        g, o = pair(n)
        g.run(ops)
        o.run(ops)
        qubits = [int(q) for q in rng.permutation(n)[: max(1, n // 2)]]
        u = rng.random(len(qubits))
        assert g.measure(qubits, u) == o.measure(qubits, u)
        assert np.max(np.abs(g.state() - o.state())) < 1e-13


@pytest.mark.parametrize("n,seed", [(5, 0), (10, 1), (14, 2)])
def test_exp_value(n, seed):
    rng = np.random.default_rng(seed + 900)
    ops = circuits.random_circuit(rng, n, 100)
    ham = circuits.random_hamiltonian(rng, n, 40)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    eg = g.exp_value(ham)
    eo = o.exp_value(ham.pauli_strings, ham.coefficients)
    assert abs(eg - eo) <= 1e-10 * abs(eo) + 1e-13
    # <I> = 1 and a product-state closed form
    from kbw_amd import Hamiltonian
    assert g.exp_value(Hamiltonian([[]], [1.0])) == pytest.approx(1.0, abs=1e-13)
    h2 = DenseHip(3)
    h2.ry(0.8, 0)
    h2.rx(-1.2, 2)
    e = h2.exp_value(Hamiltonian([[("Z", 0)], [("X", 0)], [("Z", 2)], [("Y", 2)]], [1.0, 1.0, 1.0, 1.0]))
    assert e == pytest.approx(math.cos(0.8) + math.sin(0.8) + math.cos(1.2) + math.sin(1.2), abs=1e-14)


@pytest.mark.parametrize("n", [1, 4, 6, 7, 12, 13, 18])
def test_sample_bit_exact(n):
    print(f"sample n={n}", flush=True)
    rng = np.random.default_rng(n + 40)
    ops = circuits.random_circuit(rng, n, 50)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    u = rng.random(20000)
    qubits = [int(q) for q in rng.permutation(n)]
    a = g.sample(qubits, len(u), u)
    b = o.sample_hier(qubits, len(u), u)
    assert np.array_equal(a, b)


def test_dump_bit_exact():
    rng = np.random.default_rng(3)
    n = 14
    ops = circuits.random_circuit(rng, n, 40)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    qubits = [3, 1, 13, 0]
    d = g.dump(qubits)
    k, re, im = o.dump(qubits)
    assert np.array_equal(d.keys, k) and np.array_equal(d.amplitudes_real, re) and np.array_equal(d.amplitudes_imag, im)
    # sparse state: only the nonzero amplitudes are dumped
    s = DenseHip(10)
    s.hadamard(4)
    s.pauli_x(7, [4])
    assert len(s.dump(list(range(10)))) == 2


@pytest.mark.parametrize("n,L,seed", [(6, 3, 0), (9, 5, 1), (12, 8, 2)])
def test_block_mode_matches_block_oracle(n, L, seed):
    print(f"block n={n} L={L}", flush=True)
    rng = np.random.default_rng(seed + 70)
    ops = circuits.random_circuit(rng, n, 150, max_controls=3)
    g = DenseHip(n, block=True, block_size=L)
    o = oracle.Block(n, L)
    g.run(ops)
    for op in ops:
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), ctrl_list(int(op["ctrl_mask"]), n))
    assert np.array_equal(g.qubit_map(), o.qubit_map())
    assert g.swap_count() == o.swap_count() > 0
    assert np.array_equal(g.state(), o.physical_state())
    qubits = list(range(n))
    d = g.dump(qubits)
    k, re, _ = o.dump(qubits)
    assert np.array_equal(d.keys, k) and np.array_equal(d.amplitudes_real, re)
    u = rng.random(5000)
    phys = [int(g.qubit_map()[q]) for q in qubits]
    samp = g.sample(qubits, len(u), u)
    ref = oracle.sample_hier(o.physical_state(), n, phys, len(u), u)
    assert np.array_equal(samp, ref)
    # the measure loop goes through the same LRU lookups
    um = rng.random(3)
    m = [int(x) for x in rng.permutation(n)[:3]]
    pg = [g.measure_p1(q) for q in m]
    po = [o.measure_p1(q) for q in m]
    assert np.allclose(pg, po, rtol=0, atol=1e-14)
    assert np.array_equal(g.qubit_map(), o.qubit_map())


def test_error_codes():
    with pytest.raises(KbwhError) as e:
        DenseHip(33)
    assert e.value.code == _lib.ERR_EXECUTION_FAILED
    with pytest.raises(KbwhError) as e:
        DenseHip(1, block=True)
    assert e.value.code == _lib.ERR_EXECUTION_FAILED
    g = DenseHip(4)
    with pytest.raises(KbwhError) as e:
        g.hadamard(4)
    assert e.value.code == _lib.ERR_INVALID_ARGUMENT
    with pytest.raises(KbwhError) as e:
        g.pauli_x(0, [9])
    assert e.value.code == _lib.ERR_INVALID_ARGUMENT
    g.measure_collapse(0, True, 1.0)  # |0000> projected on q0=1: all zero
    with pytest.raises(KbwhError) as e:
        g.sample([0], 4, np.zeros(4))
    assert e.value.code == _lib.ERR_ZERO_WEIGHTS


def test_full_size_30q_brickwork_prefix_bit_exact():
    """At the 30-qubit size: a prefix of the brickwork circuit, compared
    amplitude by amplitude with the oracle (16 GiB on each side)."""
    n = 30
    ops, ham = circuits.brickwork(n, 20, seed=1)
    pre = ops[:40]
    print("30q: gpu run", flush=True)
    g = DenseHip(n)
    g.run(pre)
    g.synchronize()
    print("30q: oracle run", flush=True)
    o = oracle.DenseV2(n)
    o.run(pre)
    print("30q: compare", flush=True)
    ref = o.state()
    chunk = 1 << 26
    for first in range(0, 1 << n, chunk):
        a = g.state(first, chunk)
        assert np.array_equal(a, ref[first:first + chunk]), f"chunk at {first}"
    del ref
    # a size-independent property: the norm (<I>) stays 1
    from kbw_amd import Hamiltonian
    assert g.exp_value(Hamiltonian([[]], [1.0])) == pytest.approx(1.0, abs=1e-12)


@pytest.mark.parametrize("n,seed", [(11, 0), (16, 1), (20, 2)])
def test_exp_value_grouped_sweeps(n, seed):
    """Flat states of 11+ qubits evaluate strings in grouped tile sweeps: many
    strings with X/Y qubits anywhere, Y factors, a string too wide for one
    tile (summed alone) and the identity, against the oracle."""
    from kbw_amd import Hamiltonian
    rng = np.random.default_rng(seed + 950)
    ops = circuits.random_circuit(rng, n, 120)
    ham = circuits.random_hamiltonian(rng, n, 70, max_weight=4)
    strings = ham.pauli_strings + [[("X", q) for q in range(3, min(n, 13))], [], [("Y", n - 1), ("Y", 0)]]
    coeffs = ham.coefficients + [0.37, -0.2, 1.1]
    h = Hamiltonian(strings, coeffs)
    g, o = pair(n)
    g.run(ops)
    o.run(ops)
    eg = g.exp_value(h)
    eo = o.exp_value(h.pauli_strings, h.coefficients)
    assert abs(eg - eo) <= 1e-10 * abs(eo) + 1e-13, (eg, eo)
    assert g.exp_value(h) == eg  # deterministic
