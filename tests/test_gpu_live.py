"""Queued live gates (KBWH_FLAG_QUEUE): gates given one by one are applied in
one launch (up to 12 qubits) or in tiled passes before the state is read, with
the same result bits as gate by gate, against the oracle."""
import math

import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, KbwhError, circuits

pytestmark = pytest.mark.gpu


def ctrl_list(mask, n):
    return [q for q in range(n) if mask >> q & 1]


def live(sim, ops, n):
    for op in ops:
        sim.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), ctrl_list(int(op["ctrl_mask"]), n))


@pytest.mark.parametrize("n", [1, 2, 5, 9, 11, 12, 14])
def test_queued_gates_bit_exact(n):
    print(f"queued n={n}", flush=True)
    rng = np.random.default_rng(70 + n)
    ops = circuits.random_circuit(rng, n, 300, max_controls=min(3, n - 1))
    q = DenseHip(n, queue=True)
    live(q, ops, n)
    o = oracle.DenseV2(n)
    o.run(ops)
    assert np.array_equal(q.state(), o.state())
    gates, passes = q.run_stats()
    assert gates == len(ops)
    if n <= 12:
        assert passes == 1  # the whole queue in one launch


def test_queued_grover_12q_with_measurement():
    print("queued grover 12q", flush=True)
    n = 12
    ops, target = circuits.grover(n, 50, seed=0)
    q = DenseHip(n, queue=True)
    live(q, ops, n)
    o = oracle.DenseV2(n)
    o.run(ops)
    st = q.state()
    assert np.array_equal(st, o.state())
    assert int(np.argmax(np.abs(st))) == target
    u = np.random.default_rng(1).random(n)
    qubits = list(range(n))
    assert q.measure(qubits, u) == o.measure(qubits, u)
    assert np.array_equal(q.state(), o.state())


def test_reads_between_queued_gates():
    """Every read flushes first: p1, collapse, exp_value, dump, sample and
    run see exactly the gates given before them."""
    n = 10
    rng = np.random.default_rng(3)
    a_ops = circuits.random_circuit(rng, n, 80)
    b_ops = circuits.random_circuit(rng, n, 80)
    q = DenseHip(n, queue=True)
    g = DenseHip(n)
    live(q, a_ops, n)
    live(g, a_ops, n)
    assert q.measure_p1(3) == g.measure_p1(3)
    live(q, b_ops, n)
    live(g, b_ops, n)
    ham = circuits.random_hamiltonian(rng, n, 12)
    assert q.exp_value(ham) == g.exp_value(ham)
    live(q, a_ops, n)
    live(g, a_ops, n)
    d1, d2 = q.dump([0, 4, 7]), g.dump([0, 4, 7])
    assert np.array_equal(d1.amplitudes_real, d2.amplitudes_real)
    live(q, b_ops, n)
    live(g, b_ops, n)
    u = rng.random(500)
    assert np.array_equal(q.sample(list(range(n)), 500, u), g.sample(list(range(n)), 500, u))
    live(q, a_ops, n)
    q.run(b_ops)  # the queue goes first
    live(g, a_ops, n)
    g.run(b_ops)
    assert np.array_equal(q.state(), g.state())


def test_clear_discards_the_queue_and_errors_are_immediate():
    q = DenseHip(6, queue=True)
    q.hadamard(0)
    q.clear()
    st = q.state()
    assert st[0] == 1 and np.count_nonzero(st) == 1
    with pytest.raises(KbwhError):
        q.hadamard(6)  # target out of range, reported when given
    # a Block state's queue and map go back to the start too
    b = DenseHip(8, block=True, block_size=5, queue=True)
    b.hadamard(7)
    b.pauli_x(6, [7])
    assert b.swap_count() == 2 and list(b.qubit_map()) != list(range(8))
    b.clear()
    st = b.state()
    assert st[0] == 1 and np.count_nonzero(st) == 1
    assert list(b.qubit_map()) == list(range(8))


def block_live(o, ops, n):
    for op in ops:
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), ctrl_list(int(op["ctrl_mask"]), n))


@pytest.mark.parametrize("n,L", [(6, 3), (9, 5), (12, 8), (16, 11), (20, 14)])
def test_queued_block_gates_bit_exact(n, L):
    """Block's gate queue (block.rs:179-208): the map, LRU stamps and swap
    count follow each gate as it is given (compared before anything is
    flushed); at the flush, gates and swaps run in exact tiled passes, and the
    physical state, dump and samples equal the oracle Block's bit for bit."""
    print(f"queued block n={n} L={L}", flush=True)
    rng = np.random.default_rng(500 + n)
    ops = circuits.random_circuit(rng, n, 240, max_controls=3)
    q = DenseHip(n, block=True, block_size=L, queue=True)
    live(q, ops, n)
    o = oracle.Block(n, L)
    block_live(o, ops, n)
    assert np.array_equal(q.qubit_map(), o.qubit_map())
    assert q.swap_count() == o.swap_count() > 0
    st = q.state()  # flushes
    assert np.array_equal(st, o.physical_state())
    gates, passes = q.run_stats()
    assert gates == len(ops)
    if n >= 11:
        assert passes < len(ops) // 4, passes  # many gates and swaps per pass
    qubits = list(range(n))
    d = q.dump(qubits)
    k, re, im = o.dump(qubits)
    assert np.array_equal(d.keys, k) and np.array_equal(d.amplitudes_real, re)
    assert np.array_equal(d.amplitudes_imag, im)
    live(q, ops[:60], n)
    block_live(o, ops[:60], n)
    u = rng.random(4000)
    phys = [int(x) for x in o.qubit_map()]
    samp = q.sample(qubits, len(u), u)  # flushes the 60 new gates first
    assert np.array_equal(samp, oracle.sample_hier(o.physical_state(), n, phys, len(u), u))


def oracle_block_measure(o, qubits, uniforms):
    """QubitManager::do_measure (quantum_execution.rs:272-292) over the oracle
    Block: reverse order, p0 = max(1 - p1, 0), rand's WeightedIndex pick."""
    out = 0
    for idx, q in enumerate(reversed(qubits)):
        p1 = o.measure_p1(q)
        p0 = max(1.0 - p1, 0.0)
        x = uniforms[idx] * oracle.uniform_scale(p0 + p1) + 0.0
        r = 1 if p0 <= x else 0
        prob = p1 if r else p0
        o.measure_collapse(q, bool(r), 1.0 / math.sqrt(max(prob, 2.2250738585072014e-308)))
        out |= r << idx
    return out


def test_queued_block_measure_loop():
    n, L = 14, 10
    rng = np.random.default_rng(88)
    ops = circuits.random_circuit(rng, n, 200, max_controls=2)
    q = DenseHip(n, block=True, block_size=L, queue=True)
    o = oracle.Block(n, L)
    for rnd in range(3):
        live(q, ops[rnd * 60:(rnd + 1) * 60], n)
        block_live(o, ops[rnd * 60:(rnd + 1) * 60], n)
        m = [int(x) for x in rng.permutation(n)[:5]]
        u = rng.random(5)
        assert q.measure(m, u) == oracle_block_measure(o, m, u)
        assert np.array_equal(q.qubit_map(), o.qubit_map())
        assert q.swap_count() == o.swap_count()
        # the collapse factors come from each side's own p1 (equal to a few
        # ulps), so the states agree to rounding after a measurement
        assert np.max(np.abs(q.state() - o.physical_state())) < 1e-13


def test_queued_block_collapse_before_queued_gates():
    """block.rs:311-316: measure_collapse does not flush the queue, so gates
    given before it are applied after it (the reference's order); a global
    target first flushes and swaps.  Same factor p on both sides: bit-exact."""
    n, L = 12, 8
    rng = np.random.default_rng(91)
    ops = circuits.random_circuit(rng, n, 120, max_controls=2)
    q = DenseHip(n, block=True, block_size=L, queue=True)
    o = oracle.Block(n, L)
    live(q, ops, n)
    block_live(o, ops, n)
    local_q = next(l for l in range(n) if int(o.qubit_map()[l]) < L)
    global_q = next(l for l in range(n) if int(o.qubit_map()[l]) >= L)
    for tq, res in ((local_q, True), (global_q, False)):
        q.measure_collapse(tq, res, 1.75)
        o.measure_collapse(tq, res, 1.75)
        more = circuits.random_circuit(rng, n, 30, max_controls=2)
        live(q, more, n)
        block_live(o, more, n)
    assert np.array_equal(q.qubit_map(), o.qubit_map())
    assert q.swap_count() == o.swap_count()
    assert np.array_equal(q.state(), o.physical_state())


@pytest.mark.parametrize("n", [12, 13])
@pytest.mark.parametrize("kind", range(8))
def test_queued_every_kind_every_position(kind, n):
    """Queued gates of one kind on each of qubits 0..10 under every control
    placement: the 12-qubit whole-state launch and the 13-qubit tiled passes
    (qubits 3..10 take register, lane and wave positions of the tile), bit for
    bit against the oracle."""
    print(f"queued every position kind={kind} n={n}", flush=True)
    ops = circuits.every_position(kind, n)
    rng = np.random.default_rng(300 + 10 * kind + n)
    st = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    st /= np.linalg.norm(st)
    q = DenseHip(n, queue=True)
    q.set_state(st)
    live(q, ops, n)
    o = oracle.DenseV2(n)
    o.set_state(st)
    o.run(ops)
    assert np.array_equal(q.state(), o.state())
    q.close()
