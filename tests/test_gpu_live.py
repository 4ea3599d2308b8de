"""Queued live gates (KBWH_FLAG_QUEUE): gates given one by one are applied in
one launch (up to 12 qubits) or in tiled passes before the state is read, with
the same result bits as gate by gate, against the oracle."""
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
    with pytest.raises(KbwhError):
        DenseHip(8, block=True, block_size=6, queue=True)
