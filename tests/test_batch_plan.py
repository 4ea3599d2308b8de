"""The batch planner on the host (kbwh_batch_plan needs no device): pass
counts, where Hamiltonian strings are evaluated, and plan errors."""
import pytest

from kbw_amd import circuits
from kbw_amd._lib import KbwhError, RY, X, Z
from kbw_amd.batch import Batch
from kbw_amd.circuits import OpList
from kbw_amd.dense import Hamiltonian

TILE = 11  # tile qubits; 0..2 are in every tile


def plan(circs):
    return Batch(circs, plan_only=True).plan


def empty_ops():
    return OpList().array()


def test_gate_free_circuit_gets_one_init_pass():
    g, p, a = plan([(3, empty_ops(), Hamiltonian([[("Z", 0)]], [1.0]), 0)])
    assert (g, p, a) == (0, 1, 1 << TILE)


def test_gate_free_circuit_without_strings():
    assert plan([(5, empty_ops(), Hamiltonian([], []), 0)]) == (0, 1, 1 << TILE)


def test_x_strings_spread_over_read_only_passes():
    # 20 single-X strings, no gates: the init pass holds qubits 0..10, the next
    # read-only pass 0..2 + 11..18, the last one qubit 19
    n = 20
    ham = Hamiltonian([[("X", q)] for q in range(n)], [1.0] * n)
    g, p, a = plan([(n, empty_ops(), ham, 0)])
    assert (g, p) == (0, 3)
    assert a == 1 << n  # the batch buffer: one 2^n slice


def test_diagonal_strings_need_no_extra_pass():
    n = 16
    ham = Hamiltonian([[("Z", q), ("Z", q + 1)] for q in range(n - 1)], [1.0] * (n - 1))
    ops = OpList()
    for q in range(n):
        ops.add(RY, q, angle=0.3)
    g, p, a = plan([(n, ops.array(), ham, 0)])
    gate_passes = plan([(n, ops.array(), Hamiltonian([], []), 0)])[1]
    assert p == gate_passes


def test_strings_on_last_pass_tile_need_no_extra_pass():
    # the last pass targets qubits 3..10, so X strings there ride along
    n = 14
    ops = OpList()
    for q in range(3, 11):
        ops.add(X, q)
    ham = Hamiltonian([[("X", q)] for q in range(11)], [0.5] * 11)
    assert plan([(n, ops.array(), ham, 0)])[1] == 1


def test_string_with_too_many_x_qubits_is_rejected():
    n = 20
    ham = Hamiltonian([[("X", q) for q in range(3, 12)]], [1.0])  # 9 X qubits above 2
    with pytest.raises(KbwhError):
        plan([(n, empty_ops(), ham, 0)])


def test_plan_errors():
    ops = OpList()
    ops.add(Z, 5)
    with pytest.raises(KbwhError):
        plan([(4, ops.array(), Hamiltonian([], []), 0)])  # target outside the circuit
    with pytest.raises(KbwhError):
        plan([(27, empty_ops(), Hamiltonian([], []), 0)])  # wider than a batch allows
    with pytest.raises(KbwhError):
        plan([(4, empty_ops(), Hamiltonian([[("Z", 4)]], [1.0]), 0)])
    with pytest.raises(KbwhError):
        plan([(4, empty_ops(), Hamiltonian([[("Z", 1), ("X", 1)]], [1.0]), 0)])
    ops = OpList()
    ops.add(RY, 0, param=3)
    with pytest.raises(KbwhError):
        plan([(4, ops.array(), Hamiltonian([], []), 2)])  # parameter index past the circuit's


def test_c3_plan_is_stable():
    circs, params = circuits.hea_batch(64, seed=12)
    g, p, a = plan(circs)
    assert g == sum(len(c[1]) for c in circs)
    assert len(params) == sum(c[3] for c in circs)
    # every circuit has at least one pass; the buffer holds one slice per circuit
    assert p >= len(circs)
    assert a == sum(1 << max(c[0], TILE) for c in circs)


@pytest.mark.parametrize("seed", range(6))
def test_random_batches_plan(seed):
    """Random batches (1..26 qubits, random gates with parameters, random
    Hamiltonians, gate-free circuits) plan without error, with at least one
    pass per circuit and 2^max(n, 11) amplitudes per circuit pass."""
    import numpy as np

    rng = np.random.default_rng(300 + seed)
    circs = []
    for _ in range(int(rng.integers(1, 12))):
        n = int(rng.integers(1, 27))
        ops = circuits.random_circuit(rng, n, int(rng.integers(0, 120)), max_controls=min(3, n - 1)) if n > 1 \
            else empty_ops()
        rot = np.isin(ops["kind"], [4, 5, 6, 7])
        ops["param"][rot] = np.arange(rot.sum())
        ham = circuits.random_hamiltonian(rng, n, int(rng.integers(1, 10)), max_weight=min(3, n))
        circs.append((n, ops, ham, int(rot.sum())))
    gates, passes, amps = plan(circs)
    assert passes >= len(circs)
    assert amps >= sum(2 ** max(c[0], TILE) for c in circs)
