"""Failure handling on the device side: out-of-memory from real and injected
allocation failures (kbwh_inject_alloc_failure), and recovery afterwards.
The reference's error sources are dense_v2.rs:155-157 (too many qubits) and
convert.rs:60 (zero weights); KBWH_ERR_OUT_OF_MEMORY has no reference
counterpart (the reference aborts on allocation failure)."""
import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, KbwhError, _lib, circuits
from kbw_amd.batch import Batch

pytestmark = pytest.mark.gpu

OOM = 3


def inject(after):
    _lib.check(_lib.load().kbwh_inject_alloc_failure(after))


@pytest.fixture(autouse=True)
def no_pending_failure():
    yield
    inject(-1)


def test_real_out_of_memory_is_reported():
    """A 40-qubit Block state needs 16 TiB: the allocation fails and the
    handle is not created."""
    print("40-qubit block state", flush=True)
    with pytest.raises(KbwhError) as ei:
        DenseHip(40, block=True, block_size=30)
    assert ei.value.code == OOM
    s = DenseHip(12)
    s.hadamard(3)
    assert abs(s.measure_p1(3) - 0.5) < 1e-15
    s.close()


@pytest.mark.parametrize("after", [0, 1, 2])
def test_injected_failure_in_kbwh_new(after):
    print(f"inject new {after}", flush=True)
    inject(after)
    with pytest.raises(KbwhError) as ei:
        DenseHip(14)
    assert ei.value.code == OOM
    s = DenseHip(14)  # the failure was consumed
    s.pauli_x(2)
    assert s.measure_p1(2) == 1.0
    s.close()


def test_injected_failure_in_growth_then_recovery():
    """The adjoint's vectors, then a sample's level buffers: each call fails
    with OUT_OF_MEMORY once and gives the right result when repeated."""
    print("inject growth", flush=True)
    n = 13
    ops, params = circuits.hea(n, 2, seed=3)
    ham = circuits.tfim(n, seed=4)
    s = DenseHip(n)
    inject(0)
    with pytest.raises(KbwhError) as ei:
        s.adjoint_gradient(ops, params, ham)
    assert ei.value.code == OOM
    e, g = s.adjoint_gradient(ops, params, ham)
    o = oracle.DenseV2(n)
    o.run(ops, params)
    assert abs(e - o.exp_value(ham.pauli_strings, ham.coefficients)) <= 1e-10
    s.run(ops, params)
    u = np.random.default_rng(5).random(64)
    inject(0)
    try:
        first = s.sample(list(range(n)), 64, u)
    except KbwhError as err:
        assert err.code == OOM
        first = None
    inject(-1)
    again = s.sample(list(range(n)), 64, u)
    if first is not None:  # the buffers already existed: nothing was allocated
        assert np.array_equal(first, again)
    assert np.array_equal(again, oracle.sample_hier(s.state(), n, list(range(n)), 64, u))
    s.close()


@pytest.mark.parametrize("after", [0, 3, 6, 9])
def test_injected_failure_in_batch_new(after):
    """The batch's allocations (states, tables, group tables) failing one at a
    time: OUT_OF_MEMORY and a clean teardown; then the same batch builds and
    matches the oracle."""
    print(f"inject batch {after}", flush=True)
    rng = np.random.default_rng(40 + after)
    circs, params = [], []
    for n in [12] * 130 + [11, 13]:  # 130 x 2 tiles: one structure group
        ops, p = circuits.hea(n, 2, seed=int(rng.integers(1 << 30)))
        circs.append((n, ops, circuits.tfim(n, seed=int(rng.integers(1 << 30))), len(p)))
        params.append(p)
    params = np.concatenate(params)
    inject(after)
    try:
        Batch(circs)
        built = True
    except KbwhError as err:
        assert err.code == OOM
        built = False
    inject(-1)
    assert not built  # a batch with a group makes more than 9 allocations
    b = Batch(circs)
    out = b.run(params)
    off = b.param_offsets
    assert len(b.groups()) == 1
    for i in (0, 77, 130, 131):
        n, ops, ham, _ = circs[i]
        o = oracle.DenseV2(n)
        o.run(ops, params[off[i]:off[i + 1]])
        e = o.exp_value(ham.pauli_strings, ham.coefficients)
        assert abs(out[i] - e) <= 1e-11 * max(1.0, abs(e))
    b.close()
