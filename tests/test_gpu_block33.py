"""The blocked QFT + brickwork workload (33 qubits, 128 GiB, blocks of 2^28,
5 global qubits, 10^6 shots), checked on the GPU.

The oracle's Block cannot hold 33 qubits on the host (its out-of-place swap
needs two 128 GiB copies), so the same circuit shape runs bit-exactly against
the oracle Block at 23 qubits (5 global qubits), and the 33-qubit run is
checked through size-independent properties: QFT|0> is exactly uniform, the
norm stays 1, the brickwork layers followed by their inverse return to the
uniform state, the Block qubit map / swap count follow block.rs's LRU rule,
and sampled marginals follow measure_p1.
"""
import math

import numpy as np
import pytest

import oracle
from kbw_amd import DenseHip, Hamiltonian, RUN_TILED, circuits

pytestmark = pytest.mark.gpu

X, Y, Z, H, P, RX, RY, RZ = range(8)


def inverse(ops):
    inv = ops[::-1].copy()
    rot = np.isin(inv["kind"], [P, RX, RY, RZ])
    inv["angle"][rot] = -inv["angle"][rot]
    return inv


def qft_count(n):
    return n + n * (n - 1) // 2 + 3 * (n // 2)


def lru_replay(n, L, ops):
    """block.rs:104-156 on the map only: (logical_to_physical, swaps)."""
    l2p, p2l = list(range(n)), list(range(n))
    last, counter, swaps = [0] * L, 0, 0
    for o in ops:
        t = int(o["target"])
        phys = l2p[t]
        if phys < L:
            counter += 1
            last[phys] = counter
            continue
        lq = min(range(L), key=lambda i: (last[i], i))
        old = p2l[lq]
        l2p[t], p2l[lq] = lq, t
        l2p[old], p2l[phys] = phys, old
        counter += 1
        last[lq] = counter
        swaps += 1
    return l2p, swaps


def test_block_23q_same_shape_bit_exact():
    n, L = 23, 18
    ops = circuits.qft_plus_brickwork(n, 4, seed=2)
    print(f"block23: {len(ops)} gates", flush=True)
    g = DenseHip(n, block=True, block_size=L)
    g.run(ops, mode=RUN_TILED)
    o = oracle.Block(n, L)
    for op in ops:
        m = int(op["ctrl_mask"])
        o.gate(int(op["kind"]), float(op["angle"]), int(op["target"]), [q for q in range(n) if m >> q & 1])
    assert np.array_equal(g.qubit_map(), o.qubit_map())
    assert g.swap_count() == o.swap_count() > 0
    phys = o.physical_state()
    assert np.array_equal(g.state(), phys)
    u = np.random.default_rng(5).random(100000)
    qubits = list(range(n))
    pq = [int(g.qubit_map()[q]) for q in qubits]
    assert np.array_equal(g.sample(qubits, len(u), u), oracle.sample_hier(phys, n, pq, len(u), u))


def test_block_33q_properties():
    n, L = 33, 28
    ops = circuits.qft_plus_brickwork(n, 4, seed=2)
    nq = qft_count(n)
    qft, layers = ops[:nq], ops[nq:]
    print(f"block33: qft {len(qft)} gates, layers {len(layers)} gates", flush=True)
    g = DenseHip(n, block=True, block_size=L)
    g.run(qft, mode=RUN_TILED)
    # QFT|0> is the uniform superposition; every amplitude went through the
    # same 33 products by 1/sqrt(2), so all are the same double
    f = 0.70710678118654752440
    ref = 1.0
    for _ in range(n):
        ref = ref * f
    for first in (0, 1 << 20, (1 << 32) + 12345, (1 << 33) - (1 << 16)):
        a = g.state(first, 1 << 16)
        assert np.all(a == ref), (first, a[:4], ref)
    print("block33: qft uniform ok", flush=True)
    g.run(layers, mode=RUN_TILED)
    assert g.exp_value(Hamiltonian([[]], [1.0])) == pytest.approx(1.0, abs=1e-12)
    l2p, swaps = lru_replay(n, L, ops)
    assert list(g.qubit_map()) == l2p and g.swap_count() == swaps
    # sampling 10^6 shots: marginals of the low qubits follow measure_p1
    shots = 1_000_000
    u = np.random.default_rng(9).random(shots)
    qubits = list(range(n))
    keys = g.sample(qubits, shots, u)
    assert keys.max() < (1 << n)
    for q in (0, 17, 32):
        p1 = g.measure_p1(q)
        freq = np.mean((keys >> np.uint64(n - 1 - q)) & np.uint64(1))
        assert abs(freq - p1) < 6 * math.sqrt(p1 * (1 - p1) / shots) + 1e-9, (q, freq, p1)
    print("block33: sampling ok", flush=True)
    # the inverse layers return to the uniform state
    g.run(inverse(layers), mode=RUN_TILED)
    l2p = g.qubit_map()
    for first in (0, (1 << 32) + 777):
        a = g.state(first, 1 << 16)
        assert np.max(np.abs(a - ref)) < 1e-14
    print("block33: inverse ok", flush=True)
