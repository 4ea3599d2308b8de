"""Seeded synthetic workloads for the dense path, as kbwh_op arrays.

Each generator returns the same circuit for the same seed (numpy PCG64), so
the GPU engine and the CPU oracle run identical gate lists.  Gate order
inside a layer is fixed (ascending qubits) because gates on different qubits
commute exactly in value only when one of them is a permutation.
"""
import math

import numpy as np

from ._lib import OP_DTYPE, H, P, RX, RY, RZ, X, Y, Z
from .dense import Hamiltonian


class OpList:
    def __init__(self):
        self.rows = []

    def add(self, kind, target, control=(), angle=0.0, param=-1):
        mask = 0
        for c in control:
            mask |= 1 << int(c)
        self.rows.append((int(kind), int(target), 0, int(param), float(angle), mask))

    def array(self):
        a = np.zeros(len(self.rows), dtype=OP_DTYPE)
        for i, r in enumerate(self.rows):
            a[i] = r
        return a


# ---- C0: Grover, 12 qubits ------------------------------------------------
def grover(n=12, iterations=50, seed=0):
    """H on all; per iteration: oracle (X on the target's zero bits, Z on qubit 0
    controlled by the others, undo X) then diffusion (H, X per qubit ascending;
    the multi-controlled Z; X, H per qubit descending) -- the compute/action/
    uncompute shape of examples/grover_batch.rs:34-52 with a seeded target."""
    rng = np.random.default_rng(seed)
    target = int(rng.integers(0, 1 << n))
    ops = OpList()
    for q in range(n):
        ops.add(H, q)
    zeros = [q for q in range(n) if not (target >> q) & 1]
    ctrl = list(range(1, n))
    for _ in range(iterations):
        for q in zeros:
            ops.add(X, q)
        ops.add(Z, 0, ctrl)
        for q in zeros:
            ops.add(X, q)
        for q in range(n):
            ops.add(H, q)
            ops.add(X, q)
        ops.add(Z, 0, ctrl)
        for q in reversed(range(n)):
            ops.add(X, q)
            ops.add(H, q)
    return ops.array(), target


# ---- C1: random brickwork ---------------------------------------------------
ONE_QUBIT_CHOICES = ("H", "X", "Y", "Z", "P2", "P4", "RX", "RY", "RZ")


def brickwork_layers(ops, rng, n, layers, first_layer=0):
    for l in range(first_layer, first_layer + layers):
        for q in range(n):
            g = ONE_QUBIT_CHOICES[int(rng.integers(0, len(ONE_QUBIT_CHOICES)))]
            theta = float(rng.uniform(0.0, 2.0 * math.pi))
            if g == "H":
                ops.add(H, q)
            elif g == "X":
                ops.add(X, q)
            elif g == "Y":
                ops.add(Y, q)
            elif g == "Z":
                ops.add(Z, q)
            elif g == "P2":
                ops.add(P, q, angle=math.pi / 2)
            elif g == "P4":
                ops.add(P, q, angle=math.pi / 4)
            else:
                ops.add({"RX": RX, "RY": RY, "RZ": RZ}[g], q, angle=theta)
        start = 0 if l % 2 == 0 else 1
        for a in range(start, n - 1, 2):
            if rng.integers(0, 2) == 0:
                ops.add(Z, a + 1, [a])  # CZ
            else:
                ops.add(X, a + 1, [a])  # CNOT
        if (l + 1) % 5 == 0:
            for _ in range(8):
                t = int(rng.integers(0, n))
                nc = int(rng.integers(1, 4))
                others = [q for q in range(n) if q != t]
                ctrl = rng.choice(others, size=nc, replace=False)
                ops.add(X, t, [int(c) for c in ctrl])


def random_hamiltonian(rng, n, nstrings=64, max_weight=4):
    strings, coeffs = [], []
    for _ in range(nstrings):
        w = int(rng.integers(1, max_weight + 1))
        qs = rng.choice(n, size=w, replace=False)
        strings.append([("XYZ"[int(rng.integers(0, 3))], int(q)) for q in qs])
        coeffs.append(float(rng.normal(0.0, 1.0)))
    return Hamiltonian(strings, coeffs)


def brickwork(n=30, layers=20, seed=1, nstrings=64):
    """C1: per layer one gate per qubit from {H,X,Y,Z,P(pi/2),P(pi/4),RX,RY,RZ}
    (angles U[0,2pi)), CZ or CNOT on alternating even/odd neighbour pairs,
    8 multi-controlled X (1-3 controls) every 5th layer; then a Hamiltonian of
    random Pauli strings (weight 1-4, N(0,1) coefficients)."""
    rng = np.random.default_rng(seed)
    ops = OpList()
    brickwork_layers(ops, rng, n, layers)
    return ops.array(), random_hamiltonian(rng, n, nstrings)


# ---- C2: QFT + brickwork (Block) ------------------------------------------
def qft_plus_brickwork(n=33, layers=4, seed=2):
    """C2: QFT (H and controlled-phase ladders, reversal by CNOT swaps), then
    `layers` brickwork layers."""
    rng = np.random.default_rng(seed)
    ops = OpList()
    for j in reversed(range(n)):
        ops.add(H, j)
        for k in reversed(range(j)):
            ops.add(P, j, [k], angle=math.pi / (1 << (j - k)))
    for i in range(n // 2):
        a, b = i, n - 1 - i
        ops.add(X, b, [a])
        ops.add(X, a, [b])
        ops.add(X, b, [a])
    brickwork_layers(ops, rng, n, layers)
    return ops.array()


# ---- C3 / C4: hardware-efficient ansatz ---------------------------------
def hea(n, layers, seed, angle_low=-math.pi, angle_high=math.pi, ring=True):
    """Layers of RY then RZ on every qubit (one parameter each), then a CZ
    ring.  Returns (ops with param ids, params)."""
    rng = np.random.default_rng(seed)
    ops = OpList()
    nparam = 0
    for _ in range(layers):
        for q in range(n):
            ops.add(RY, q, param=nparam)
            nparam += 1
        for q in range(n):
            ops.add(RZ, q, param=nparam)
            nparam += 1
        if ring:
            for q in range(n if n > 2 else n - 1):
                ops.add(Z, (q + 1) % n, [q])
    params = rng.uniform(angle_low, angle_high, size=nparam)
    a = ops.array()
    a["angle"] = params[a["param"]] * (a["param"] >= 0)
    return a, params


def tfim(n, seed, j=1.0, h_mean=1.0, h_std=0.1):
    """Open transverse-field Ising chain: sum J Z_i Z_i+1 + sum h_i X_i."""
    rng = np.random.default_rng(seed)
    strings = [[("Z", i), ("Z", i + 1)] for i in range(n - 1)]
    coeffs = [j] * (n - 1)
    for i in range(n):
        strings.append([("X", i)])
        coeffs.append(float(rng.normal(h_mean, h_std)))
    return Hamiltonian(strings, coeffs)


def xxz(n, seed):
    """XXZ Heisenberg chain, n-1 bonds: Jxy (XX + YY) + Jz ZZ."""
    rng = np.random.default_rng(seed)
    strings, coeffs = [], []
    for i in range(n - 1):
        jxy = float(rng.normal(1.0, 0.1))
        jz = float(rng.normal(0.5, 0.1))
        strings += [[("X", i), ("X", i + 1)], [("Y", i), ("Y", i + 1)], [("Z", i), ("Z", i + 1)]]
        coeffs += [jxy, jxy, jz]
    return Hamiltonian(strings, coeffs)


def hea_batch(num=4096, seed=12, n_low=10, n_high=20, layers=8):
    """The batched workload: `num` hardware-efficient-ansatz circuits with
    qubit counts U{n_low..n_high}, `layers` layers of RY then RZ on every
    qubit (angles U[-pi, pi)) and a CZ ring per layer, each ending with the
    expectation value of an open transverse-field Ising chain (J = 1,
    h_i ~ N(1, 0.1)).  Returns (circuits, params) for kbw_amd.batch.Batch:
    circuits[i] = (n, ops, hamiltonian, num_params)."""
    rng = np.random.default_rng(seed)
    circuits, params = [], []
    for i in range(num):
        n = int(rng.integers(n_low, n_high + 1))
        ops, p = hea(n, layers, seed=int(rng.integers(0, 2 ** 31)))
        ham = tfim(n, seed=int(rng.integers(0, 2 ** 31)))
        circuits.append((n, ops, ham, len(p)))
        params.append(p)
    return circuits, np.concatenate(params)


def every_position(kind, n=13, params=False):
    """Tests: one gate of `kind` on each of qubits 0..10 (every position of an
    11-qubit tile: a pass's planner places all eleven), each uncontrolled, with
    a control on qubit 0 or 1, on qubit 3 or 4 and on qubit n - 1, with angles
    of every sign and magnitude form; params=True gives each rotation its own
    parameter (returns (ops, params))."""
    ops = OpList()
    vals = []
    for q in range(11):
        lane_c = [0] if q != 0 else [1]
        tile_c = [3] if q != 3 else [4]
        for ctrl, ang in zip(([], lane_c, tile_c, [n - 1]), (0.9, 2.5, 4.0, 5.9)):
            a = ang + 0.01 * q
            if params and kind in (P, RX, RY, RZ):
                ops.add(kind, q, ctrl, angle=a, param=len(vals))
                vals.append(a)
            else:
                ops.add(kind, q, ctrl, angle=a)
    return (ops.array(), np.array(vals)) if params else ops.array()


def random_circuit(rng, n, ngates, max_controls=3):
    """Random gates of every kind with random targets and controls (tests)."""
    ops = OpList()
    for _ in range(ngates):
        kind = int(rng.integers(0, 8))
        t = int(rng.integers(0, n))
        nc = int(rng.integers(0, min(max_controls, n - 1) + 1))
        others = [q for q in range(n) if q != t]
        ctrl = [int(c) for c in rng.choice(others, size=nc, replace=False)] if nc else []
        ops.add(kind, t, ctrl, angle=float(rng.uniform(-2 * math.pi, 2 * math.pi)))
    return ops.array()


# ---- algorithmic bytes (per launch) ----------------------------------------
def gate_bytes(n, kind, ctrl_mask, target):
    """HBM bytes the plain path must move for one gate: read + write of each
    touched amplitude (16 B each way)."""
    c = bin(int(ctrl_mask) & ~(1 << int(target))).count("1")
    if kind in (Z, P):
        return 32 * 2 ** (n - c - 1)
    return 32 * 2 ** (n - c)


def circuit_bytes(n, ops):
    return sum(gate_bytes(n, int(o["kind"]), int(o["ctrl_mask"]), int(o["target"])) for o in ops)
