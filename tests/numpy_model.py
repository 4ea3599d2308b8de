"""Independent brute-force model for small registers (test checker).

Operators are built as full 2^n x 2^n matrices from Kronecker products of
the textbook 2x2 gate matrices (qubit 0 is the least significant index bit),
with controls as |1><1| projectors: U_full = I + (P_c (x) (U - I)).  This is
a different formulation from the oracle's pair loops, so a sign, index or
transposition mistake in either shows up as a mismatch.
"""
import numpy as np

I2 = np.eye(2, dtype=complex)
P1 = np.array([[0, 0], [0, 1]], dtype=complex)


def gate_matrix(kind, angle):
    c, s = np.cos(angle / 2), np.sin(angle / 2)
    return {
        0: np.array([[0, 1], [1, 0]], dtype=complex),
        1: np.array([[0, -1j], [1j, 0]], dtype=complex),
        2: np.array([[1, 0], [0, -1]], dtype=complex),
        3: np.array([[1, 1], [1, -1]], dtype=complex) / np.sqrt(2),
        4: np.array([[1, 0], [0, np.exp(1j * angle)]], dtype=complex),
        5: np.array([[c, -1j * s], [-1j * s, c]], dtype=complex),
        6: np.array([[c, -s], [s, c]], dtype=complex),
        7: np.array([[np.exp(-0.5j * angle), 0], [0, np.exp(0.5j * angle)]], dtype=complex),
    }[kind]


def full_operator(n, kind, angle, target, controls):
    u = gate_matrix(kind, angle)
    controls = [c for c in controls if c != target]
    op = np.ones((1, 1), dtype=complex)
    for q in reversed(range(n)):  # most significant qubit first in kron order
        if q == target:
            f = u - I2
        elif q in controls:
            f = P1
        else:
            f = I2
        op = np.kron(op, f)
    return np.eye(2 ** n, dtype=complex) + op


def apply(state, n, kind, angle, target, controls=()):
    return full_operator(n, kind, angle, target, list(controls)) @ state


def pauli_operator(n, string):
    mats = {"X": np.array([[0, 1], [1, 0]], dtype=complex),
            "Y": np.array([[0, -1j], [1j, 0]], dtype=complex),
            "Z": np.array([[1, 0], [0, -1]], dtype=complex)}
    per = {q: mats[p] for p, q in string}
    op = np.ones((1, 1), dtype=complex)
    for q in reversed(range(n)):
        op = np.kron(op, per.get(q, I2))
    return op


def expectation(state, n, hamiltonian):
    total = 0.0
    for s, c in zip(hamiltonian.pauli_strings, hamiltonian.coefficients):
        total += c * np.vdot(state, pauli_operator(n, s) @ state).real
    return total


def zero_state(n):
    s = np.zeros(2 ** n, dtype=complex)
    s[0] = 1
    return s


def run_ops(n, ops, params=None):
    s = zero_state(n)
    for o in ops:
        ang = float(o["angle"])
        if params is not None and int(o["param"]) >= 0:
            ang = float(params[int(o["param"])])
        m = int(o["ctrl_mask"])
        ctrl = [q for q in range(n) if m >> q & 1]
        s = apply(s, n, int(o["kind"]), ang, int(o["target"]), ctrl)
    return s
