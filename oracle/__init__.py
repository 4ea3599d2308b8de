"""CPU oracle: ctypes binding of oracle/liboracle.so.

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's DenseV2<f64>
(src/dense_v2.rs), Dense<f64> (src/dense.rs), the QubitManager measure loop
and trait defaults (src/quantum_execution.rs, src/convert.rs) and
Block<(), DenseV2> (src/block.rs) in plain C; see each C function's
citation.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg import this package, as the checker (or the timed CPU baseline); the
product (kbw_amd, libkbwh.so) never does.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

X, Y, Z, H, P, RX, RY, RZ = range(8)
PAULI = {"X": 1, "Y": 2, "Z": 3}
OP_DTYPE = np.dtype([("kind", "u1"), ("target", "u1"), ("reserved", "<u2"), ("param", "<i4"),
                     ("angle", "<f8"), ("ctrl_mask", "<u8")], align=True)

_vp, _u32, _u64, _f64, _i32, _p = C.c_void_p, C.c_uint32, C.c_uint64, C.c_double, C.c_int, C.c_void_p
_SIG = {
    "orc_v2_new": ([_u32], _vp), "orc_v2_free": ([_vp], None), "orc_v2_data": ([_vp], _vp),
    "orc_v2_clear": ([_vp], None), "orc_v2_gate_mask": ([_vp, _i32, _f64, _u32, _u64], None),
    "orc_v2_pauli_x": ([_vp, _u32, _p, _u32], None), "orc_v2_pauli_y": ([_vp, _u32, _p, _u32], None),
    "orc_v2_pauli_z": ([_vp, _u32, _p, _u32], None), "orc_v2_hadamard": ([_vp, _u32, _p, _u32], None),
    "orc_v2_phase": ([_vp, _f64, _u32, _p, _u32], None), "orc_v2_rx": ([_vp, _f64, _u32, _p, _u32], None),
    "orc_v2_ry": ([_vp, _f64, _u32, _p, _u32], None), "orc_v2_rz": ([_vp, _f64, _u32, _p, _u32], None),
    "orc_v2_run": ([_vp, _p, _u64, _p], None), "orc_v2_measure_p1": ([_vp, _u32], _f64),
    "orc_v2_measure_collapse": ([_vp, _u32, _i32, _f64], None),
    "orc_v2_dump": ([_vp, _p, _u32, _p, _p, _p], _u64),
    "orc_v1_new": ([_u32], _vp), "orc_v1_free": ([_vp], None), "orc_v1_data": ([_vp], _vp),
    "orc_v1_gate_mask": ([_vp, _i32, _f64, _u32, _u64], None), "orc_v1_measure_p1": ([_vp, _u32], _f64),
    "orc_v1_measure_collapse": ([_vp, _u32, _i32, _f64], None),
    "orc_uniform_scale": ([_f64], _f64), "orc_measure": ([_vp, _p, _u32, _p], _u64),
    "orc_exp_value": ([_vp, _p, _p, _p, _p, _u32], _f64),
    "orc_exp_value_chunked": ([_vp, _p, _p, _p, _p, _u32], _f64),
    "orc_sample_sequential": ([_vp, _p, _u32, _u64, _p, _p], _i32),
    "orc_sample_hier": ([_p, _u32, _p, _u32, _u64, _p, _p], _i32),
    "orc_block_new": ([_u32, _u32], _vp), "orc_block_free": ([_vp], None),
    "orc_block_gate": ([_vp, _i32, _f64, _u32, _p, _u32], None),
    "orc_block_measure_p1": ([_vp, _u32], _f64), "orc_block_measure_collapse": ([_vp, _u32, _i32, _f64], None),
    "orc_block_dump": ([_vp, _p, _u32, _p, _p, _p], _u64),
    "orc_block_logical_state": ([_vp, _p], None), "orc_block_physical_state": ([_vp, _p], None),
    "orc_block_map": ([_vp, _p], None), "orc_block_swap_count": ([_vp], _u64),
    "orc_block_clear": ([_vp], None),
}

_LIB = None


def load():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-s", "-C", HERE], check=True)
        lib = C.CDLL(LIB_PATH)
        for name, (args, res) in _SIG.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _LIB = lib
    return _LIB


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _u32(seq):
    return np.ascontiguousarray(np.asarray(list(seq), dtype=np.uint32))


def flatten_hamiltonian(pauli_strings, coefficients):
    paulis, qubits, offsets = [], [], [0]
    for s in pauli_strings:
        for p, q in s:
            paulis.append(PAULI[str(p).upper()])
            qubits.append(int(q))
        offsets.append(len(paulis))
    return (np.asarray(paulis, np.uint8), np.asarray(qubits, np.uint32),
            np.asarray(offsets, np.uint32), np.asarray(coefficients, np.float64))


def to_ops(ops):
    if isinstance(ops, np.ndarray):
        return np.ascontiguousarray(ops.view(OP_DTYPE) if ops.dtype != OP_DTYPE else ops)
    a = np.zeros(len(ops), dtype=OP_DTYPE)
    for i, op in enumerate(ops):
        kind, angle, target, control = op[0], op[1], op[2], op[3]
        param = op[4] if len(op) > 4 else -1
        a[i] = (kind, target, 0, param, angle, sum(1 << c for c in control))
    return a


class DenseV2:
    """Oracle DenseV2<f64> (dense_v2.rs)."""

    def __init__(self, n):
        self._l = load()
        self._s = self._l.orc_v2_new(n)
        if not self._s:
            raise ValueError("ExecutionFailed: more than 32 qubits")
        self.n = n

    def __del__(self):
        if getattr(self, "_s", None):
            self._l.orc_v2_free(self._s)
            self._s = None

    def _gate(self, kind, angle, target, control):
        c = _u32(control)
        mask = 0
        for q in c:
            mask |= 1 << int(q)
        self._l.orc_v2_gate_mask(self._s, kind, float(angle), target, mask)

    def pauli_x(self, t, control=()):
        c = _u32(control)
        self._l.orc_v2_pauli_x(self._s, t, _ptr(c), len(c))

    def pauli_y(self, t, control=()):
        c = _u32(control)
        self._l.orc_v2_pauli_y(self._s, t, _ptr(c), len(c))

    def pauli_z(self, t, control=()):
        c = _u32(control)
        self._l.orc_v2_pauli_z(self._s, t, _ptr(c), len(c))

    def hadamard(self, t, control=()):
        c = _u32(control)
        self._l.orc_v2_hadamard(self._s, t, _ptr(c), len(c))

    def phase(self, lam, t, control=()):
        c = _u32(control)
        self._l.orc_v2_phase(self._s, float(lam), t, _ptr(c), len(c))

    def rx(self, th, t, control=()):
        c = _u32(control)
        self._l.orc_v2_rx(self._s, float(th), t, _ptr(c), len(c))

    def ry(self, th, t, control=()):
        c = _u32(control)
        self._l.orc_v2_ry(self._s, float(th), t, _ptr(c), len(c))

    def rz(self, th, t, control=()):
        c = _u32(control)
        self._l.orc_v2_rz(self._s, float(th), t, _ptr(c), len(c))

    def gate(self, kind, angle, target, control=()):
        self._gate(kind, angle, target, control)

    def gate_mask(self, kind, angle, target, mask):
        self._l.orc_v2_gate_mask(self._s, kind, float(angle), target, mask)

    def run(self, ops, params=None):
        a = to_ops(ops)
        p = None if params is None else np.ascontiguousarray(params, np.float64)
        self._l.orc_v2_run(self._s, _ptr(a), len(a), _ptr(p))

    def measure_p1(self, t):
        return self._l.orc_v2_measure_p1(self._s, t)

    def measure_collapse(self, t, result, p):
        self._l.orc_v2_measure_collapse(self._s, t, 1 if result else 0, float(p))

    def measure(self, qubits, uniforms):
        q = _u32(qubits)
        u = np.ascontiguousarray(uniforms, np.float64)
        return self._l.orc_measure(self._s, _ptr(q), len(q), _ptr(u))

    def dump(self, qubits):
        q = _u32(qubits)
        m = self._l.orc_v2_dump(self._s, _ptr(q), len(q), None, None, None)
        keys = np.zeros(m, np.uint64)
        re = np.zeros(m)
        im = np.zeros(m)
        self._l.orc_v2_dump(self._s, _ptr(q), len(q), _ptr(keys), _ptr(re), _ptr(im))
        return keys, re, im

    def exp_value(self, pauli_strings, coefficients):
        pa, qu, off, co = flatten_hamiltonian(pauli_strings, coefficients)
        return self._l.orc_exp_value(self._s, _ptr(pa), _ptr(qu), _ptr(off), _ptr(co), len(co))

    def exp_value_chunked(self, pauli_strings, coefficients):
        """exp_value with each string's sum split into 2^20-amplitude chunks
        summed in parallel (same bits as exp_value for n <= 20)."""
        pa, qu, off, co = flatten_hamiltonian(pauli_strings, coefficients)
        return self._l.orc_exp_value_chunked(self._s, _ptr(pa), _ptr(qu), _ptr(off), _ptr(co), len(co))

    def sample_sequential(self, qubits, shots, uniforms):
        q = _u32(qubits)
        u = np.ascontiguousarray(uniforms, np.float64)
        out = np.zeros(shots, np.uint64)
        rc = self._l.orc_sample_sequential(self._s, _ptr(q), len(q), shots, _ptr(u), _ptr(out))
        if rc:
            raise ValueError("all weights zero")
        return out

    def sample_hier(self, qubits, shots, uniforms):
        return sample_hier(self.state(), self.n, qubits, shots, uniforms)

    def clear(self):
        self._l.orc_v2_clear(self._s)

    def state(self):
        n = 1 << self.n
        buf = (C.c_double * (2 * n)).from_address(self._l.orc_v2_data(self._s))
        return np.frombuffer(buf, dtype=np.complex128).copy()

    def state_view(self, first, count):
        """Amplitudes [first, first+count) as a view of the oracle's buffer
        (no copy; valid while this object lives and the state is unchanged)."""
        assert 0 <= first and first + count <= (1 << self.n)
        addr = self._l.orc_v2_data(self._s) + 16 * first
        buf = (C.c_double * (2 * count)).from_address(addr)
        return np.frombuffer(buf, dtype=np.complex128)

    def set_state(self, amps):
        n = 1 << self.n
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        assert a.shape == (n,)
        C.memmove(self._l.orc_v2_data(self._s), a.ctypes.data, a.nbytes)


class DenseV1:
    """Oracle Dense<f64> (dense.rs), the double-buffered twin."""

    def __init__(self, n):
        self._l = load()
        self._s = self._l.orc_v1_new(n)
        self.n = n

    def __del__(self):
        if getattr(self, "_s", None):
            self._l.orc_v1_free(self._s)
            self._s = None

    def gate(self, kind, angle, target, control=()):
        mask = 0
        for q in control:
            mask |= 1 << int(q)
        self._l.orc_v1_gate_mask(self._s, kind, float(angle), target, mask)

    def measure_p1(self, t):
        return self._l.orc_v1_measure_p1(self._s, t)

    def measure_collapse(self, t, result, p):
        self._l.orc_v1_measure_collapse(self._s, t, 1 if result else 0, float(p))

    def state(self):
        n = 1 << self.n
        buf = (C.c_double * (2 * n)).from_address(self._l.orc_v1_data(self._s))
        return np.frombuffer(buf, dtype=np.complex128).copy()


class Block:
    """Oracle Block<(), DenseV2<f64>> (block.rs)."""

    def __init__(self, n, block_size):
        self._l = load()
        self._s = self._l.orc_block_new(n, block_size)
        if not self._s:
            raise ValueError("ExecutionFailed")
        self.n = n

    def __del__(self):
        if getattr(self, "_s", None):
            self._l.orc_block_free(self._s)
            self._s = None

    def gate(self, kind, angle, target, control=()):
        c = _u32(control)
        self._l.orc_block_gate(self._s, kind, float(angle), target, _ptr(c), len(c))

    def measure_p1(self, t):
        return self._l.orc_block_measure_p1(self._s, t)

    def measure_collapse(self, t, result, p):
        self._l.orc_block_measure_collapse(self._s, t, 1 if result else 0, float(p))

    def dump(self, qubits):
        q = _u32(qubits)
        m = self._l.orc_block_dump(self._s, _ptr(q), len(q), None, None, None)
        keys = np.zeros(m, np.uint64)
        re = np.zeros(m)
        im = np.zeros(m)
        self._l.orc_block_dump(self._s, _ptr(q), len(q), _ptr(keys), _ptr(re), _ptr(im))
        return keys, re, im

    def logical_state(self):
        out = np.zeros(1 << self.n, np.complex128)
        self._l.orc_block_logical_state(self._s, _ptr(out))
        return out

    def physical_state(self):
        out = np.zeros(1 << self.n, np.complex128)
        self._l.orc_block_physical_state(self._s, _ptr(out))
        return out

    def qubit_map(self):
        out = np.zeros(self.n, np.uint32)
        self._l.orc_block_map(self._s, _ptr(out))
        return out

    def swap_count(self):
        return self._l.orc_block_swap_count(self._s)

    def clear(self):
        self._l.orc_block_clear(self._s)


def uniform_scale(total):
    return load().orc_uniform_scale(float(total))


def sample_hier(amps, n, qubits, shots, uniforms):
    """Hierarchical-order sampling over a physical-order amplitude array."""
    a = np.ascontiguousarray(amps, dtype=np.complex128)
    q = _u32(qubits)
    u = np.ascontiguousarray(uniforms, np.float64)
    out = np.zeros(shots, np.uint64)
    rc = load().orc_sample_hier(_ptr(a), n, _ptr(q), len(q), shots, _ptr(u), _ptr(out))
    if rc:
        raise ValueError("all weights zero")
    return out
