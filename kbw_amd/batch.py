"""Batches of circuits in one persistent launch (kbwh_batch_*).

`Batch(circuits)` takes a list of (num_qubits, ops, hamiltonian, num_params)
and plans every circuit once; `run(params)` returns <H_c> for each circuit,
starting each from |0...0>.  `use_graph(True)` captures a run in a hipGraph
and replays it for later parameter sets (parameter sweeps).
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import check, ptr
from .dense import ops_array


class Batch:
    def __init__(self, circuits, plan_only=False):
        lib = _lib.load()
        recs = np.zeros(len(circuits), dtype=_lib.CIRCUIT_DTYPE)
        all_ops, paulis, qubits, offsets, coeffs = [], [], [], [0], []
        nops = nparams = nstr = 0
        for i, (n, ops, ham, np_c) in enumerate(circuits):
            a = ops_array(ops)
            p, q, off, co = ham.flatten()
            recs[i] = (n, len(co), nops, len(a), nstr, nparams)
            all_ops.append(a)
            paulis.append(p)
            qubits.append(q)
            offsets.extend((off[1:] + offsets[-1]).tolist())
            coeffs.append(co)
            nops += len(a)
            nparams += np_c
            nstr += len(co)
        self._ops = np.ascontiguousarray(np.concatenate(all_ops)) if all_ops else np.zeros(0, _lib.OP_DTYPE)
        self._paulis = np.ascontiguousarray(np.concatenate(paulis)) if paulis else np.zeros(0, np.uint8)
        self._qubits = np.ascontiguousarray(np.concatenate(qubits)) if qubits else np.zeros(0, np.uint32)
        self._offsets = np.ascontiguousarray(np.asarray(offsets, np.uint32))
        self._coeffs = np.ascontiguousarray(np.concatenate(coeffs)) if coeffs else np.zeros(0)
        self._recs = recs
        self.num_circuits = len(circuits)
        self.num_params = nparams
        self.param_offsets = np.concatenate([[0], np.cumsum([c[3] for c in circuits])]).astype(np.int64)
        self._lib = lib
        if plan_only:
            g, p, a = C.c_uint64(), C.c_uint64(), C.c_uint64()
            check(lib.kbwh_batch_plan(ptr(recs), len(recs), ptr(self._ops), ptr(self._paulis),
                                      ptr(self._qubits), ptr(self._offsets), ptr(self._coeffs), nparams,
                                      C.byref(g), C.byref(p), C.byref(a)))
            self.plan = (g.value, p.value, a.value)
            self._h = None
            return
        h = C.c_void_p()
        check(lib.kbwh_batch_new(ptr(recs), len(recs), ptr(self._ops), ptr(self._paulis), ptr(self._qubits),
                                 ptr(self._offsets), ptr(self._coeffs), nparams, C.byref(h)))
        self._lib, self._h = lib, h

    def run(self, params):
        p = np.ascontiguousarray(params, dtype=np.float64)
        if len(p) != self.num_params:
            raise ValueError("one value per batch parameter")
        out = np.zeros(self.num_circuits)
        check(self._lib.kbwh_batch_run(self._h, ptr(p), ptr(out)))
        return out

    def use_graph(self, enable=True):
        check(self._lib.kbwh_batch_use_graph(self._h, 1 if enable else 0))

    def state(self, circuit, count):
        out = np.zeros(count, dtype=np.complex128)
        check(self._lib.kbwh_batch_state(self._h, circuit, count, ptr(out)))
        return out

    def stats(self):
        g, p, a = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(self._lib.kbwh_batch_stats(self._h, C.byref(g), C.byref(p), C.byref(a)))
        return g.value, p.value, a.value

    def groups(self):
        """The structure groups run as sets: a list of dicts (num_qubits,
        num_sets, tile_targets, compiled passes, passes, sweeps, tuned_ms)."""
        n = C.c_uint32()
        check(self._lib.kbwh_batch_groups(self._h, None, 0, C.byref(n)))
        info = np.zeros(n.value, dtype=_lib.GROUP_DTYPE)
        check(self._lib.kbwh_batch_groups(self._h, ptr(info), n.value, C.byref(n)))
        return [{k: info[k][i].item() for k in info.dtype.names} for i in range(len(info))]

    def configure(self, tile_targets=8, compiled=False):
        check(self._lib.kbwh_batch_configure(self._h, tile_targets, 1 if compiled else 0))

    def autotune(self, params):
        """Times every group's candidates (generic kernel at 8, 6 and 4
        targets per pass, compiled kernels at the best) and keeps the fastest;
        returns the candidates' times, one row of 4 per group (ms)."""
        p = np.ascontiguousarray(params, dtype=np.float64)
        g = len(self.groups())
        ms = np.zeros(max(4 * g, 1))
        check(self._lib.kbwh_batch_autotune(self._h, ptr(p), ptr(ms), 4 * g))
        return ms[:4 * g].reshape(g, 4)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.kbwh_batch_free(self._h)
            self._h = None

    def __del__(self):
        self.close()
