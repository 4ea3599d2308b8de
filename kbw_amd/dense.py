"""Host-side mirror of the reference's QuantumExecution trait over libkbwh.so.

`DenseHip` has the method names, argument meanings and error behaviour of the
trait (src/quantum_execution.rs:48-162) as DenseV2 implements it
(src/dense_v2.rs:150-342); with ``block=True`` it follows Block<_, DenseV2>
(src/block.rs:211-381).  `QubitManager` mirrors the adapter of
quantum_execution.rs:171-427 (measure loop, sample, exp_value, batch replay).
Every call goes to the HIP library; nothing here computes amplitudes.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import KbwhError, check, ptr

PAULI_CODES = {"X": _lib.PAULI_X, "Y": _lib.PAULI_Y, "Z": _lib.PAULI_Z}


class Hamiltonian:
    """libket's Hamiltonian: ``pauli_strings[k]`` is a list of ``(pauli, qubit)``
    with pauli in ``"XYZ"``; ``coefficients[k]`` is its real weight."""

    def __init__(self, pauli_strings, coefficients):
        if len(pauli_strings) != len(coefficients):
            raise ValueError("one coefficient per Pauli string")
        self.pauli_strings = [[(str(p).upper(), int(q)) for p, q in s] for s in pauli_strings]
        self.coefficients = [float(c) for c in coefficients]

    def flatten(self):
        """(paulis u8, qubits u32, offsets u32 [len+1], coeffs f64) for the C ABI."""
        paulis, qubits, offsets = [], [], [0]
        for s in self.pauli_strings:
            for p, q in s:
                paulis.append(PAULI_CODES[p])
                qubits.append(q)
            offsets.append(len(paulis))
        return (np.asarray(paulis, dtype=np.uint8), np.asarray(qubits, dtype=np.uint32),
                np.asarray(offsets, dtype=np.uint32), np.asarray(self.coefficients, dtype=np.float64))


class DumpData:
    """libket's DumpData: one entry per amplitude with |a| > 1e-15."""

    def __init__(self, keys, re, im):
        self.keys = keys
        self.amplitudes_real = re
        self.amplitudes_imag = im

    @property
    def basis_states(self):
        return [[int(k)] for k in self.keys]

    def __len__(self):
        return len(self.keys)


def _u32(seq):
    return np.ascontiguousarray(np.asarray(list(seq), dtype=np.uint32))


def ops_array(ops):
    """Coerce a gate list to a contiguous kbwh_op array."""
    if isinstance(ops, np.ndarray) and ops.dtype == _lib.OP_DTYPE:
        return np.ascontiguousarray(ops)
    a = np.zeros(len(ops), dtype=_lib.OP_DTYPE)
    for i, op in enumerate(ops):
        kind, angle, target, control = op[0], op[1], op[2], op[3]
        param = op[4] if len(op) > 4 else -1
        a[i] = (kind, target, 0, param, angle, sum(1 << c for c in control))
    return a


def reorder_plan(ops, num_qubits, tile_targets=64):
    """The pass plan of RUN_REORDER (host only): (order, passes), where
    order lists the op indices in execution order and passes[p] is the
    (start, end) range of pass p in order."""
    a = ops_array(ops)
    order = np.zeros(len(a), dtype=np.uint32)
    ends = np.zeros(max(len(a), 1), dtype=np.uint32)
    npass = C.c_uint64()
    check(_lib.load().kbwh_reorder_plan(ptr(a), len(a), num_qubits, tile_targets, ptr(order), ptr(ends),
                                        C.byref(npass)))
    ends = ends[:npass.value].astype(np.int64)
    starts = np.concatenate([[0], ends[:-1]])
    return order.astype(np.int64), list(zip(starts.tolist(), ends.tolist()))


def compiled_stats():
    """Pass-specialised kernels of RUN_COMPILED in this process: a dict of
    kernels compiled, cache hits, failures, passes run by the generic kernel
    and seconds spent compiling."""
    v = [C.c_uint64() for _ in range(4)]
    sec = C.c_double()
    check(_lib.load().kbwh_compiled_stats(*[C.byref(x) for x in v], C.byref(sec)))
    return {"compiled": v[0].value, "cache_hits": v[1].value, "failures": v[2].value,
            "generic_passes": v[3].value, "compile_seconds": sec.value}


def set_order_tuning(enable=True):
    """Tile-order tuning of the compiled passes for the process
    (kbwh_set_order_tuning): each pass layout tries the candidate orders on
    its first runs and keeps the fastest.  The order does not change results."""
    check(_lib.load().kbwh_set_order_tuning(1 if enable else 0))


def order_tuning_stats():
    """A dict of pass layouts tuned, layouts still trying orders, candidate
    launches timed and chosen orders other than the default."""
    v = [C.c_uint64() for _ in range(4)]
    check(_lib.load().kbwh_order_tuning_stats(*[C.byref(x) for x in v]))
    return {"tuned": v[0].value, "pending": v[1].value, "trials": v[2].value, "non_default": v[3].value}


def compiled_pass_code(ops, num_qubits, pass_index, reverse=False):
    """The gfx950 code object (bytes) of the RUN_COMPILED kernel of one pass
    (reverse=True: of the adjoint's reverse sweep), compiled on the host."""
    a = ops_array(ops)
    size = C.c_uint64()
    lib = _lib.load()
    fn = lib.kbwh_compiled_reverse_pass_code if reverse else lib.kbwh_compiled_pass_code
    check(fn(ptr(a), len(a), num_qubits, pass_index, None, C.byref(size)))
    buf = C.create_string_buffer(size.value)
    check(fn(ptr(a), len(a), num_qubits, pass_index, buf, C.byref(size)))
    return buf.raw[:size.value]


class DenseHip:
    """QuantumExecution on one MI355X: the dense c128 state lives in HBM."""

    def __init__(self, num_qubits, block=False, block_size=0, queue=False):
        """queue=True: live gates are queued and applied before the next read
        of the state (KBWH_FLAG_QUEUE; one launch for up to 12 qubits)."""
        lib = _lib.load()
        h = C.c_void_p()
        flags = (_lib.FLAG_BLOCK if block else 0) | (_lib.FLAG_QUEUE if queue else 0)
        check(lib.kbwh_new(num_qubits, flags, block_size, C.byref(h)))
        self._lib = lib
        self._h = h
        self.num_qubits = num_qubits
        self.block = block

    @classmethod
    def new(cls, num_qubits):
        """DenseV2::new (dense_v2.rs:151-164): raises KbwhError(1) when n > 32."""
        return cls(num_qubits)

    # ---- lifetime ------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            self._lib.kbwh_free(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def clear(self):
        check(self._lib.kbwh_clear(self._h))

    # ---- gates (quantum_execution.rs:57-79) -----------------------------
    def gate(self, kind, angle, target, control=()):
        c = _u32(control)
        check(self._lib.kbwh_gate(self._h, int(kind), float(angle), int(target), ptr(c), len(c)))

    def pauli_x(self, target, control=()):
        self.gate(_lib.X, 0.0, target, control)

    def pauli_y(self, target, control=()):
        self.gate(_lib.Y, 0.0, target, control)

    def pauli_z(self, target, control=()):
        self.gate(_lib.Z, 0.0, target, control)

    def hadamard(self, target, control=()):
        self.gate(_lib.H, 0.0, target, control)

    def phase(self, lam, target, control=()):
        self.gate(_lib.P, lam, target, control)

    def rx(self, theta, target, control=()):
        self.gate(_lib.RX, theta, target, control)

    def ry(self, theta, target, control=()):
        self.gate(_lib.RY, theta, target, control)

    def rz(self, theta, target, control=()):
        self.gate(_lib.RZ, theta, target, control)

    def run(self, ops, params=None, mode=_lib.RUN_PER_GATE):
        a = ops_array(ops)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        check(self._lib.kbwh_run(self._h, ptr(a), len(a), ptr(p), mode))

    def set_compiled(self, enable=True):
        """Adjoint gradients on kernels compiled per pass, and exp_value's
        grouped sweeps on kernels compiled per string group (kbwh_set_compiled)."""
        check(self._lib.kbwh_set_compiled(self._h, 1 if enable else 0))

    def set_plan_search(self, trials):
        """Randomised plans tried by dependency-ordered runs, once per circuit
        structure (kbwh_set_plan_search; < 0: the default, 0 for RUN_REORDER
        and 256 for RUN_COMPILED)."""
        check(self._lib.kbwh_set_plan_search(self._h, int(trials)))

    def set_run_config(self, fusion_width=5, tile_targets=64):
        """Qubits per fused block (4 or 5) and most non-diagonal targets per
        tiled pass (1..64; 64 = as many as the tile has free positions)."""
        check(self._lib.kbwh_set_run_config(self._h, fusion_width, tile_targets))

    def autotune(self, ops, params=None, allow_inexact=False):
        """Time the run configurations on a scratch buffer and keep the fastest
        for RUN_AUTO; returns (mode, fusion_width, tile_targets, ms per candidate)."""
        a = ops_array(ops)
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float64)
        out = np.zeros(1, dtype=_lib.TUNE_DTYPE)
        check(self._lib.kbwh_autotune(self._h, ptr(a), len(a), ptr(p), 1 if allow_inexact else 0, ptr(out)))
        r = out[0]
        return int(r["mode"]), int(r["fusion_width"]), int(r["tile_targets"]), r["ms"][: int(r["num_candidates"])].copy()

    # ---- measurement ----------------------------------------------------
    def measure_p1(self, target):
        out = C.c_double()
        check(self._lib.kbwh_measure_p1(self._h, target, C.byref(out)))
        return out.value

    def measure_collapse(self, target, result, p):
        check(self._lib.kbwh_measure_collapse(self._h, target, 1 if result else 0, float(p)))

    def measure(self, qubits, uniforms):
        """QubitManager::do_measure (quantum_execution.rs:272-292) with host uniforms."""
        q = _u32(qubits)
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        if len(u) < len(q):
            raise ValueError("one uniform per measured qubit")
        out = C.c_uint64()
        check(self._lib.kbwh_measure(self._h, ptr(q), len(q), ptr(u), C.byref(out)))
        return out.value

    # ---- dump / sample / exp_value -------------------------------------
    def dump(self, qubits):
        q = _u32(qubits)
        cnt = C.c_uint64()
        check(self._lib.kbwh_dump(self._h, ptr(q), len(q), 0, C.byref(cnt), None, None, None))
        m = cnt.value
        keys = np.zeros(m, dtype=np.uint64)
        re = np.zeros(m, dtype=np.float64)
        im = np.zeros(m, dtype=np.float64)
        if m:
            check(self._lib.kbwh_dump(self._h, ptr(q), len(q), m, C.byref(cnt), ptr(keys), ptr(re),
                                      ptr(im)))
        return DumpData(keys, re, im)

    def sample(self, qubits, shots, uniforms):
        """One dump key per shot (the reference returns their histogram)."""
        q = _u32(qubits)
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        if len(u) < shots:
            raise ValueError("one uniform per shot")
        out = np.zeros(shots, dtype=np.uint64)
        check(self._lib.kbwh_sample(self._h, ptr(q), len(q), shots, ptr(u), ptr(out)))
        return out

    def exp_value(self, hamiltonian):
        paulis, qubits, offsets, coeffs = hamiltonian.flatten()
        out = C.c_double()
        check(self._lib.kbwh_exp_value(self._h, ptr(paulis), ptr(qubits), ptr(offsets), ptr(coeffs),
                                       len(coeffs), C.byref(out)))
        return out.value

    def adjoint_gradient(self, ops, params, hamiltonian):
        """Energy and d<H>/dparams by adjoint differentiation (two extra vectors)."""
        a = ops_array(ops)
        p = np.ascontiguousarray(params, dtype=np.float64)
        paulis, qubits, offsets, coeffs = hamiltonian.flatten()
        grad = np.zeros(len(p), dtype=np.float64)
        e = C.c_double()
        check(self._lib.kbwh_adjoint_gradient(self._h, ptr(a), len(a), ptr(p), len(p), ptr(paulis),
                                              ptr(qubits), ptr(offsets), ptr(coeffs), len(coeffs),
                                              C.byref(e), ptr(grad)))
        return e.value, grad

    def adjoint_batch(self, ops, param_sets, hamiltonian):
        """Energies and gradients for many parameter sets of one circuit
        (rows of param_sets) in one sweep over all sets."""
        a = ops_array(ops)
        p = np.ascontiguousarray(np.atleast_2d(param_sets), dtype=np.float64)
        nsets, nparams = p.shape
        paulis, qubits, offsets, coeffs = hamiltonian.flatten()
        energies = np.zeros(nsets, dtype=np.float64)
        grads = np.zeros((nsets, nparams), dtype=np.float64)
        check(self._lib.kbwh_adjoint_batch(self._h, ptr(a), len(a), ptr(p), nparams, nsets, ptr(paulis),
                                           ptr(qubits), ptr(offsets), ptr(coeffs), len(coeffs), ptr(energies),
                                           ptr(grads)))
        return energies, grads

    def adjoint_stats(self):
        """(forward passes, lambda sweeps, reverse passes) of the last adjoint gradient."""
        f, l, r = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check(self._lib.kbwh_adjoint_stats(self._h, C.byref(f), C.byref(l), C.byref(r)))
        return f.value, l.value, r.value

    # ---- raw access ---------------------------------------------------
    def state(self, first=0, count=None):
        """Physical-order amplitudes as complex128 (a host copy)."""
        n = 1 << self.num_qubits
        count = n - first if count is None else count
        out = np.zeros(count, dtype=np.complex128)
        check(self._lib.kbwh_read_amplitudes(self._h, first, count, ptr(out)))
        return out

    def set_state(self, amps, first=0):
        a = np.ascontiguousarray(amps, dtype=np.complex128)
        check(self._lib.kbwh_write_amplitudes(self._h, first, len(a), ptr(a)))

    def qubit_map(self):
        out = np.zeros(self.num_qubits, dtype=np.uint32)
        check(self._lib.kbwh_qubit_map(self._h, ptr(out)))
        return out

    def logical_state(self):
        """Amplitudes in logical qubit order (undoes the Block permutation)."""
        phys = self.state()
        l2p = self.qubit_map()
        idx = np.arange(1 << self.num_qubits, dtype=np.uint64)
        src = np.zeros_like(idx)
        for q, p in enumerate(l2p):
            src |= ((idx >> np.uint64(q)) & np.uint64(1)) << np.uint64(p)
        return phys[src]

    def swap_count(self):
        """Block swaps made since the handle was created (clear() keeps the count)."""
        out = C.c_uint64()
        check(self._lib.kbwh_swap_count(self._h, C.byref(out)))
        return out.value

    def run_stats(self):
        """(gates, HBM passes) of the last run()."""
        g, p = C.c_uint64(), C.c_uint64()
        check(self._lib.kbwh_run_stats(self._h, C.byref(g), C.byref(p)))
        return g.value, p.value

    def synchronize(self):
        check(self._lib.kbwh_synchronize(self._h))

    def stream(self):
        out = C.c_void_p()
        check(self._lib.kbwh_stream(self._h, C.byref(out)))
        return out.value

    def device_buffer(self):
        p, b = C.c_void_p(), C.c_uint64()
        check(self._lib.kbwh_device_buffer(self._h, C.byref(p), C.byref(b)))
        return p.value, b.value


def kbw_seed():
    """KBW_SEED as quantum_execution.rs:183-186 reads it, or None."""
    try:
        return int(os.environ.get("KBW_SEED", ""))
    except ValueError:
        return None


class QubitManager:
    """The adapter of quantum_execution.rs:171-427 over DenseHip.

    Uniforms come from a numpy PCG64 stream seeded from KBW_SEED (the
    reference uses rand's StdRng; the engine only consumes the uniforms).
    """

    def __init__(self, num_qubits, block=False, block_size=0, seed=None, queue=False):
        self.num_qubits = num_qubits
        self.block = block
        self.block_size = block_size
        self.simulator = DenseHip(num_qubits, block=block, block_size=block_size, queue=queue)
        s = kbw_seed() if seed is None else seed
        self.rng = np.random.default_rng(s)

    def _uniforms(self, n):
        return self.rng.random(n)

    def compute_gate(self, kind, angle, target, control=()):
        self.simulator.gate(kind, angle, target, control)

    def compute_native_gates(self, gates):
        """compute_gates (quantum_execution.rs:342-364): clear, then named gates."""
        self.simulator.clear()
        for name, angles, qubits in gates:
            control, target = list(qubits[:-1]), qubits[-1]
            if name in ("x", "y", "z", "h"):
                self.simulator.gate(_lib.GATE_NAMES[name], 0.0, target)
            elif name in ("rx", "ry", "rz", "p"):
                self.simulator.gate(_lib.GATE_NAMES[name], angles[0], target)
            elif name == "cnot":
                self.simulator.pauli_x(target, control)
            elif name == "cz":
                self.simulator.pauli_z(target, control)
            else:
                raise KbwhError(_lib.ERR_INVALID_ARGUMENT, f"native gate {name!r} unsupported")

    def measure(self, qubits, uniforms=None):
        u = self._uniforms(len(qubits)) if uniforms is None else uniforms
        return self.simulator.measure(qubits, u)

    def sample(self, qubits, shots, uniforms=None):
        """(basis_states, counts) histogram, as SampleData."""
        if not len(qubits):
            raise KbwhError(_lib.ERR_INVALID_ARGUMENT, "sample needs at least one qubit")
        u = self._uniforms(shots) if uniforms is None else uniforms
        keys = self.simulator.sample(qubits, shots, u)
        states, counts = np.unique(keys, return_counts=True)
        return [[int(s)] for s in states], [int(c) for c in counts]

    def dump(self, qubits):
        return self.simulator.dump(qubits)

    def exp_value(self, hamiltonian):
        return self.simulator.exp_value(hamiltonian)

    def batch_exp_value(self, ops, hamiltonians, params=None):
        """BatchExecution::exp_value (quantum_execution.rs:392-405): a fresh state, replay, evaluate."""
        sim = DenseHip(self.num_qubits, block=self.block, block_size=self.block_size)
        try:
            sim.run(ops, params)
            return [sim.exp_value(h) for h in hamiltonians]
        finally:
            sim.close()

    def batch_sample(self, ops, qubits, shots, uniforms=None, params=None):
        """BatchExecution::sample (quantum_execution.rs:368-379)."""
        sim = DenseHip(self.num_qubits, block=self.block, block_size=self.block_size)
        try:
            sim.run(ops, params)
            u = self._uniforms(shots) if uniforms is None else uniforms
            return sim.sample(qubits, shots, u)
        finally:
            sim.close()
