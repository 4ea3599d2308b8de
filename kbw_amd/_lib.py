"""ctypes binding of libkbwh.so (include/kbwh.h).

The library is the product: there is no Python or CPU fallback.  Loading
fails loudly when kbw_amd/libkbwh.so has not been built.
"""
import ctypes as C
import os

import numpy as np

# KBWH_LIB: another build of the library (the sanitized host-only build of
# the CPU tests); default the in-tree build
LIB_PATH = os.environ.get("KBWH_LIB") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "libkbwh.so")

OK = 0
ERR_EXECUTION_FAILED = 1
ERR_INVALID_ARGUMENT = 2
ERR_OUT_OF_MEMORY = 3
ERR_DEVICE = 4
ERR_NO_DEVICE = 5
ERR_ZERO_WEIGHTS = 6
ERR_CAPACITY = 7

X, Y, Z, H, P, RX, RY, RZ = range(8)
GATE_NAMES = {"x": X, "y": Y, "z": Z, "h": H, "p": P, "rx": RX, "ry": RY, "rz": RZ}
PAULI_X, PAULI_Y, PAULI_Z = 1, 2, 3
FLAG_BLOCK, FLAG_QUEUE = 1, 2
RUN_PER_GATE, RUN_TILED, RUN_FUSED, RUN_AUTO, RUN_REORDER, RUN_COMPILED = 0, 1, 2, 3, 4, 5

# kbwh_op, 24 bytes
OP_DTYPE = np.dtype([("kind", "u1"), ("target", "u1"), ("reserved", "<u2"), ("param", "<i4"),
                     ("angle", "<f8"), ("ctrl_mask", "<u8")], align=True)
assert OP_DTYPE.itemsize == 24

# kbwh_circuit, 32 bytes
CIRCUIT_DTYPE = np.dtype([("num_qubits", "<u4"), ("num_strings", "<u4"), ("first_op", "<u8"),
                          ("num_ops", "<u8"), ("first_string", "<u4"), ("first_param", "<u4")], align=True)
assert CIRCUIT_DTYPE.itemsize == 32

# kbwh_tune_result, 80 bytes
GROUP_DTYPE = np.dtype([("num_qubits", "<u4"), ("num_sets", "<u4"), ("tile_targets", "<u4"), ("compiled", "<u4"),
                        ("passes", "<u4"), ("sweeps", "<u4"), ("tuned_ms", "<f8")])
assert GROUP_DTYPE.itemsize == 32
TUNE_DTYPE = np.dtype([("mode", "<u4"), ("fusion_width", "<u4"), ("tile_targets", "<u4"),
                       ("num_candidates", "<u4"), ("ms", "<f8", (8,))], align=True)
assert TUNE_DTYPE.itemsize == 80

# Every exported symbol: (name, argtypes).  All return int32 except kbwh_last_error.
_vp, _u32, _u64, _i32, _f64 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_double
_p = C.c_void_p
SIGNATURES = {
    "kbwh_new": [_u32, _u32, _u32, C.POINTER(_vp)],
    "kbwh_free": [_vp],
    "kbwh_clear": [_vp],
    "kbwh_num_qubits": [_vp, C.POINTER(_u32)],
    "kbwh_gate": [_vp, _i32, _f64, _u32, _p, _u32],
    "kbwh_gate_mask": [_vp, _i32, _f64, _u32, _u64],
    "kbwh_run": [_vp, _p, _u64, _p, _u32],
    "kbwh_measure_p1": [_vp, _u32, C.POINTER(_f64)],
    "kbwh_measure_collapse": [_vp, _u32, _i32, _f64],
    "kbwh_measure": [_vp, _p, _u32, _p, C.POINTER(_u64)],
    "kbwh_dump": [_vp, _p, _u32, _u64, C.POINTER(_u64), _p, _p, _p],
    "kbwh_sample": [_vp, _p, _u32, _u64, _p, _p],
    "kbwh_exp_value": [_vp, _p, _p, _p, _p, _u32, C.POINTER(_f64)],
    "kbwh_exp_value_noalloc": [_vp, _p, _p, _p, _p, _u32, C.POINTER(_f64)],
    "kbwh_adjoint_gradient": [_vp, _p, _u64, _p, _u32, _p, _p, _p, _p, _u32, C.POINTER(_f64), _p],
    "kbwh_adjoint_batch": [_vp, _p, _u64, _p, _u32, _u32, _p, _p, _p, _p, _u32, _p, _p],
    "kbwh_adjoint_stats": [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)],
    "kbwh_read_amplitudes": [_vp, _u64, _u64, _p],
    "kbwh_write_amplitudes": [_vp, _u64, _u64, _p],
    "kbwh_qubit_map": [_vp, _p],
    "kbwh_swap_count": [_vp, C.POINTER(_u64)],
    "kbwh_synchronize": [_vp],
    "kbwh_run_stats": [_vp, C.POINTER(_u64), C.POINTER(_u64)],
    "kbwh_set_run_config": [_vp, _u32, _u32],
    "kbwh_autotune": [_vp, _p, _u64, _p, _i32, _p],
    "kbwh_stream": [_vp, C.POINTER(_vp)],
    "kbwh_device_buffer": [_vp, C.POINTER(_vp), C.POINTER(_u64)],
    "kbwh_batch_new": [_p, _u32, _p, _p, _p, _p, _p, _u64, C.POINTER(_vp)],
    "kbwh_batch_plan": [_p, _u32, _p, _p, _p, _p, _p, _u64, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)],
    "kbwh_reorder_plan": [_p, _u64, _u32, _u32, _p, _p, C.POINTER(_u64)],
    "kbwh_set_compiled": [_vp, _i32],
    "kbwh_set_plan_search": [_vp, _i32],
    "kbwh_compiled_stats": [C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_f64)],
    "kbwh_set_order_tuning": [_i32],
    "kbwh_order_tuning_stats": [C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)],
    "kbwh_compiled_pass_code": [_p, _u64, _u32, _u32, _p, C.POINTER(_u64)],
    "kbwh_compiled_reverse_pass_code": [_p, _u64, _u32, _u32, _p, C.POINTER(_u64)],
    "kbwh_batch_run": [_vp, _p, _p],
    "kbwh_batch_use_graph": [_vp, _i32],
    "kbwh_batch_state": [_vp, _u32, _u64, _p],
    "kbwh_batch_stats": [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)],
    "kbwh_batch_groups": [_vp, _p, _u32, C.POINTER(_u32)],
    "kbwh_inject_alloc_failure": [C.c_int64],
    "kbwh_batch_configure": [_vp, _u32, _i32],
    "kbwh_batch_autotune": [_vp, _p, _p, _u32],
    "kbwh_batch_free": [_vp],
    "kbwh_build_info": [C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t)],
    "kbwh_last_error": [],
}


class KbwhError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, code, message):
        super().__init__(f"kbwh error {code}: {message}")
        self.code = code


_LIB = None


def load():
    """Load libkbwh.so once and attach the signatures."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_char_p if name == "kbwh_last_error" else C.c_int32
    _LIB = lib
    return lib


def check(rc):
    if rc != OK:
        msg = load().kbwh_last_error()
        raise KbwhError(rc, msg.decode() if msg else "")
    return rc


def ptr(a):
    """Data pointer of a contiguous numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)


def build_info():
    lib = load()
    msg = C.POINTER(C.c_uint8)()
    size = C.c_size_t()
    check(lib.kbwh_build_info(C.byref(msg), C.byref(size)))
    return C.string_at(msg, size.value).decode()
