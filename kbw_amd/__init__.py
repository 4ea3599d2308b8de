"""MI355X-native dense state-vector engine for KBW.

The compute path is libkbwh.so (hand-written HIP for gfx950 behind the C ABI
in include/kbwh.h); this package is its host-side mirror of the reference's
QuantumExecution interface.
"""
from ._lib import (H, OP_DTYPE, P, PAULI_X, PAULI_Y, PAULI_Z, RUN_AUTO, RUN_COMPILED, RUN_FUSED,
                   RUN_PER_GATE, RUN_REORDER, RUN_TILED, RX, RY, RZ, X, Y, Z, KbwhError, build_info, load)
from .dense import (DenseHip, DumpData, Hamiltonian, QubitManager, compiled_pass_code, compiled_stats, ops_array,
                    order_tuning_stats, set_order_tuning)

__all__ = [
    "DenseHip", "DumpData", "Hamiltonian", "QubitManager", "KbwhError", "OP_DTYPE", "ops_array",
    "build_info", "load", "X", "Y", "Z", "H", "P", "RX", "RY", "RZ", "PAULI_X", "PAULI_Y",
    "PAULI_Z", "RUN_PER_GATE", "RUN_TILED", "RUN_FUSED", "RUN_AUTO", "RUN_REORDER", "RUN_COMPILED",
    "compiled_stats", "compiled_pass_code", "order_tuning_stats", "set_order_tuning",
]
