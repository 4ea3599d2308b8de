"""The C-ABI library loads and exports every entry point include/kbwh.h declares (CPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import kbw_amd
from kbw_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kbwh.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kbwh_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = kbw_amd.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for name in syms:
        assert hasattr(lib, name), name
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signatures cover exactly the header"


def test_op_layout_matches_header():
    assert _lib.OP_DTYPE.itemsize == 24
    assert [_lib.OP_DTYPE.fields[f][1] for f in ("kind", "target", "reserved", "param", "angle", "ctrl_mask")] == [0, 1, 2, 4, 8, 16]


def test_build_info():
    assert "kbwh" in kbw_amd.build_info()


def test_constants_match_header():
    text = open(os.path.join(ROOT, "include", "kbwh.h")).read()
    consts = dict(re.findall(r"#define (KBWH_[A-Z0-9_]+) (\d+)u?", text))
    assert int(consts["KBWH_RX"]) == _lib.RX and int(consts["KBWH_RZ"]) == _lib.RZ
    assert int(consts["KBWH_ERR_EXECUTION_FAILED"]) == _lib.ERR_EXECUTION_FAILED
    assert int(consts["KBWH_PAULI_Y"]) == _lib.PAULI_Y
    assert int(consts["KBWH_RUN_TILED"]) == _lib.RUN_TILED
    assert int(consts["KBWH_RUN_FUSED"]) == _lib.RUN_FUSED and int(consts["KBWH_RUN_AUTO"]) == _lib.RUN_AUTO
    assert int(consts["KBWH_RUN_REORDER"]) == _lib.RUN_REORDER


def test_tune_result_layout_matches_header():
    assert _lib.TUNE_DTYPE.itemsize == 80
    assert [_lib.TUNE_DTYPE.fields[f][1] for f in ("mode", "fusion_width", "tile_targets", "num_candidates", "ms")] == [0, 4, 8, 12, 16]


def test_invalid_arguments_rejected_without_device():
    lib = kbw_amd.load()
    assert lib.kbwh_new(4, 0, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert lib.kbwh_gate_mask(None, 0, 0.0, 0, 0) == _lib.ERR_INVALID_ARGUMENT
    assert lib.kbwh_set_run_config(None, 5, 8) == _lib.ERR_INVALID_ARGUMENT
    assert lib.kbwh_autotune(None, None, 0, None, 0, None) == _lib.ERR_INVALID_ARGUMENT
    assert b"null" in lib.kbwh_last_error()


def test_hamiltonian_flatten():
    h = kbw_amd.Hamiltonian([[("X", 0), ("Z", 3)], [("Y", 1)]], [0.5, -2.0])
    p, q, o, c = h.flatten()
    assert p.tolist() == [1, 3, 2] and q.tolist() == [0, 3, 1] and o.tolist() == [0, 2, 3]
    assert c.tolist() == [0.5, -2.0]
