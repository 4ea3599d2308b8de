"""Code generation guards on the built library (CPU; reads the gfx950 code
objects inside kbw_amd/libkbwh.so).

The register-tile kernels keep a thread's amplitudes in registers.  When the
optimizer indexes that array through a select of addresses it goes to scratch
memory and a pass runs about 10x slower while every result stays correct, so
parity tests cannot catch it.  These tests read each kernel's metadata:
no kernel may use scratch memory, and the tile kernels must stay within the
register budget that keeps their occupancy (5 waves per SIMD for the pass
kernels, 3 for the adjoint sweep)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "kbw_amd", "libkbwh.so")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"
MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"


def _tools():
    need = [shutil.which("objcopy"), os.path.join(LLVM, "clang-offload-bundler"), os.path.join(LLVM, "llvm-readelf")]
    return all(t and os.path.exists(t) for t in need)


@pytest.fixture(scope="module")
def kernels(tmp_path_factory):
    if not os.path.exists(LIB):
        pytest.skip("kbw_amd/libkbwh.so is not built")
    if not _tools():
        pytest.skip("objcopy / clang-offload-bundler / llvm-readelf not available")
    d = tmp_path_factory.mktemp("codeobj")
    fat = d / "fatbin.bin"
    subprocess.run(["objcopy", "-O", "binary", "--only-section=.hip_fatbin", LIB, str(fat)], check=True)
    data = fat.read_bytes()
    starts = [m.start() for m in re.finditer(re.escape(MAGIC), data)]
    assert starts, "no offload bundle in .hip_fatbin"
    meta = {}
    for i, s in enumerate(starts):
        e = starts[i + 1] if i + 1 < len(starts) else len(data)
        part, co = d / f"b{i}.bin", d / f"b{i}.co"
        part.write_bytes(data[s:e])
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", "--type=o", f"--targets={TARGET}",
                        f"--input={part}", f"--output={co}"], check=True, capture_output=True)
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(co)], check=True,
                               capture_output=True, text=True).stdout
        name = None
        for line in notes.splitlines():
            m = re.search(r"^\s+\.name:\s+(\S+)", line)
            if m:
                name = m.group(1)
                meta.setdefault(name, {})
                continue
            m = re.search(r"^\s+\.(private_segment_fixed_size|vgpr_count):\s+(\d+)", line)
            if m and name:
                meta[name][m.group(1)] = int(m.group(2))
    return meta


def test_every_kernel_has_metadata(kernels):
    tiles = [k for k in kernels if "k_tile" in k]
    assert len(tiles) >= 6, sorted(kernels)  # k_tile<PERM, PF> x 4, k_tile_fma x 2
    for k in ("k_batch", "k_whole", "k_adj_tile", "k_pauli_tile"):
        assert any(k in n for n in kernels), k


def test_no_kernel_uses_scratch(kernels):
    bad = {k: v.get("private_segment_fixed_size") for k, v in kernels.items()
           if v.get("private_segment_fixed_size", 0) != 0}
    assert not bad, bad


@pytest.mark.parametrize("pattern, limit", [("k_tile", 96), ("k_tile_fma", 96), ("k_adj_tile", 168)])
def test_tile_kernels_keep_their_register_budget(kernels, pattern, limit):
    found = {k: v["vgpr_count"] for k, v in kernels.items() if re.search(pattern + r"[IE]", k)}
    assert found, pattern
    over = {k: n for k, n in found.items() if n > limit}
    assert not over, over


def _notes(path):
    notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", str(path)], check=True,
                           capture_output=True, text=True).stdout
    out = {}
    for key in ("private_segment_fixed_size", "vgpr_count"):
        m = re.search(r"^\s+\." + key + r":\s+(\d+)", notes, re.M)
        out[key] = int(m.group(1)) if m else None
    return out


@pytest.mark.parametrize("pass_index", [0, 7, 15])
def test_compiled_pass_kernels_keep_registers(tmp_path, pass_index):
    """Kernels compiled at run time for single passes of the 30-qubit brickwork
    (12-qubit tiles, 16 planned passes): no scratch memory, and within the
    register budget of 2 waves per SIMD (256 VGPRs; a kernel that spills at 3
    waves is built again for 2)."""
    if not os.path.exists(LIB):
        pytest.skip("kbw_amd/libkbwh.so is not built")
    if not _tools():
        pytest.skip("llvm-readelf not available")
    from kbw_amd import circuits, compiled_pass_code

    ops, _ = circuits.brickwork(30, 20, seed=1)
    co = tmp_path / "pass.co"
    co.write_bytes(compiled_pass_code(ops, 30, pass_index))
    meta = _notes(co)
    assert meta["private_segment_fixed_size"] == 0, meta
    assert meta["vgpr_count"] is not None and meta["vgpr_count"] <= 256, meta


def test_compiled_narrow_pass_kernels_keep_registers(tmp_path):
    """On 11 qubits (one 11-qubit tile) the compiled passes keep the 11-qubit
    kernels' budget: no scratch memory, at most 96 VGPRs."""
    if not os.path.exists(LIB):
        pytest.skip("kbw_amd/libkbwh.so is not built")
    if not _tools():
        pytest.skip("llvm-readelf not available")
    from kbw_amd import circuits, compiled_pass_code

    ops, _ = circuits.brickwork(11, 6, seed=1)
    co = tmp_path / "pass.co"
    co.write_bytes(compiled_pass_code(ops, 11, 0))
    meta = _notes(co)
    assert meta["private_segment_fixed_size"] == 0, meta
    assert meta["vgpr_count"] is not None and meta["vgpr_count"] <= 96, meta


@pytest.mark.parametrize("pass_index", [0, 3, 12, 24])
def test_compiled_reverse_pass_kernels_use_no_scratch(tmp_path, pass_index):
    """Reverse-sweep kernels compiled for passes of the 24-qubit, 576-parameter
    ansatz gradient: no scratch memory (a kernel that spills is not used at run
    time, and its pass falls back to the generic kernel)."""
    if not os.path.exists(LIB):
        pytest.skip("kbw_amd/libkbwh.so is not built")
    if not _tools():
        pytest.skip("llvm-readelf not available")
    from kbw_amd import circuits, compiled_pass_code

    ops, _ = circuits.hea(24, 12, seed=7)
    co = tmp_path / "rev.co"
    co.write_bytes(compiled_pass_code(ops, 24, pass_index, reverse=True))
    meta = _notes(co)
    assert meta["private_segment_fixed_size"] == 0, meta
