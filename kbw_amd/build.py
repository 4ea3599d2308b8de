"""Builds libkbwh.so (HIP kernels for gfx950 + the C ABI) in-tree.

The library is compiled with hipcc for gfx950 only.  FP contraction is off in
every translation unit: the plain gate path reproduces the reference's
arithmetic operand for operand, and a fused multiply-add would change the
rounding.  Packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libkbwh.so")
BUILD_DIR = os.path.join(ROOT, "_build", "kbwh")

HIP_SOURCES = ["kernels.hip", "state.hip", "observables.hip", "tiled.hip", "adjoint.hip", "batch.hip", "fused.hip", "autotune.hip", "live.hip", "adjoint_tiled.hip", "tiled_fma.hip", "tiled_pass.hip", "tile_jit.hip", "tiled_wide.hip"]
CXX_SOURCES = ["coef.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", "-Wno-unused-parameter", "-Wno-unused-command-line-argument",
]
# Units holding register-tile kernels (tile_core.h) keep the optimizer from
# sinking code shared by the per-gate variants: merged variants index the
# register tile through a select of addresses, which puts it in scratch memory.
# The tiled pass kernels also keep their uniform per-gate branches
# unstructurized: each gate path then updates the register tile in place,
# instead of writing a second register set that is copied back after every
# gate.
_NO_SINK = ["-mllvm", "-simplifycfg-sink-common=false"]
_UNIFORM_BRANCHES = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]
HIP_SOURCE_FLAGS = {
    "tiled_pass.hip": _NO_SINK + _UNIFORM_BRANCHES,
    "tiled_fma.hip": _NO_SINK + _UNIFORM_BRANCHES,
    "batch.hip": _NO_SINK + _UNIFORM_BRANCHES,
    "live.hip": _NO_SINK + _UNIFORM_BRANCHES,
    "adjoint_tiled.hip": _NO_SINK + _UNIFORM_BRANCHES,
    "fused.hip": _NO_SINK,
    "tiled.hip": _NO_SINK + _UNIFORM_BRANCHES,
}
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return out


def build(jobs=8, verbose=False):
    """Compile every source and link kbw_amd/libkbwh.so (always rebuilds)."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    procs, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        procs.append([HIPCC, *HIP_FLAGS, *HIP_SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(BUILD_DIR, src.replace(".cpp", ".o"))
        objs.append(obj)
        procs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    # the device tile header as a string, for the kernels compiled at run time
    for hdr, sym in (("tile_dev.h", "kTileDevSource"), ("adj_dev.h", "kAdjDevSource")):
        emb = embed_header(os.path.join(CSRC, hdr), sym, os.path.join(BUILD_DIR, hdr.replace(".h", "_src.cpp")))
        objs.append(emb[:-4] + ".o")
        procs.append(["g++", *CXX_FLAGS, "-c", emb, "-o", objs[-1]])
    running = []
    for cmd in procs:
        running.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(running) >= jobs:
            _wait(running.pop(0), verbose)
    while running:
        _wait(running.pop(0), verbose)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-ldl"])
    os.replace(tmp, LIB)
    return LIB


def embed_header(path, symbol, out):
    """Writes a C++ source defining kbwh::<symbol> as the text of `path`."""
    text = open(path).read()
    delim = "KBWHSRC"
    assert ")" + delim not in text
    with open(out, "w") as f:
        f.write(f'namespace kbwh {{\nextern const char {symbol}[];\nconst char {symbol}[] = R"{delim}(')
        f.write(text)
        f.write(f'){delim}";\n}}  // namespace kbwh\n')
    return out


def _wait(item, verbose):
    cmd, p = item
    out, _ = p.communicate()
    if p.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + out)
        raise RuntimeError(f"compile failed: {cmd[-3]}")
    if verbose and out.strip():
        sys.stderr.write(out)


SAN_DIR = os.path.join(ROOT, "_build", "asan")
SAN_LIB = os.path.join(SAN_DIR, "libkbwh.so")
CLANGXX = "/opt/rocm/llvm/bin/clang++"


def asan_runtime():
    """The clang AddressSanitizer runtime the sanitized build links to (to
    preload into a process that loads it), or None."""
    try:
        out = subprocess.run([CLANGXX, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True,
                             text=True).stdout.strip()
    except OSError:
        return None
    if out and os.path.isabs(out) and os.path.exists(out):
        return out
    import glob

    hits = glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    return hits[0] if hits else None


def build_sanitized(jobs=8, verbose=False):
    """A build of the library with AddressSanitizer and UBSan on its host
    code (planners, plan checks, argument checks, the C ABI; device code is
    built as usual, uninstrumented) for the CPU tests.  Writes
    _build/asan/libkbwh.so."""
    os.makedirs(SAN_DIR, exist_ok=True)
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
    host_flags = [f for f in HIP_FLAGS if not f.startswith("--offload-arch")]
    procs, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(SAN_DIR, src.replace(".hip", ".o"))
        objs.append(obj)
        procs.append([HIPCC, f"--offload-arch={ARCH}", *host_flags, *HIP_SOURCE_FLAGS.get(src, []), "-g", *san,
                      "-c", os.path.join(CSRC, src), "-o", obj])
    cflags = ["-O1", "-g", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
              "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
    for src in CXX_SOURCES:
        obj = os.path.join(SAN_DIR, src.replace(".cpp", ".o"))
        objs.append(obj)
        procs.append([CLANGXX, *cflags, "-c", os.path.join(CSRC, src), "-o", obj])
    for hdr, sym in (("tile_dev.h", "kTileDevSource"), ("adj_dev.h", "kAdjDevSource")):
        emb = embed_header(os.path.join(CSRC, hdr), sym, os.path.join(SAN_DIR, hdr.replace(".h", "_src.cpp")))
        objs.append(emb[:-4] + ".o")
        procs.append([CLANGXX, *cflags, "-c", emb, "-o", objs[-1]])
    running = []
    for cmd in procs:
        running.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len(running) >= jobs:
            _wait(running.pop(0), verbose)
    while running:
        _wait(running.pop(0), verbose)
    tmp = SAN_LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address", "-Xarch_host",
          "-fsanitize=undefined", "-shared-libsan", "-o", tmp, *objs, "-ldl"])
    os.replace(tmp, SAN_LIB)
    return SAN_LIB


def build_oracle():
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], check=True)


if __name__ == "__main__":
    build(verbose=True)
    build_oracle()
    print(LIB)
