"""Build libflagfft.so (gfx950 HIP kernels + C++ host runtime) in-tree.

Every source under flagfft_amd/csrc is compiled by hipcc for gfx950 and linked
into flagfft_amd/libflagfft.so, the C-ABI library declared by include/flagfft.h
and include/flagfft_ext.h.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libflagfft.so")
ARCH = os.environ.get("FLAGFFT_OFFLOAD_ARCH", "gfx950")

SOURCES = ["kernels.hip", "leaf.hip", "stage.hip", "colpass.hip", "bluestein_lds.hip", "leaf_split.hip", "mfma.hip", "planner.cpp", "exec.cpp", "api.cpp", "serve.cpp", "tunedb.cpp"]
HEADERS = ["fft_args.hpp", "codelets.hpp", "dev_common.hpp", "spec_common.hpp", "kernels.hpp", "colpass_body.hpp",
           "planner.hpp", "exec.hpp", "tunedb.hpp"]

COMMON_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # no packed-FP32 VALU instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


# Sources whose hashes fingerprint the tuned-plan rows: a row applies only to
# the planner and kernels it was measured with.
PLANNER_SOURCES = ["planner.cpp", "planner.hpp", "exec.cpp", "exec.hpp"]
KERNEL_SOURCES = [s for s in SOURCES if s.endswith(".hip")] + [
    "fft_args.hpp", "codelets.hpp", "dev_common.hpp", "spec_common.hpp", "kernels.hpp", "colpass_body.hpp"]


def source_hash(names) -> str:
    """First 16 hex digits of the SHA-256 over the named csrc files (name and bytes)."""
    import hashlib

    h = hashlib.sha256()
    for name in sorted(names):
        h.update(name.encode() + b"\0")
        with open(os.path.join(CSRC, name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def fingerprint_flags():
    return [f'-DFLAGFFT_PLANNER_HASH="{source_hash(PLANNER_SOURCES)}"',
            f'-DFLAGFFT_KERNEL_HASH="{source_hash(KERNEL_SOURCES)}"']


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _newest_dep() -> float:
    deps = [os.path.join(CSRC, h) for h in HEADERS] + [
        os.path.join(ROOT, "include", "flagfft.h"), os.path.join(ROOT, "include", "flagfft_ext.h"),
        os.path.abspath(__file__)]
    return max(os.path.getmtime(p) for p in deps if os.path.exists(p))


def _compile(src: str, force: bool) -> str:
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    if not force and os.path.exists(obj):
        deps = [path] + ([os.path.join(CSRC, n) for n in PLANNER_SOURCES + KERNEL_SOURCES] if src == "tunedb.cpp" else [])
        if os.path.getmtime(obj) >= max(max(os.path.getmtime(d) for d in deps), _newest_dep()):
            return obj
    flags = COMMON_FLAGS + (fingerprint_flags() if src == "tunedb.cpp" else [])
    cmd = [hipcc()] + flags + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [hipcc()] + flags + ["-x", "hip", "-c", path, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed for {src}")
    return obj


def build(force: bool = True, jobs: int = 4) -> str:
    """Compile all HIP/C++ sources for gfx950 and link libflagfft.so; returns its path."""
    os.makedirs(OBJ, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(lambda s: _compile(s, force), SOURCES))
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(o) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-ldl", "-lpthread"]
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError("link of libflagfft.so failed")
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(force="--force" in sys.argv))
