"""Builds libnnpe.so (the HIP kernels + C ABI) in-tree for gfx950 with hipcc.

Every .hip source under csrc/ is compiled to an object (in parallel) and linked into
nnpengine_amd/libnnpe.so.  Flags:
  -ffp-contract=off   no FMA contraction: distances, cell indices, table lerps and g(r) bins
                      round exactly as the reference's FP32 code (bit-exact lists and counts);
                      kernels that want FMAs call fmaf / MFMA builtins explicitly.
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed-FP32 instructions.
"""
from __future__ import annotations

import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnnpe.so")
ARCH = os.environ.get("NNPE_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument", "-Wno-unused-result"]


def _compile(src, force):
    obj = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o"))
    deps = [src] + glob.glob(os.path.join(CSRC, "*.h")) + [os.path.join(HERE, "..", "include", "nnpe.h")]
    if not force and os.path.exists(obj) and os.path.getmtime(obj) >= max(os.path.getmtime(d) for d in deps):
        return obj, ""
    cmd = [hipcc(), "-c", src, "-o", obj] + COMMON
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr[-6000:]}")
    return obj, res.stderr


def build(force=False, verbose=False):
    """Compile every HIP source and link libnnpe.so.  Returns the library path."""
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        results = list(ex.map(lambda s: _compile(s, force), srcs))
    objs = [r[0] for r in results]
    if verbose:
        for o, err in results:
            if err.strip():
                print(f"[{os.path.basename(o)}] {err.strip()[:2000]}", file=sys.stderr)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(o) for o in objs):
        tmp = LIB + ".tmp"
        cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"link failed:\n{res.stderr[-4000:]}")
        os.replace(tmp, LIB)
    build_host(force)
    return LIB


HOST_EXE = os.path.join(HERE, "nnpe_run")


def build_host(force=False):
    """The reference-style C++ driver (host/nnpe_run.cpp over host/nnpe_host.cpp), linked
    against libnnpe.so (rpath $ORIGIN)."""
    srcs = [os.path.join(HERE, "host", f) for f in ("nnpe_host.cpp", "nnpe_run.cpp")]
    deps = srcs + [os.path.join(HERE, "host", "nnpe_host.h"), LIB]
    if not force and os.path.exists(HOST_EXE) and os.path.getmtime(HOST_EXE) >= max(os.path.getmtime(d) for d in deps):
        return HOST_EXE
    cmd = [shutil.which("g++") or "g++", "-O2", "-std=c++17", "-o", HOST_EXE] + srcs + \
          ["-L" + HERE, "-lnnpe", "-Wl,-rpath,$ORIGIN"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host driver build failed:\n{res.stderr[-4000:]}")
    return HOST_EXE


if __name__ == "__main__":
    print(build(force="--force" in sys.argv, verbose=True))
