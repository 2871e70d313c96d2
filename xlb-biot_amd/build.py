"""Build ``lib/libxlb_biot.so`` (gfx950) from ``csrc/`` with hipcc.

Usage: ``python xlb-biot_amd/build.py``.  Packed-FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); the host pass warns that it
does not know the feature, which is harmless.  Floating-point contraction is off
(``-ffp-contract=off``): the collision is inlined in several places of the sweep
(pipeline prologue, unrolled loop bodies), and with contraction the compiler may
fuse multiply-adds differently in each copy, so results would depend in the last
bit on the launch shape; without it every cell gets the same arithmetic wherever
it is computed (the kernels are bandwidth-bound, the extra instructions are free).

The source compiles as ``PARTS`` objects in parallel (``-DXB_PART=k``: each defines
one group of the C entry points and instantiates only their kernels), linked into
one shared library.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "xlb_biot.hip")]
HEADERS = [os.path.join(HERE, "csrc", "xb_device.h"), os.path.join(ROOT, "include", "xlb_biot.h")]
OUT = os.path.join(HERE, "lib", "libxlb_biot.so")
PARTS = 11


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}\n{err}")
    return err


def build(verbose=False, jobs=None):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    objdir = os.path.join(HERE, "lib", "obj")
    os.makedirs(objdir, exist_ok=True)
    flags = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-function",
             "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc")]
    objs = [os.path.join(objdir, f"xlb_biot_part{k}.o") for k in range(PARTS)]
    cmds = [[hipcc()] + flags + [f"-DXB_PART={k}", "-c", "-o", o] + SOURCES for k, o in enumerate(objs)]
    jobs = jobs or min(PARTS, max(1, int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        errs = list(ex.map(_run, cmds))
    tmp = OUT + ".tmp"
    errs.append(_run([hipcc(), "--offload-arch=gfx950", "-fPIC", "-shared", "-o", tmp] + objs))
    err = "\n".join(e for e in errs if e.strip())
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
