"""Build ``lib/libxlb_biot.so`` (gfx950) from ``csrc/`` with hipcc.

Usage: ``python xlb-biot_amd/build.py``.  Packed-FP32 instructions are disabled
(``-Xclang -target-feature -Xclang -packed-fp32-ops``); the host pass warns that it
does not know the feature, which is harmless.  Floating-point contraction is off
(``-ffp-contract=off``): the collision is inlined in several places of the sweep
(pipeline prologue, unrolled loop bodies), and with contraction the compiler may
fuse multiply-adds differently in each copy, so results would depend in the last
bit on the launch shape; without it every cell gets the same arithmetic wherever
it is computed (the kernels are bandwidth-bound, the extra instructions are free).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = [os.path.join(HERE, "csrc", "xlb_biot.hip")]
HEADERS = [os.path.join(HERE, "csrc", "xb_device.h"), os.path.join(ROOT, "include", "xlb_biot.h")]
OUT = os.path.join(HERE, "lib", "libxlb_biot.so")


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if cand and (os.path.sep not in cand or os.path.exists(cand)):
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose=False):
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"), "-o", tmp] + SOURCES
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{err}")
    if verbose and err.strip():
        print(err, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
