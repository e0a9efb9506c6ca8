"""Build libvfs_hip.so (gfx950) in-tree with hipcc.

Usage: python build_lib.py [--arch gfx950] [--out PATH]

All HIP sources under csrc/ are compiled into one shared library next to this
file.  Packed-FP32 instruction selection is disabled for every HIP source.

Floating-point contraction is limited to single expressions (-ffp-contract=on)
instead of hipcc's default of fusing a multiply into an add across statements.
The solve's paths (per-kernel, blocked, batched) and the adjoint's sweeps inline
the same per-pixel functions into different kernels; with cross-statement fusion
the compiler may fuse a product into a later addition in one kernel and not in
another, so the paths could differ in the last bit.  Within one expression the
choice is made from the source alone, so every kernel rounds each shared
function the same way.
"""

import argparse
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_NAME = "libvfs_hip.so"


def hip_sources():
    return sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))


def build(arch="gfx950", out=None, verbose=False):
    out = out or os.path.join(HERE, LIB_NAME)
    srcs = hip_sources()
    if not srcs:
        raise RuntimeError("no HIP sources found under csrc/")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    tmp = out + ".tmp"
    cmd = [hipcc, f"--offload-arch={arch}", "-O3", "-std=c++17", "-fPIC", "-shared",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-ffp-contract=on",
           "-Wall", "-Wno-unused-function",
           "-I", os.path.join(HERE, "csrc"), "-I", os.path.join(HERE, "..", "include"),
           *srcs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed with exit code {res.returncode}")
    os.replace(tmp, out)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    print(build(a.arch, a.out, verbose=True))


if __name__ == "__main__":
    main()
