"""ctypes binding of oracle/vfs_oracle.c, the float64 C restatement of the solve.

TEST INFRASTRUCTURE ONLY: loaded by tests/ and by the cpu_baseline leg of bench.py,
never by the product (vfs-python_amd/).  ``build()`` compiles the library with gcc
next to this file; ``solve_stereo`` has the arguments of vfs_oracle.solve_stereo and
returns the same flow to rounding (tests/test_oracle_c.py pins the two to each other).
"""

import ctypes
import os
import subprocess

import numpy as np

from . import vfs_oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "vfs_oracle.c")
LIB = os.path.join(HERE, "libvfs_oracle.so")


class _Params(ctypes.Structure):
    _fields_ = [("beta", ctypes.c_double), ("gamma", ctypes.c_double), ("alpha0", ctypes.c_double),
                ("alpha1", ctypes.c_double), ("Lambda", ctypes.c_double), ("f_scale", ctypes.c_double),
                ("limit_range", ctypes.c_double), ("n_level", ctypes.c_int32), ("n_warp", ctypes.c_int32),
                ("n_iter", ctypes.c_int32), ("median_k", ctypes.c_int32), ("n_threads", ctypes.c_int32)]


def build():
    """gcc -O3 (IEEE double semantics: no fast-math, no FMA contraction) into libvfs_oracle.so next to the source."""
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O3", "-ffp-contract=off", "-fopenmp", "-std=c99", "-Wall", "-fPIC", "-shared", SRC, "-o", tmp, "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        p = ctypes.c_void_p
        L.vfs_oracle_solve.argtypes = [p, p, p, p, p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_Params), p, p, p, p]
        L.vfs_oracle_solve.restype = ctypes.c_int
        _lib = L
    return _lib


def solve_stereo(img_ref, img_other, mask, tvf, cvf, prm, median_k=O.MEDIAN_KERNEL_SIZE, threads=1):
    """vfs_oracle.solve_stereo in C: (H, W, 2) float64 flow.  ``threads``: OpenMP threads over
    image rows (the result does not depend on it)."""
    ref = np.ascontiguousarray(img_ref, dtype=np.uint8)
    oth = np.ascontiguousarray(img_other, dtype=np.uint8)
    msk = np.ascontiguousarray(mask, dtype=np.uint8)
    tv = np.ascontiguousarray(tvf, dtype=np.float32)
    cv = np.ascontiguousarray(cvf, dtype=np.float32)
    h, w = ref.shape
    assert oth.shape == msk.shape == (h, w) and tv.shape == cv.shape == (h, w, 2)
    taus, sigmas, mus = (np.ascontiguousarray(a, dtype=np.float64)
                         for a in O.pd_schedule(prm.nSolverIters, prm.timeStepLambda))
    P = _Params(float(prm.beta), float(prm.gamma), float(prm.alpha0), float(prm.alpha1), float(prm.Lambda),
                float(prm.fScale), float(prm.limitRange), int(prm.nLevel), int(prm.nWarpIters),
                int(prm.nSolverIters), int(median_k), int(threads))
    out = np.empty((h, w, 2), np.float64)
    ptr = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = lib().vfs_oracle_solve(ptr(ref), ptr(oth), ptr(msk), ptr(tv), ptr(cv), h, w, ctypes.byref(P), ptr(taus),
                                ptr(sigmas), ptr(mus), ptr(out))
    if rc != 0:
        raise RuntimeError(f"vfs_oracle_solve returned {rc}")
    return out
