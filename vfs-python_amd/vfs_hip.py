"""ctypes binding of libvfs_hip.so (the C ABI declared in include/vfs_hip.h).

The library is loaded from the directory of this file.  There is no fallback:
if the library is missing or a call fails, a RuntimeError is raised.  Device
memory is owned by torch tensors; this module only passes their addresses and
the current HIP stream.
"""

import ctypes
import os

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvfs_hip.so")
VFS_EINVAL = -22
VFS_ETOOBIG = -7
VFS_NOALLOC_MAX_PARTS = 1024

_lib = None

c_f = ctypes.c_float
c_i = ctypes.c_int
c_p = ctypes.c_void_p
c_i64 = ctypes.c_int64


class VfsParams(ctypes.Structure):
    """struct vfs_params (include/vfs_hip.h)."""

    _fields_ = [
        ("beta", c_f), ("gamma", c_f), ("alpha0", c_f), ("alpha1", c_f),
        ("eta_p", c_f), ("eta_q", c_f), ("lambda_", c_f), ("limit_range", c_f),
        ("f_scale", c_f), ("min_grad", c_f), ("min_tensor", c_f), ("a11", c_f), ("a004", c_f),
        ("median_k", ctypes.c_int32),
    ]


class VfsLevel(ctypes.Structure):
    """struct vfs_level (include/vfs_hip.h)."""

    _fields_ = [("H", ctypes.c_int32), ("W", ctypes.c_int32)] + [
        (n, c_p) for n in ("ref", "other", "tvf", "mask", "flow", "u", "u_bar", "ref_smoothed", "abc", "eta",
                           "tv", "warped", "deriv", "p", "q", "tp", "v_bar", "u0", "uf")
    ]


class VfsPdWs(ctypes.Structure):
    """struct vfs_pd_ws (include/vfs_hip.h)."""

    _fields_ = [(n, c_p) for n in ("u_th0", "u_th1", "u_bar2", "v_bar2", "p2", "q2")]


class VfsAdjOpts(ctypes.Structure):
    """struct vfs_adj_opts (include/vfs_hip.h)."""

    _fields_ = [(n, ctypes.c_int32) for n in ("fwd_tile_x", "fwd_tile_y", "fwd_k", "fwd_threads", "rec_tile_x",
                                              "rec_tile_y", "rec_k", "rec_threads", "sweep", "reverse_warps")] + [
        ("probe_events", ctypes.POINTER(ctypes.c_void_p)), ("probe_units", ctypes.POINTER(ctypes.c_int32)),
        ("probe_count", ctypes.POINTER(ctypes.c_int32)), ("probe_max", ctypes.c_int32), ("tape_save", c_p),
        ("tape_warps", ctypes.c_int32)]


class VfsAdjPd(ctypes.Structure):
    """struct vfs_adj_pd (include/vfs_hip.h)."""

    _fields_ = [("H", ctypes.c_int32), ("W", ctypes.c_int32)] + [
        (n, c_p) for n in ("gu", "gub", "gvb", "gp", "gq", "um", "vm", "qmsg", "tmsg", "gabc", "geta", "gderiv",
                           "gpar", "p", "q", "u", "p_prev", "q_prev", "u_prev", "u_bar_prev", "v_bar_prev", "abc",
                           "eta", "deriv", "mask")
    ]


# name -> argtypes (restype is int for all but vfs_version)
_SIGS = {
    "vfs_cast_and_scale": [c_p, c_p, c_p, c_i, c_i],
    "vfs_downsample_image": [c_p, c_p, c_i, c_i, c_p, c_i, c_i],
    "vfs_downsample_vector_field": [c_p, c_p, c_i, c_i, c_p, c_i, c_i],
    "vfs_downsample_mask": [c_p, c_p, c_i, c_i, c_p, c_i, c_i],
    "vfs_warp_image": [c_p, c_p, c_p, c_i, c_i, c_p, c_i, c_i],
    "vfs_gaussian_masked": [c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_compute_tensor_eta": [c_p, c_p, c_p, c_p, c_p, c_i, c_i, ctypes.POINTER(VfsParams)],
    "vfs_find_warping_vector": [c_p, c_p, c_p, c_i, c_i, c_p, c_i, c_i],
    "vfs_compute_derivatives": [c_p, c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_update_dual": [c_p, c_p, c_p, c_p, c_f, c_p, c_p, c_p, c_p, c_i, c_i, ctypes.POINTER(VfsParams)],
    "vfs_l1_thresholding": [c_p, c_p, c_f, c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i, ctypes.POINTER(VfsParams)],
    "vfs_update_primal": [c_p, c_p, c_p, c_f, c_f, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i,
                          ctypes.POINTER(VfsParams)],
    "vfs_median_filter": [c_p, c_p, c_p, c_p, c_i, c_i, c_i],
    "vfs_optical_flow_with_clamp": [c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i, c_f],
    "vfs_upsample": [c_p, c_p, c_i, c_i, c_p, c_i, c_i, c_i, c_f],
    "vfs_triangulate_rays": [c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_fill": [c_p, c_p, c_f, c_i64],
    "vfs_copy": [c_p, c_p, c_p, c_i64],
    "vfs_zero_masked": [c_p, c_p, c_p, c_i, c_i, c_i],
    "vfs_cam_get_rays": [c_p, c_p, c_i, c_i, c_f, c_f, c_f],
    "vfs_cam_get_uv_from_xyz": [c_p, c_p, c_p, c_i, c_i, c_f, c_f, c_f],
    "vfs_cam_get_xyz_from_uv": [c_p, c_p, c_p, c_i, c_i, c_f, c_f, c_f],
    "vfs_cam_get_thetas": [c_p, c_p, c_p, c_i, c_i],
    "vfs_cam_xyz_from_distance": [c_p, c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_vector_fields": [c_p, c_p, c_p, c_p, c_p, c_i, c_i, c_p, c_p, c_p, c_p, c_p],
    "vfs_solve_level_plain": [c_p, ctypes.POINTER(VfsLevel), ctypes.POINTER(VfsParams), c_p, c_p, c_p, c_i, c_i],
    "vfs_plan_create_level_plain": [ctypes.POINTER(c_p), ctypes.POINTER(VfsLevel), ctypes.POINTER(VfsParams),
                                    c_p, c_p, c_p, c_i, c_i],
    "vfs_pd_variants": [c_p, c_i],
    "vfs_pd_blocks": [c_p, c_i, c_i] + [c_p] * 14 + [c_i, c_p, c_p, c_p, c_i, c_i, c_i, c_i, c_i,
                                                     ctypes.POINTER(VfsParams)],
    "vfs_solve_level_blocked": [c_p, ctypes.POINTER(VfsLevel), ctypes.POINTER(VfsPdWs), ctypes.POINTER(VfsParams),
                                c_p, c_p, c_p, c_i, c_i, c_i, c_i, c_i, c_i],
    "vfs_batch_prepare_level": [c_p, ctypes.c_size_t, c_p, c_p, c_i, c_i, c_i, c_i, c_i],
    "vfs_solve_level_blocked_batch": [c_p, c_p, c_p, c_p, c_i, ctypes.POINTER(VfsParams), c_p, c_p, c_p, c_i, c_i,
                                      c_i, c_i, c_i, c_i],
    "vfs_pd_blocks_batch": [c_p, c_p, c_p, c_p, c_i, ctypes.POINTER(VfsParams), c_p, c_p, c_p, c_i, c_i, c_i, c_i,
                            c_i],
    "vfs_adj_pd_iteration": [c_p, ctypes.POINTER(VfsAdjPd), ctypes.POINTER(VfsParams), c_f, c_f, c_f, c_i],
    "vfs_adj_pd_finish": [c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_adj_clamp": [c_p] * 10 + [c_i, c_i, c_f, c_i],
    "vfs_adj_median": [c_p, c_p, c_p, c_p, c_i, c_i, c_i],
    "vfs_adj_derivatives": [c_p, c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_adj_warp": [c_p] * 7 + [c_i, c_i],
    "vfs_adj_tensor": [c_p] * 7 + [c_i, c_i, ctypes.POINTER(VfsParams)],
    "vfs_adj_gaussian": [c_p, c_p, c_p, c_p, c_i, c_i],
    "vfs_adj_downsample": [c_p, c_p, c_i, c_i, c_p, c_i, c_i],
    "vfs_adj_upsample": [c_p, c_p, c_i, c_i, c_p, c_i, c_i, c_i, c_f],
    "vfs_depth_loss": [c_p] * 6 + [c_i, c_i, c_f, c_f, c_f, c_f, c_f],
    "vfs_reduce_sum": [c_p, c_p, c_p, c_i, c_p, c_i64, c_i],
    "vfs_residual_norm": [c_p, c_p, c_p, c_i, c_p, c_p, c_p, c_i64, c_i],
    "vfs_residual_norm_noalloc": [c_p, c_p, c_i, c_p, c_p, c_p, c_i64, c_i],
    "vfs_adj_forward_level": [c_p, ctypes.POINTER(VfsLevel), ctypes.POINTER(VfsParams), c_p, c_p, c_p, c_i, c_i, c_i,
                              c_p, c_p, ctypes.POINTER(VfsAdjOpts)],
    "vfs_adj_reverse_level": [c_p, ctypes.POINTER(VfsLevel), ctypes.POINTER(VfsParams), c_p, c_p, c_p, c_i, c_i, c_i,
                              c_p, c_p, c_p, c_p, c_p, c_p, c_p, ctypes.POINTER(VfsAdjOpts)],
    "vfs_capture_begin": [ctypes.POINTER(c_p)],
    "vfs_capture_end": [c_p, ctypes.POINTER(c_p)],
    "vfs_plan_launch": [c_p, c_p],
    "vfs_plan_destroy": [c_p],
}


# entry points returning size_t
_SIZE_SIGS = {
    "vfs_batch_workspace_bytes": [c_i],
    "vfs_adj_save_bytes": [c_i, c_i, c_i, c_i, c_i],
    "vfs_adj_workspace_bytes": [c_i, c_i, c_i],
    "vfs_adj_tape_bytes": [c_i, c_i, c_i, c_i, c_i],
}


def exported_symbols():
    """Names of the C ABI entry points this binding expects (include/vfs_hip.h)."""
    return sorted(list(_SIGS) + list(_SIZE_SIGS) + ["vfs_version"])


def lib():
    """Load libvfs_hip.so once; raise if it is absent (there is no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with build_lib.py (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = c_i
        L.vfs_version.restype = ctypes.c_char_p
        L.vfs_version.argtypes = []
        for name, args in _SIZE_SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_size_t
        _lib = L
    return _lib


def check(rc, name):
    if rc != 0:
        if rc == VFS_EINVAL:
            raise RuntimeError(f"{name}: invalid argument (sizes or null pointer)")
        if rc == VFS_ETOOBIG:
            raise RuntimeError(f"{name}: size beyond the library's fixed capacity")
        raise RuntimeError(f"{name}: HIP error {rc}")


def call(name, *args):
    check(getattr(lib(), name)(*args), name)


def stream():
    """Handle of torch's current HIP stream on the current device."""
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    """Device address of a contiguous CUDA tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise RuntimeError("tensor is not on the GPU")
    if not t.is_contiguous():
        raise RuntimeError("tensor is not C-contiguous")
    return ctypes.c_void_p(t.data_ptr())


def make_params(param, median_k=5):
    """vfs_params from a vfs_params-like object; constants as in Python/vfs.py:52-68."""
    p = VfsParams()
    p.beta = float(param.beta)
    p.gamma = float(param.gamma)
    p.alpha0 = float(param.alpha0)
    p.alpha1 = float(param.alpha1)
    p.eta_p = 3.0
    p.eta_q = 2.0
    p.lambda_ = float(param.Lambda)
    p.limit_range = float(param.limitRange)
    p.f_scale = float(param.fScale)
    p.min_grad = 1e-8
    p.min_tensor = 1e-8
    # Python/vfs.cu:230-231 evaluates these in double before the float conversion
    p.a11 = float(np.float32(float(param.alpha1) * float(param.alpha1)))
    p.a004 = float(np.float32(4 * float(param.alpha0) * float(param.alpha0)))
    p.median_k = int(median_k)
    return p


def pd_variants():
    """(tile_x, tile_y, k, threads) of every blocked primal-dual variant in the library."""
    L = lib()
    n = L.vfs_pd_variants(None, 0)
    buf = (ctypes.c_int * (4 * n))()
    n = L.vfs_pd_variants(ctypes.cast(buf, c_p), n)
    return [tuple(buf[4 * i:4 * i + 4]) for i in range(n)]


def pd_schedule(n_iter, time_step_lambda):
    """tau, sigma, mu per primal-dual iteration (Python/vfs.py:494-515).

    Evaluated with the reference's own scalar expressions under NumPy >= 2
    rules (mu is float32 and keeps tau, sigma float32 after the first step).
    Returns three float32 arrays of length n_iter.
    """
    taus, sigmas, mus = [], [], []
    tau = 1.0
    sigma = 1.0 / tau
    for _ in range(n_iter):
        if sigma < 1000:
            mu = 1 / np.sqrt(1 + 0.7 * tau * time_step_lambda).astype(np.float32)
        else:
            mu = 1
        taus.append(np.float32(tau))
        sigmas.append(np.float32(sigma))
        mus.append(np.float32(mu))
        sigma = sigma / mu
        tau = tau * mu
    return (np.array(taus, dtype=np.float32), np.array(sigmas, dtype=np.float32),
            np.array(mus, dtype=np.float32))
