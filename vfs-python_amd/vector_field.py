"""Epipolar vector fields of an equidistant stereo pair (Python/vector_field.py:28-65).

``get_vector_fields`` keeps the reference's signature and return convention
(host NumPy arrays).  The normaliser callable is applied to the unprojected
rays as in the reference; everything after it (the two projections of
XYZ +- t, the unit normalisation of the translation field, the rotation and
the calibration field) runs in one fused HIP kernel.
"""

import ctypes

import numpy as np
import torch

import vfs_hip as H


def unit_normalizer(xyz):
    """x / |x| per pixel: the normaliser of Python/main_equidistant.py:60."""
    return xyz / torch.linalg.vector_norm(xyz, dim=2, keepdim=True)


def get_vector_fields(grid, camera_ref, camera_other, R, t, t_scale, normalizer):
    """Python/vector_field.py:28-56.  Returns (tvf, cvf) as (H, W, 2) float32 NumPy arrays."""
    if isinstance(t_scale, (float, int)):
        t_scale = (t_scale, t_scale)
    camera_ref.setup_module()
    camera_other.setup_module()
    grid = grid.contiguous()
    XYZ = camera_ref.get_XYZs(grid)
    XYZ = normalizer(XYZ).to(torch.float32).contiguous()
    t = np.asarray(t, dtype=np.float64).reshape(3)
    st0 = (t_scale[0] * t).astype(np.float32)
    st1 = (t_scale[1] * t).astype(np.float32)
    Rm = np.ascontiguousarray(np.asarray(R, dtype=np.float32).reshape(3, 3))
    cr = camera_ref.get_intrincs_array()
    co = camera_other.get_intrincs_array()
    Hh, Ww = grid.shape[:2]
    tvf = torch.empty((Hh, Ww, 2), dtype=torch.float32, device=grid.device)
    cvf = torch.empty((Hh, Ww, 2), dtype=torch.float32, device=grid.device)

    def fp(a):
        return a.ctypes.data_as(ctypes.c_void_p)

    H.call("vfs_vector_fields", H.stream(), H.ptr(tvf), H.ptr(cvf), H.ptr(XYZ), H.ptr(grid), Hh, Ww,
           fp(cr), fp(co), fp(Rm), fp(st0), fp(st1))
    return tvf.cpu().numpy(), cvf.cpu().numpy()


def get_grid_zero(sz):
    """Python/vector_field.py:58-62: (H, W, 2) float32 pixel grid (x, y) on the GPU.
    Built on the device (integers below 2^24 are exact in float32), so no host array is
    copied over."""
    h, w = int(sz[0]), int(sz[1])
    uv = torch.empty((h, w, 2), dtype=torch.float32, device="cuda")
    uv[:, :, 0] = torch.arange(w, dtype=torch.float32, device="cuda")[None, :]
    uv[:, :, 1] = torch.arange(h, dtype=torch.float32, device="cuda")[:, None]
    return uv


def get_grid_one(sz):
    """Python/vector_field.py:64-65."""
    return get_grid_zero(sz) + 1
