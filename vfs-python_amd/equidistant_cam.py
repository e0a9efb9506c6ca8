"""Equidistant fisheye camera r = f * theta (Python/equidistant_cam.py:34-206).

Same class name, constructor and methods as the reference.  Arrays are torch
float32 tensors on the GPU; the intrinsics travel as kernel arguments instead
of a __constant__ upload (Python/util_cuda.py:27-30).
"""

import numpy as np
import torch

import texture
import vfs_hip as H


def _as_device(a, dtype=torch.float32):
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.cuda()
        return t.to(dtype).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


class EquidistantCam:
    def __init__(self, p):
        self.f = p[0]
        self.u0 = p[1]
        self.v0 = p[2]
        self.gpu_module = None
        self.rays = None
        self.mask = None
        self.sz = None

    def get_intrincs_array(self):
        """Python/equidistant_cam.py:44-46."""
        return np.array([self.f, self.u0, self.v0], dtype=np.float32)

    def compile_module(self):
        """Loads libvfs_hip.so (prebuilt for gfx950; nothing is compiled at run time)."""
        self.gpu_module = H.lib()

    def setup_module(self):
        if self.gpu_module is None:
            self.compile_module()

    def _intr(self):
        p = self.get_intrincs_array()
        return float(p[0]), float(p[1]), float(p[2])

    def get_rays(self, sz=None):
        """getRays, Python/equidistant_cam.py:72-92: unit rays of the integer pixel grid."""
        self.setup_module()
        if sz is None:
            sz = self.sz
        rays = torch.empty((sz[0], sz[1], 3), dtype=torch.float32, device="cuda")
        H.call("vfs_cam_get_rays", H.stream(), H.ptr(rays), sz[0], sz[1], *self._intr())
        self.rays = rays
        return rays

    def get_UVs(self, xyz):
        """getUVFromXYZ, Python/equidistant_cam.py:94-114."""
        self.setup_module()
        assert xyz.is_contiguous()
        assert xyz.dtype == torch.float32
        uv = torch.empty((*xyz.shape[:2], 2), dtype=torch.float32, device=xyz.device)
        H.call("vfs_cam_get_uv_from_xyz", H.stream(), H.ptr(uv), H.ptr(xyz), uv.shape[0], uv.shape[1], *self._intr())
        return uv

    def get_XYZs(self, uv):
        """getXYZFromUV, Python/equidistant_cam.py:116-136."""
        self.setup_module()
        assert uv.is_contiguous()
        assert uv.dtype == torch.float32
        xyz = torch.empty((*uv.shape[:2], 3), dtype=torch.float32, device=uv.device)
        H.call("vfs_cam_get_xyz_from_uv", H.stream(), H.ptr(xyz), H.ptr(uv), xyz.shape[0], xyz.shape[1], *self._intr())
        return xyz

    def get_thetas(self):
        """getThetas, Python/equidistant_cam.py:138-159."""
        self.setup_module()
        if self.rays is None:
            self.get_rays()
        thetas = torch.empty_like(self.rays)
        H.call("vfs_cam_get_thetas", H.stream(), H.ptr(thetas), H.ptr(self.rays), thetas.shape[0], thetas.shape[1])
        return thetas

    def create_default_mask(self, sz=None):
        """Python/equidistant_cam.py:161-170: all-ones mask (Border/Point texture)."""
        if sz is None:
            sz = self.sz
        mask = torch.ones((sz[0], sz[1]), dtype=torch.uint8, device="cuda")
        self.mask = mask
        self.to_mask = texture.create_texture_object(mask, addressMode=texture.cudaAddressModeBorder,
                                                     filterMode=texture.cudaFilterModePoint)

    def set_mask(self, mask):
        """Python/equidistant_cam.py:172-178."""
        mask = _as_device(mask, torch.uint8)
        self.mask = mask
        self.to_mask = texture.create_texture_object(mask, addressMode=texture.cudaAddressModeBorder,
                                                     filterMode=texture.cudaFilterModePoint)

    def get_xyz_from_distance(self, D):
        """getXYZFromDistance, Python/equidistant_cam.py:180-206: D * ray, (0, 0, -1) outside the mask."""
        self.setup_module()
        D = _as_device(D)
        if self.rays is None:
            self.get_rays(D.shape[:2])
        if self.mask is None:
            self.create_default_mask(D.shape[:2])
        assert tuple(D.shape) == tuple(self.rays.shape[:2])
        xyz = torch.empty_like(self.rays)
        m = self.to_mask.tensor
        H.call("vfs_cam_xyz_from_distance", H.stream(), H.ptr(xyz), H.ptr(self.rays), H.ptr(D), H.ptr(m),
               xyz.shape[0], xyz.shape[1])
        return xyz
