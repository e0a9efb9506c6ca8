"""Variational fisheye stereo solver on gfx950 (drop-in for Python/vfs.py:33-559).

Same class, method names, argument order and attributes as the reference's
``VFS``.  Arrays are torch tensors on the GPU instead of CuPy arrays, and the
``*_to`` arguments are ``texture.SampledImage`` descriptors instead of CUDA
texture objects.  Every kernel runs in libvfs_hip.so (hand-written HIP for
gfx950) through the C ABI in include/vfs_hip.h; there is no CPU fallback.

Parameters are read at each call (the reference bakes them into the CUDA
source at the first ``setup_module``, Python/vfs.py:38-74).

``solve_stereo`` runs the reference's loop (Python/vfs.py:460-533).  The
attribute ``solver`` selects how it is driven:

* ``"blocked"`` (default) / ``"blocked_eager"``: the primal-dual loop temporally
  blocked (k iterations per launch on LDS-exchanged register tiles, see
  csrc/vfs_pd_block.hip) and the per-warp kernels fused, recorded once into a HIP
  graph and replayed per call, or enqueued per call.  The tile shape and k per
  pyramid level come from ``pd_tiles`` (a dict {level: (tile_x, tile_y, k[, threads])})
  or from the autotuned table (autotune.py).  The result equals the per-kernel
  solve's bit for bit;
* ``"graph"``: the whole solve with the reference's kernel decomposition --
  pyramids, every level's warp loop, upsampling, final masking -- recorded once
  into a HIP graph and replayed per call;
* ``"eager"``: the same launches enqueued from C++ on each call;
* ``"python"``: the reference's host loop literally, one Python call per
  kernel through the per-kernel methods below.

A recorded solve holds at most ``GRAPH_MAX_LAUNCHES`` kernel launches; a
larger one (the per-kernel solve at the default 11 levels x 100 warps x 100
iterations issues ~337 k) is enqueued eagerly instead, with the same kernels
and bitwise the same result.  ``launch_mode`` tells which ran.
"""

import ctypes
import math

import numpy as np
import torch

import texture
import vfs_hip as H

_DEV = "cuda"

# Largest solve recorded into one HIP graph.  Recorded solves of up to ~47 k
# launches (the blocked solve at the default parameters, 800x800) replay normally; the
# recording of a ~337 k-launch per-kernel solve ended the process with a
# segmentation fault inside graph capture / instantiation, so solves above this
# bound are launched eagerly.
GRAPH_MAX_LAUNCHES = 50_000


def _u8_device(img):
    if isinstance(img, torch.Tensor):
        t = img if img.is_cuda else img.cuda()
        return t.to(torch.uint8).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(img, dtype=np.uint8)), device=_DEV)


def _f32_device(a):
    if isinstance(a, torch.Tensor):
        t = a if a.is_cuda else a.cuda()
        return t.to(torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float32)), device=_DEV)


def _empty(shape, dtype=torch.float32):
    return torch.empty(shape, dtype=dtype, device=_DEV)


class VFS:
    def __init__(self, param):
        self.param = param
        self.gpu_module = None
        self.solver = "blocked"
        # solver="python" only: per warp, the RMS over the mask of (thresholded u -
        # warp-start u), as (level, warp, rms) in self.residuals after a solve
        self.track_residuals = False
        self.residuals = []
        self.pd_tiles = None
        self.median_filter_kernel_size = 5
        self.launch_mode = None  # "graph" or "eager": how the last solve_stereo ran
        self._plan = None
        self._plan_key = None
        self._bufs = None
        self._t_dev = None  # (translation values, their device copy) of the last triangulation

    # -- module setup (Python/vfs.py:38-74) ---------------------------------------------
    def compile_module(self):
        """Loads the prebuilt gfx950 library; parameters are runtime arguments."""
        self.gpu_module = H.lib()

    def setup_module(self):
        if self.gpu_module is None:
            self.compile_module()

    def _prm(self):
        return H.make_params(self.param, self.median_filter_kernel_size)

    @staticmethod
    def compute_optical_flow(u, vec):
        """Python/vfs.py:76-82."""
        assert len(u.shape) == 2
        assert len(vec.shape) == 3
        assert tuple(u.shape) == tuple(vec.shape[:2])
        assert vec.shape[2] == 2
        return u[:, :, None] * vec

    def compute_pyramid_sizes(self):
        """Python/vfs.py:84-91."""
        sizes = list()
        h, w = self.param.height, self.param.width
        for k in range(self.param.nLevel):
            sizes.append((h, w))
            h = int(h / self.param.fScale)
            w = int(w / self.param.fScale)
        self.sizes = sizes

    # -- pyramids (Python/vfs.py:93-218) -------------------------------------------------
    def _image_pyramid_into(self, stream, pyr):
        for k in range(1, len(pyr)):
            d, s = pyr[k], pyr[k - 1]
            H.call("vfs_downsample_image", stream, H.ptr(d), d.shape[0], d.shape[1], H.ptr(s), s.shape[0], s.shape[1])

    def create_image_pyramid_(self, img_f32):
        """Python/vfs.py:93-119."""
        self.setup_module()
        pyr = [img_f32] + [_empty(sz) for sz in self.sizes[1:]]
        self._image_pyramid_into(H.stream(), pyr)
        return pyr, [texture.wrap(x) for x in pyr]

    def create_image_pyramid(self, img_ref, img_other):
        """Python/vfs.py:121-157: u8 -> float/256, de-rotate the other view by the calibration field, pyramids."""
        self.setup_module()
        assert tuple(img_ref.shape) == tuple(img_other.shape)
        s = H.stream()
        r8, o8 = _u8_device(img_ref), _u8_device(img_other)
        ref0, other0 = _empty(r8.shape), _empty(o8.shape)
        for d, src in ((ref0, r8), (other0, o8)):
            H.call("vfs_cast_and_scale", s, H.ptr(d), H.ptr(src), d.shape[0], d.shape[1])
        refs, refs_to = self.create_image_pyramid_(ref0)
        other0_warp = self.warping_image(texture.wrap(other0), self.calibration)
        others, others_to = self.create_image_pyramid_(other0_warp)
        self.refs, self.others = refs, others
        self.refs_to, self.others_to = refs_to, others_to

    def create_mask_pyramid(self, mask):
        """Python/vfs.py:159-187."""
        self.setup_module()
        masks = [_u8_device(mask)]
        s = H.stream()
        for sz in self.sizes[1:]:
            d = _empty(sz, torch.uint8)
            src = masks[-1]
            H.call("vfs_downsample_mask", s, H.ptr(d), sz[0], sz[1], H.ptr(src), src.shape[0], src.shape[1])
            masks.append(d)
        self.masks = masks
        self.masks_to = [texture.wrap(m, texture.cudaFilterModePoint) for m in masks]

    def create_vector_field_pyramid(self, translation, calibration):
        """Python/vfs.py:189-218."""
        self.setup_module()
        self.calibration = _f32_device(calibration)
        translations = [_f32_device(translation)]
        s = H.stream()
        for sz in self.sizes[1:]:
            d = _empty((*sz, 2))
            src = translations[-1]
            H.call("vfs_downsample_vector_field", s, H.ptr(d), sz[0], sz[1], H.ptr(src), src.shape[0], src.shape[1])
            translations.append(d)
        self.translations = translations
        self.translations_to = [texture.wrap(t) for t in translations]

    # -- per-kernel wrappers (Python/vfs.py:220-458) ----------------------------------------
    def warping_image(self, to, warp):
        """warpingImage, Python/vfs.py:220-236."""
        self.setup_module()
        src = to.require(texture.cudaAddressModeMirror)
        assert warp.is_contiguous()
        sz = tuple(warp.shape[:2])
        d = _empty(sz)
        H.call("vfs_warp_image", H.stream(), H.ptr(d), H.ptr(warp), sz[0], sz[1], H.ptr(src), src.shape[0],
               src.shape[1])
        return d

    def gaussian(self, to, to_mask, sz):
        """gaussianMasked, Python/vfs.py:238-253."""
        self.setup_module()
        src = to.require(texture.cudaAddressModeMirror)
        m = to_mask.require(texture.cudaAddressModeMirror)
        sz = tuple(sz)
        assert tuple(src.shape[:2]) == sz and tuple(m.shape[:2]) == sz
        d = _empty(sz)
        H.call("vfs_gaussian_masked", H.stream(), H.ptr(d), H.ptr(src), H.ptr(m), sz[0], sz[1])
        return d

    def compute_tensor_eta(self, img, to_mask):
        """computeTensorEtaMasked, Python/vfs.py:255-275.  Returns (abc, eta), each (H, W, 3)."""
        self.setup_module()
        img = img.contiguous()
        m = to_mask.require(texture.cudaAddressModeMirror)
        assert tuple(m.shape[:2]) == tuple(img.shape)
        d_tensor = _empty((*img.shape, 3))
        d_eta = _empty((*img.shape, 3))
        prm = self._prm()
        H.call("vfs_compute_tensor_eta", H.stream(), H.ptr(d_tensor), H.ptr(d_eta), H.ptr(img), H.ptr(m),
               img.shape[0], img.shape[1], ctypes.byref(prm))
        return d_tensor, d_eta

    def find_warping_vector(self, warpUV, to_translation):
        """findWarpingVector, Python/vfs.py:277-293."""
        self.setup_module()
        assert warpUV.is_contiguous()
        tvf = to_translation.require(texture.cudaAddressModeMirror)
        d = torch.empty_like(warpUV)
        H.call("vfs_find_warping_vector", H.stream(), H.ptr(d), H.ptr(warpUV), d.shape[0], d.shape[1], H.ptr(tvf),
               tvf.shape[0], tvf.shape[1])
        return d

    def compute_derivatives(self, wapred_other, to_ref, to_translation):
        """computeDerivatives, Python/vfs.py:295-313."""
        self.setup_module()
        other = wapred_other.contiguous()
        ref = to_ref.require(texture.cudaAddressModeMirror)
        tvf = to_translation.require(texture.cudaAddressModeMirror)
        sz = tuple(other.shape[:2])
        assert tuple(ref.shape[:2]) == sz and tuple(tvf.shape[:2]) == sz
        d = _empty((*sz, 2))
        H.call("vfs_compute_derivatives", H.stream(), H.ptr(d), H.ptr(ref), H.ptr(other), H.ptr(tvf), sz[0], sz[1])
        return d

    def update_dual_variables(self, p, q, sigma, u, v, tensor_abc, to_mask):
        """updateDualVariables, Python/vfs.py:315-335 (p and q in place)."""
        self.setup_module()
        for a in (p, q, u, v, tensor_abc):
            assert a.is_contiguous()
        assert tuple(u.shape) == tuple(v.shape[:2])
        m = to_mask.require(texture.cudaAddressModeMirror)
        tp = torch.empty_like(p)
        prm = self._prm()
        H.call("vfs_update_dual", H.stream(), H.ptr(p), H.ptr(tp), H.ptr(q), float(np.float32(sigma)), H.ptr(u),
               H.ptr(v), H.ptr(tensor_abc), H.ptr(m), p.shape[0], p.shape[1], ctypes.byref(prm))
        return p, tp, q

    def l1_thresholding(self, tau, tp, u, u_, derivatives, eta, to_mask):
        """l1Thresholding, Python/vfs.py:337-357."""
        self.setup_module()
        for a in (tp, u, u_, derivatives, eta):
            assert a.is_contiguous()
        assert tuple(u.shape) == tuple(u_.shape)
        m = to_mask.require(texture.cudaAddressModeMirror)
        us = torch.empty_like(u)
        prm = self._prm()
        H.call("vfs_l1_thresholding", H.stream(), H.ptr(us), float(np.float32(tau)), H.ptr(tp), H.ptr(u), H.ptr(u_),
               H.ptr(derivatives), H.ptr(eta), H.ptr(m), us.shape[0], us.shape[1], ctypes.byref(prm))
        return us

    def update_primal_variables(self, tau, mu, p, q, u, u_, v_, tensor_abc, eta, to_mask):
        """updatePrimalVariables, Python/vfs.py:359-382."""
        self.setup_module()
        for a in (p, q, u, u_, v_, tensor_abc, eta):
            assert a.is_contiguous()
        assert tuple(u_.shape) == tuple(v_.shape[:2])
        m = to_mask.require(texture.cudaAddressModeMirror)
        us = torch.empty_like(u_)
        vs = torch.empty_like(v_)
        prm = self._prm()
        H.call("vfs_update_primal", H.stream(), H.ptr(us), H.ptr(vs), float(np.float32(tau)), float(np.float32(mu)),
               H.ptr(p), H.ptr(q), H.ptr(u), H.ptr(u_), H.ptr(v_), H.ptr(tensor_abc), H.ptr(eta), H.ptr(m),
               us.shape[0], us.shape[1], ctypes.byref(prm))
        return us, vs

    def median_filter(self, u, to_mask):
        """medianFilter, Python/vfs.py:384-400."""
        self.setup_module()
        assert u.is_contiguous()
        m = to_mask.require(texture.cudaAddressModeMirror)
        d = torch.empty_like(u)
        H.call("vfs_median_filter", H.stream(), H.ptr(d), H.ptr(u), H.ptr(m), d.shape[0], d.shape[1],
               self.median_filter_kernel_size)
        return d

    def compute_optical_flow_with_clamp(self, warpUV_, u_, u, translation_vector, to_mask):
        """computeOpticalFlowWithClamp, Python/vfs.py:402-418 (warpUV_ and u_ in place)."""
        self.setup_module()
        for a in (warpUV_, u_, u, translation_vector):
            assert a.is_contiguous()
        m = to_mask.require(texture.cudaAddressModeMirror)
        H.call("vfs_optical_flow_with_clamp", H.stream(), H.ptr(warpUV_), H.ptr(u_), H.ptr(u),
               H.ptr(translation_vector), H.ptr(m), warpUV_.shape[0], warpUV_.shape[1],
               float(self.param.limitRange))
        return warpUV_, u_

    def upsampling_vector_field1(self, u, shape):
        """upSamplingVectorField1, Python/vfs.py:420-438."""
        self.setup_module()
        u = u.contiguous()
        d = _empty(tuple(shape))
        H.call("vfs_upsample", H.stream(), H.ptr(d), shape[0], shape[1], H.ptr(u), u.shape[0], u.shape[1], 1,
               float(self.param.fScale))
        return d

    def upsampling_vector_field(self, warp, shape):
        """upSamplingVectorField, Python/vfs.py:440-458."""
        self.setup_module()
        warp = warp.contiguous()
        d = _empty((*shape, 2))
        H.call("vfs_upsample", H.stream(), H.ptr(d), shape[0], shape[1], H.ptr(warp), warp.shape[0], warp.shape[1],
               2, float(self.param.fScale))
        return d

    # -- the solve (Python/vfs.py:460-533) -----------------------------------------------
    def solve_stereo(self, img_ref, img_other):
        """Coarse-to-fine solve; returns the flow (H, W, 2) and stores it as ``warpUV``."""
        self.setup_module()
        if self.solver == "python":
            self.launch_mode = "python"
            return self._solve_stereo_python(img_ref, img_other)
        if self.solver not in ("graph", "eager", "blocked", "blocked_eager"):
            raise ValueError(f"unknown solver {self.solver!r}")
        b = self._buffers()
        b["img_ref"].copy_(_u8_device(img_ref))
        b["img_other"].copy_(_u8_device(img_other))
        if self.solver in ("eager", "blocked_eager") or self.launch_count() > GRAPH_MAX_LAUNCHES:
            self.launch_mode = "eager"
            self._enqueue_solve(H.stream(), b)
        else:
            self.launch_mode = "graph"
            key = self._key(b)
            if self._plan is None or self._plan_key != key:
                self.release_plan()
                cs = ctypes.c_void_p()
                H.call("vfs_capture_begin", ctypes.byref(cs))
                try:
                    self._enqueue_solve(cs, b)
                finally:
                    plan = ctypes.c_void_p()
                    H.call("vfs_capture_end", cs, ctypes.byref(plan))
                self._plan, self._plan_key = plan, key
            H.call("vfs_plan_launch", self._plan, H.stream())
        self.refs, self.others = b["refs"], b["others"]
        self.refs_to = [texture.wrap(x) for x in b["refs"]]
        self.others_to = [texture.wrap(x) for x in b["others"]]
        # a fresh tensor per call, as the reference's solve returns a new array
        # (Python/vfs.py:462-533); the level buffers are reused by the next solve
        self.warpUV = b["result"].clone()
        return self.warpUV

    def release_plan(self):
        if self._plan is not None:
            H.lib().vfs_plan_destroy(self._plan)
            self._plan = None
            self._plan_key = None

    def __del__(self):
        try:
            self.release_plan()
        except Exception:
            pass

    def _key(self, b):
        p = self.param
        return (tuple(self.sizes), int(p.nWarpIters), int(p.nSolverIters), float(p.timeStepLambda),
                float(p.beta), float(p.gamma), float(p.alpha0), float(p.alpha1), float(p.Lambda),
                float(p.limitRange), float(p.fScale), self.median_filter_kernel_size,
                tuple(m.data_ptr() for m in self.masks), tuple(t.data_ptr() for t in self.translations),
                self.calibration.data_ptr(), id(b), self.solver, tuple(self._tiles()))

    def launch_count(self):
        """Kernel launches of one solve_stereo with the current sizes, solver and schedule."""
        p = self.param
        n_warp, n_iter = int(p.nWarpIters), int(p.nSolverIters)
        n = 4 * len(self.sizes) + 4  # pyramids, level starts, result
        blocked = self.solver in ("blocked", "blocked_eager")
        tiles = self._tiles() if blocked else [None] * len(self.sizes)
        for t in tiles:
            n += 3
            n += n_warp * ((2 + -(-n_iter // t[2])) if blocked else (6 + 3 * n_iter))
        return n

    def _tiles(self):
        """(tile_x, tile_y, k, threads) of the blocked primal-dual loop for every pyramid level."""
        import autotune
        out = []
        for lvl, (h, w) in enumerate(self.sizes):
            if self.pd_tiles is not None and lvl in self.pd_tiles:
                out.append(autotune.variant(self.pd_tiles[lvl]))
            else:
                out.append(autotune.tile_for(h, w, int(self.param.nSolverIters)))
        return out

    def _buffers(self):
        """Persistent device buffers of the solve, sized by self.sizes (reused across calls)."""
        sizes = tuple(self.sizes)
        if self._bufs is not None and self._bufs["sizes"] == sizes:
            return self._bufs
        self.release_plan()
        h0, w0 = sizes[0]
        b = dict(sizes=sizes)
        b["img_ref"] = _empty((h0, w0), torch.uint8)
        b["img_other"] = _empty((h0, w0), torch.uint8)
        b["ref0"] = _empty((h0, w0))
        b["other0"] = _empty((h0, w0))
        b["refs"] = [b["ref0"]] + [_empty(sz) for sz in sizes[1:]]
        b["others"] = [_empty(sz) for sz in sizes]
        b["levels"] = []
        for sz in sizes:
            e = lambda *c: _empty((*sz, *c))  # noqa: E731
            b["levels"].append(dict(flow=e(2), u=e(), u_bar=e(), ref_smoothed=e(), abc=e(3), eta=e(3), tv=e(2),
                                    warped=e(), deriv=e(2), p=e(2), q=e(4), tp=e(2), v_bar=e(2), u0=e(), uf=e()))
            b.setdefault("ws", []).append(dict(u_th0=e(), u_th1=e(), u_bar2=e(), v_bar2=e(2), p2=e(2), q2=e(4)))
        b["result"] = _empty((h0, w0, 2))
        self._bufs = b
        return b

    def _enqueue_pyramids(self, s, b):
        """castAndScale of both views, de-rotation of the other view, both pyramids (Python/vfs.py:121-157)."""
        for d, src in ((b["ref0"], b["img_ref"]), (b["other0"], b["img_other"])):
            H.call("vfs_cast_and_scale", s, H.ptr(d), H.ptr(src), d.shape[0], d.shape[1])
        self._image_pyramid_into(s, b["refs"])
        o0 = b["others"][0]
        H.call("vfs_warp_image", s, H.ptr(o0), H.ptr(self.calibration), o0.shape[0], o0.shape[1], H.ptr(b["other0"]),
               b["other0"].shape[0], b["other0"].shape[1])
        self._image_pyramid_into(s, b["others"])

    def _enqueue_level_init(self, s, b, k):
        """State at the start of level k (1-based): zeros at the coarsest, else upsampled (Python/vfs.py:466-469,523-528)."""
        p = self.param
        sizes = self.sizes
        L = b["levels"][k - 1]
        hh, ww = sizes[k - 1]
        if k == len(sizes):
            for name in ("flow", "u", "u_bar"):
                H.call("vfs_fill", s, H.ptr(L[name]), 0.0, L[name].numel())
        else:
            P = b["levels"][k]
            ph, pw = sizes[k]
            H.call("vfs_upsample", s, H.ptr(L["u_bar"]), hh, ww, H.ptr(P["u"]), ph, pw, 1, float(p.fScale))
            H.call("vfs_copy", s, H.ptr(L["u"]), H.ptr(L["u_bar"]), L["u"].numel())
            H.call("vfs_upsample", s, H.ptr(L["flow"]), hh, ww, H.ptr(P["flow"]), ph, pw, 2, float(p.fScale))

    def _enqueue_result(self, s, b):
        """flow[mask0 == 0] = 0 into the result buffer (Python/vfs.py:531-533)."""
        L0 = b["levels"][0]
        H.call("vfs_copy", s, H.ptr(b["result"]), H.ptr(L0["flow"]), L0["flow"].numel())
        H.call("vfs_zero_masked", s, H.ptr(b["result"]), H.ptr(self.masks[0]), self.sizes[0][0], self.sizes[0][1], 2)

    def _enqueue_solve(self, s, b):
        """Enqueue the whole solve on stream ``s`` (recorded when ``s`` is a capture stream)."""
        p = self.param
        prm = self._prm()
        sizes = self.sizes
        nL = len(sizes)
        blocked = self.solver in ("blocked", "blocked_eager")
        tiles = self._tiles() if blocked else None
        self._enqueue_pyramids(s, b)
        taus, sigmas, mus = H.pd_schedule(int(p.nSolverIters), p.timeStepLambda)
        sched = [a.ctypes.data_as(ctypes.c_void_p) for a in (taus, sigmas, mus)]
        for k in range(nL, 0, -1):
            L = b["levels"][k - 1]
            hh, ww = sizes[k - 1]
            self._enqueue_level_init(s, b, k)
            lv = H.VfsLevel()
            lv.H, lv.W = hh, ww
            lv.ref = b["refs"][k - 1].data_ptr()
            lv.other = b["others"][k - 1].data_ptr()
            lv.tvf = self.translations[k - 1].data_ptr()
            lv.mask = self.masks[k - 1].data_ptr()
            for name, t in L.items():
                setattr(lv, name, t.data_ptr())
            if blocked:
                ws = H.VfsPdWs()
                for name, t in b["ws"][k - 1].items():
                    setattr(ws, name, t.data_ptr())
                tx, ty, kk, nt = tiles[k - 1]
                H.call("vfs_solve_level_blocked", s, ctypes.byref(lv), ctypes.byref(ws), ctypes.byref(prm), *sched,
                       int(p.nSolverIters), int(p.nWarpIters), tx, ty, kk, nt)
            else:
                H.call("vfs_solve_level_plain", s, ctypes.byref(lv), ctypes.byref(prm), *sched,
                       int(p.nSolverIters), int(p.nWarpIters))
        self._enqueue_result(s, b)

    def _solve_stereo_python(self, img_ref, img_other):
        """Python/vfs.py:460-533 literally: one Python call per kernel."""
        self.create_image_pyramid(img_ref, img_other)
        taus, sigmas, mus = H.pd_schedule(int(self.param.nSolverIters), self.param.timeStepLambda)
        warpUV_ref = u = u_ = None
        res = []
        for s, k in zip(self.sizes[::-1], range(len(self.sizes), 0, -1)):
            s2 = (*s, 2)
            s4 = (*s, 4)
            if s == self.sizes[-1]:
                warpUV_ref = torch.zeros(s2, dtype=torch.float32, device=_DEV)
                u = torch.zeros(s, dtype=torch.float32, device=_DEV)
                u_ = torch.zeros(s, dtype=torch.float32, device=_DEV)
            ref_smoothed = self.gaussian(self.refs_to[k - 1], self.masks_to[k - 1], s)
            abc, eta = self.compute_tensor_eta(ref_smoothed, self.masks_to[k - 1])
            for kw in range(self.param.nWarpIters):
                translation_vector_ref = self.find_warping_vector(warpUV_ref, self.translations_to[k - 1])
                other_warped = self.warping_image(self.others_to[k - 1], warpUV_ref)
                derivatives = self.compute_derivatives(other_warped, self.refs_to[k - 1], self.translations_to[k - 1])
                p = torch.zeros(s2, dtype=torch.float32, device=_DEV)
                q = torch.zeros(s4, dtype=torch.float32, device=_DEV)
                v_ = torch.zeros(s2, dtype=torch.float32, device=_DEV)
                u_before_solver = u.clone()
                for ks in range(self.param.nSolverIters):
                    p, tp, q = self.update_dual_variables(p, q, sigmas[ks], u_, v_, abc, self.masks_to[k - 1])
                    u = self.l1_thresholding(taus[ks], tp, u, u_, derivatives, eta, self.masks_to[k - 1])
                    u_, v_ = self.update_primal_variables(taus[ks], mus[ks], p, q, u, u_, v_, abc, eta,
                                                          self.masks_to[k - 1])
                if self.track_residuals:
                    res.append((k - 1, kw, self.residual_norm(u, u_before_solver, self.masks[k - 1])))
                u_filtered = self.median_filter(u, self.masks_to[k - 1])
                warpUV_ref, u = self.compute_optical_flow_with_clamp(warpUV_ref, u_before_solver, u_filtered,
                                                                     translation_vector_ref, self.masks_to[k - 1])
            if k > 1:
                ss = self.sizes[k - 2]
                u_ = self.upsampling_vector_field1(u, ss)
                u = u_.clone()
                warpUV_ref = self.upsampling_vector_field(warpUV_ref, ss)
        warpUV_ref[self.masks[0] == 0, :] = 0
        self.warpUV = warpUV_ref
        self.residuals = [(lv, w, float(r)) for lv, w, r in res]
        return warpUV_ref

    def residual_norm(self, a, b, mask=None):
        """sqrt(mean over mask pixels of |a - b|^2) as a 0-d float64 device tensor (no host
        synchronisation); a, b: (H, W[, C]) float32 on the device, mask (H, W) uint8 or None."""
        assert a.shape == b.shape and a.dtype == torch.float32 and b.dtype == torch.float32
        a, b = a.contiguous(), b.contiguous()
        n = a.shape[0] * a.shape[1]
        ch = 1 if a.dim() == 2 else a.shape[2]
        nparts = max(1, min(1024, (n + 255) // 256))
        part = torch.empty(2 * nparts, dtype=torch.float64, device=_DEV)
        out = torch.empty(2, dtype=torch.float64, device=_DEV)
        m = None
        if mask is not None:
            m = mask.tensor if isinstance(mask, texture.SampledImage) else mask
            assert m.dtype == torch.uint8 and tuple(m.shape) == tuple(a.shape[:2])
            m = m.contiguous()
        H.call("vfs_residual_norm", H.stream(), H.ptr(out), H.ptr(part), nparts, H.ptr(a), H.ptr(b),
               H.ptr(m) if m is not None else None, n, ch)
        return torch.sqrt(out[0] / torch.clamp(out[1], min=1.0))

    # -- triangulation (Python/vfs.py:535-559) -------------------------------------------
    def triangulate_flow(self, camera, translation):
        """Returns xyz (H, W, 3) in the reference camera frame; z is the depth."""
        from vector_field import get_grid_zero
        camera.setup_module()
        rays = camera.get_rays(tuple(self.warpUV.shape[:2]))
        uv = get_grid_zero(tuple(self.warpUV.shape[:2]))
        rays_other = camera.get_XYZs((uv + self.warpUV).contiguous())
        assert rays.is_contiguous() and rays_other.is_contiguous()
        assert rays.dtype == torch.float32 and rays_other.dtype == torch.float32
        # the translation's device copy is kept per value: a copy from host memory makes
        # torch wait for the stream, which would stall the host behind the solve
        tkey = tuple(float(x) for x in np.asarray(translation, dtype=np.float32).reshape(-1))
        if self._t_dev is None or self._t_dev[0] != tkey:
            self._t_dev = (tkey, _f32_device(np.asarray(tkey, dtype=np.float32)))
        translation_ = self._t_dev[1]
        d = torch.empty_like(rays)
        m = self.masks_to[0].require(texture.cudaAddressModeMirror)
        H.call("vfs_triangulate_rays", H.stream(), H.ptr(d), H.ptr(rays), H.ptr(rays_other), H.ptr(translation_),
               H.ptr(m), d.shape[0], d.shape[1])
        return d
