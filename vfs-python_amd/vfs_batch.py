"""Many stereo pairs of uneven size in one solve.

``VFSBatch`` solves N pairs that share the solver parameters (lambda, alpha0,
alpha1, beta, gamma, pyramid ratio and depth, warp and primal-dual counts) but
may differ in image size, mask, vector fields and pose.  The pyramids are built
per pair.  At every pyramid level, all pairs go through the batched kernels of
``vfs_solve_level_blocked_batch`` together: one derivatives launch, the
temporally blocked primal-dual launches over the tiles of all pairs, and one
median / clamp / next-warp launch per warp.  The whole batch solve is recorded
into one HIP graph and replayed per call.

Each pair's result is identical, bit for bit, to a single-pair
``VFS.solve_stereo`` with ``solver="blocked"`` and the same tiles.
"""

import copy
import ctypes

import numpy as np
import torch

import autotune
import texture
import vfs as V
import vfs_hip as H


class VFSBatch:
    def __init__(self, param, sizes):
        """param: a vfs_params-like object (its width/height are ignored); sizes: list of (H, W)."""
        self.param = param
        self.pair_sizes = [tuple(s) for s in sizes]
        self.pd_tiles = None
        self.vfs = []
        for (h, w) in self.pair_sizes:
            p = copy.copy(param)
            p.height, p.width = h, w
            v = V.VFS(p)
            v.solver = "blocked"
            v.setup_module()
            v.compute_pyramid_sizes()
            self.vfs.append(v)
        self._plan = None
        self._plan_key = None
        self._ws = None

    def setup_pair(self, i, mask, translation, calibration):
        """Mask and vector-field pyramids of pair i (Python/vfs.py:159-218)."""
        v = self.vfs[i]
        v.create_mask_pyramid(mask)
        v.create_vector_field_pyramid(translation, calibration)
        self._plan_key = None

    def _tiles(self):
        """One (tile_x, tile_y, k, threads) per pyramid level for all pairs.  A batched launch
        covers the tiles of every pair, so the choice is the tuned one for a level of the
        same total pixel count."""
        n_level = len(self.vfs[0].sizes)
        out = []
        for lvl in range(n_level):
            if self.pd_tiles is not None and lvl in self.pd_tiles:
                out.append(autotune.variant(self.pd_tiles[lvl]))
                continue
            total = sum(v.sizes[lvl][0] * v.sizes[lvl][1] for v in self.vfs)
            out.append(autotune.tile_for_pixels(total, int(self.param.nSolverIters)))
        return out

    def solve(self, img_refs, img_others):
        """Solve every pair; returns the list of flows (H_i, W_i, 2) (views of persistent buffers)."""
        assert len(img_refs) == len(img_others) == len(self.vfs)
        bufs = [v._buffers() for v in self.vfs]
        for b, r, o in zip(bufs, img_refs, img_others):
            b["img_ref"].copy_(V._u8_device(r))
            b["img_other"].copy_(V._u8_device(o))
        key = (tuple(v._key(b) for v, b in zip(self.vfs, bufs)), tuple(self._tiles()))
        if self._plan is None or self._plan_key != key:
            self.release_plan()
            self._prepare(bufs)
            cs = ctypes.c_void_p()
            H.call("vfs_capture_begin", ctypes.byref(cs))
            try:
                self._enqueue(cs, bufs)
            finally:
                plan = ctypes.c_void_p()
                H.call("vfs_capture_end", cs, ctypes.byref(plan))
            self._plan, self._plan_key = plan, key
        H.call("vfs_plan_launch", self._plan, H.stream())
        for v, b in zip(self.vfs, bufs):
            v.refs, v.others = b["refs"], b["others"]
            v.refs_to = [texture.wrap(x) for x in b["refs"]]
            v.others_to = [texture.wrap(x) for x in b["others"]]
            v.warpUV = b["result"]
        return [v.warpUV for v in self.vfs]

    def triangulate(self, cameras, translations):
        return [v.triangulate_flow(c, t) for v, c, t in zip(self.vfs, cameras, translations)]

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

    def _level_structs(self, bufs, lvl):
        levels = (H.VfsLevel * len(self.vfs))()
        wss = (H.VfsPdWs * len(self.vfs))()
        for i, (v, b) in enumerate(zip(self.vfs, bufs)):
            L = b["levels"][lvl]
            lv = levels[i]
            lv.H, lv.W = v.sizes[lvl]
            lv.ref = b["refs"][lvl].data_ptr()
            lv.other = b["others"][lvl].data_ptr()
            lv.tvf = v.translations[lvl].data_ptr()
            lv.mask = v.masks[lvl].data_ptr()
            for name, t in L.items():
                setattr(lv, name, t.data_ptr())
            for name, t in b["ws"][lvl].items():
                setattr(wss[i], name, t.data_ptr())
        return levels, wss

    def _prepare(self, bufs):
        """Descriptor tables of every level, uploaded once before the graph is recorded."""
        n = len(self.vfs)
        nbytes = H.lib().vfs_batch_workspace_bytes(n)
        tiles = self._tiles()
        self._ws = []
        self._structs = []
        for lvl in range(len(self.vfs[0].sizes)):
            ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
            levels, wss = self._level_structs(bufs, lvl)
            tx, ty, k, nt = tiles[lvl]
            torch.cuda.synchronize()
            H.call("vfs_batch_prepare_level", H.ptr(ws), nbytes, ctypes.cast(levels, ctypes.c_void_p),
                   ctypes.cast(wss, ctypes.c_void_p), n, tx, ty, k, nt)
            self._ws.append(ws)
            self._structs.append((levels, wss))

    def _enqueue(self, s, bufs):
        p = self.param
        prm = H.make_params(p, self.vfs[0].median_filter_kernel_size)
        n = len(self.vfs)
        n_level = len(self.vfs[0].sizes)
        tiles = self._tiles()
        taus, sigmas, mus = H.pd_schedule(int(p.nSolverIters), p.timeStepLambda)
        sched = [a.ctypes.data_as(ctypes.c_void_p) for a in (taus, sigmas, mus)]
        for v, b in zip(self.vfs, bufs):
            v._enqueue_pyramids(s, b)
        for k in range(n_level, 0, -1):
            for v, b in zip(self.vfs, bufs):
                v._enqueue_level_init(s, b, k)
            levels, wss = self._structs[k - 1]
            tx, ty, kk, nt = tiles[k - 1]
            H.call("vfs_solve_level_blocked_batch", s, H.ptr(self._ws[k - 1]), ctypes.cast(levels, ctypes.c_void_p),
                   ctypes.cast(wss, ctypes.c_void_p), n, ctypes.byref(prm), *sched, int(p.nSolverIters),
                   int(p.nWarpIters), tx, ty, kk, nt)
        for v, b in zip(self.vfs, bufs):
            v._enqueue_result(s, b)
