"""Gradients of a depth loss through the whole VFS solve (hand-written adjoint).

``VFSDepthAdjoint`` runs the coarse-to-fine solve of one pair (Python/vfs.py:460-533,
the reference's kernel decomposition) with the native level driver
``vfs_adj_forward_level``, which also records the warp-start state (u, u_bar,
flow) of every warp and a checkpoint of the primal-dual state every
``checkpoint`` iterations.  ``backward`` walks the levels from fine to coarse with
``vfs_adj_reverse_level``: per warp, in reverse, it recomputes the warp's inputs,
recomputes each checkpoint segment into a tape and runs the reverse kernels of
csrc/vfs_adjoint.hip (primal-dual iterations, median, clamp, derivatives, warp),
then the level's tensor and Gaussian.  The level transitions (upsampling) and the
image pyramid are reversed here.

The loss is L = mean over valid mask pixels of |log z - log z_gt|, with z the
depth that triangulate_flow (Python/vfs.py:535-559) returns.  Gradients are with
respect to lambda, alpha0, alpha1, beta (the diffusion weight "a"), gamma (the
exponent "b") and the level-0 reference image (u8 / 256 as float32).

``DepthLoss`` exposes this as a torch.autograd.Function of (ref0, theta).

Memory: per level, n_warp x (4 + 10 ceil(n_iter / checkpoint)) floats per pixel
of saves, plus one workspace of (43 + 10 checkpoint) floats per pixel of level 0
(800 x 800, 10 warps x 50 iterations, checkpoint 25: 0.61 GB of saves for level
0 and 0.75 GB of workspace), plus, by default (``tape_last_segment``), the tape of
each reversed warp's last checkpoint segment, 10 floats per pixel per iteration of
that segment (6.4 GB for level 0 of that workload, 8.4 GB for its three levels):
the forward writes it while computing those states, and the reverse recomputes
only the earlier segments.
"""

import copy
import ctypes
import threading

import numpy as np
import torch

import vfs_hip as H

PARAMS = ("Lambda", "alpha0", "alpha1", "beta", "gamma")

# temporally blocked reverse sweeps: mode -> (tile_x, tile_y, units per launch, threads)
# (kRevVariants of csrc/vfs_adjoint_solver.hip; 41-43 are the lean forms of 31-33)
_REV_TILES = {21: (30, 30, 1, 1024), 22: (28, 28, 2, 1024), 23: (26, 26, 3, 1024), 24: (24, 24, 4, 1024),
              25: (22, 22, 5, 1024), 26: (20, 20, 6, 1024), 31: (30, 14, 1, 512), 32: (28, 12, 2, 512),
              33: (26, 10, 3, 512), 41: (30, 14, 1, 512), 42: (28, 12, 2, 512), 43: (26, 10, 3, 512)}


def _z(shape, ch=None, dtype=torch.float32):
    s = tuple(shape) if ch is None else (*shape, ch)
    return torch.zeros(s, dtype=dtype, device="cuda")


def _e(shape, ch=None):
    s = tuple(shape) if ch is None else (*shape, ch)
    return torch.empty(s, dtype=torch.float32, device="cuda")


class VFSDepthAdjoint:
    def __init__(self, param, mask, tvf, cvf, cam, t, z_gt, checkpoint=25, fused=True, rows=None, blocked=True,
                 rec_tile="auto", workspace=None, reverse_warps=0, reverse_block="auto", tape_last_segment=True):
        import vfs as V
        self.param = copy.copy(param)
        self.checkpoint = int(checkpoint)
        self.fused = bool(fused)
        if rows not in (None, 1, 2, 4):
            raise ValueError("rows must be None, 1, 2 or 4")
        # reverse sweep: 1 = one kernel per step; 0 = fused, rows per thread by level size;
        # 10 + rows = fused with that many rows per thread
        self._mode = 1 if not self.fused else (0 if rows is None else 10 + rows)
        self._auto_block = False
        # temporally blocked sweep: reverse_block = K (1-6) units per launch on a 32 x 32
        # extended tile (1024 threads; sweep 20 + K), or (K, 512) for 1-3 units on 32 x 16;
        # a dict {level: choice} sets it per pyramid level (missing levels: fused); (K, 512,
        # "lean") is the 512-thread tile with one forward-state buffer in LDS and no tape
        # prefetch, three workgroups per CU instead of two; "auto" picks by level size (the
        # lean three-unit 512-thread tile at >= 400 k pixels, three units on the 1024-thread
        # tile at >= 100 k, six below: measured at the 800 x 800 gradient workload)
        self._level_mode = {}
        if reverse_block == "auto":
            reverse_block = None if (not self.fused or rows is not None) else "auto"
        if reverse_block == "auto":
            self._auto_block = True
        elif reverse_block is not None:
            if not self.fused or rows is not None:
                raise ValueError("reverse_block replaces the fused sweep's middle launches; keep fused=True, rows=None")
            modes = {1: 21, 2: 22, 3: 23, 4: 24, 5: 25, 6: 26, (1, 512): 31, (2, 512): 32, (3, 512): 33,
                     (1, 512, "lean"): 41, (2, 512, "lean"): 42, (3, 512, "lean"): 43}
            per = reverse_block if isinstance(reverse_block, dict) else None
            for lvl, rb in (per.items() if per is not None else [(None, reverse_block)]):
                key = rb if isinstance(rb, int) else tuple(x if isinstance(x, str) else int(x) for x in rb)
                if key not in modes:
                    raise ValueError(f"reverse_block must be one of {list(modes)}")
                if lvl is None:
                    self._mode = modes[key]
                else:
                    self._level_mode[int(lvl)] = modes[key]
        # forward: blocked primal-dual launches (variant per level from the autotune table)
        # or the per-kernel steps; tape recompute: blocked launches that also write
        # their intermediate states, with the forward's variant ("auto") or a given
        # (tile_x, tile_y, k[, threads]), or the per-kernel steps (None)
        self.blocked = bool(blocked)
        self.rec_tile = rec_tile if rec_tile in (None, "auto") else tuple(int(x) for x in rec_tile)
        # 0: exact gradient; r > 0: reverse only the last r warps of each level with their
        # start state held fixed (so only level 0 reaches the loss)
        self.reverse_warps = int(reverse_warps)
        # the forward keeps the tape of each reversed warp's last checkpoint segment (it
        # computes those states anyway), so the reverse recomputes only the earlier
        # segments: n_iter - ckpt more states per warp in HBM, the same gradient bits
        # (blocked forward only)
        self.tape_last_segment = bool(tape_last_segment)
        if self.checkpoint < 1:
            raise ValueError("checkpoint must be >= 1")
        self.v = V.VFS(self.param)
        self.v.setup_module()
        self.v.compute_pyramid_sizes()
        self.v.create_mask_pyramid(mask)
        self.v.create_vector_field_pyramid(tvf, cvf)
        self.sizes = self.v.sizes
        self.cam = tuple(float(x) for x in cam)
        self.t = np.asarray(t, dtype=np.float64).reshape(3)
        self.z_gt = torch.as_tensor(np.asarray(z_gt, dtype=np.float32), device="cuda").contiguous()
        self.s = None
        self._b = None
        self._done_forward = False
        # forwards run so far: DepthLoss checks that a loss's backward reads the state
        # of its own forward
        self.generation = 0
        # optional shared workspace: a dict that instances solved one after another on the
        # same stream may share ({"ws": uint8 device tensor}, grown on demand)
        self._shared_ws = workspace

    # -- helpers ------------------------------------------------------------------------------
    def _prm(self):
        return H.make_params(self.param, self.v.median_filter_kernel_size)

    def set_params(self, theta):
        for name, val in zip(PARAMS, theta):
            setattr(self.param, name, float(val))
        self._done_forward = False

    def _schedule(self):
        p = self.param
        n_iter, n_warp = int(p.nSolverIters), int(p.nWarpIters)
        taus, sigmas, mus = (np.ascontiguousarray(a, dtype=np.float32) for a in H.pd_schedule(n_iter, p.timeStepLambda))
        return n_iter, n_warp, taus, sigmas, mus

    def _tiles(self, k):
        import autotune
        t = H.VfsAdjOpts()
        hh, ww = self.sizes[k]
        best = autotune.tile_for(hh, ww, int(self.param.nSolverIters))
        if self.blocked:
            t.fwd_tile_x, t.fwd_tile_y, t.fwd_k, t.fwd_threads = best
        if self.rec_tile is not None:
            rec = best if self.rec_tile == "auto" else autotune.variant(self.rec_tile)
            t.rec_tile_x, t.rec_tile_y, t.rec_k, t.rec_threads = rec
        if self._auto_block:
            npx = hh * ww
            t.sweep = 43 if npx >= 400_000 else (23 if npx >= 100_000 else 26)
        else:
            t.sweep = self._level_mode.get(k, self._mode)
        t.reverse_warps = self.reverse_warps
        tw = self._tape_warps(k)
        if tw > 0:
            t.tape_warps = tw
            t.tape_save = self._b["tape"][k].data_ptr()
        return t

    def _tape_warps(self, k):
        """Warps of level k whose last segment's tape the forward keeps: every warp for the
        exact gradient, the reversed ones of the finest level for the truncated one."""
        if not (self.tape_last_segment and self.blocked):
            return 0
        n_warp = int(self.param.nWarpIters)
        if self.reverse_warps > 0:
            return min(self.reverse_warps, n_warp) if k == 0 else 0
        return n_warp

    def _buffers(self):
        """Device buffers for one pair, allocated once (shapes are fixed by the pyramid)."""
        p = self.param
        n_iter, n_warp = int(p.nSolverIters), int(p.nWarpIters)
        key = (n_iter, n_warp, self.checkpoint, tuple(self._tape_warps(k) for k in range(len(self.sizes))))
        if self._b is not None and self._b["key"] == key:
            return self._b
        L = H.lib()
        sizes = self.sizes
        b = dict(key=key, refs=[_e(sz) for sz in sizes], others=[_e(sz) for sz in sizes], level=[], save=[],
                 gu=[], gub=[], gf=[], gref=[], gpar=[], other0=None, tape=[])
        for k, sz in enumerate(sizes):
            lv = dict(flow=_z(sz, 2), u=_z(sz), u_bar=_z(sz), ref_smoothed=_e(sz), abc=_e(sz, 3), eta=_e(sz, 3),
                      tv=_e(sz, 2), warped=_e(sz), deriv=_e(sz, 2), p=_e(sz, 2), q=_e(sz, 4), tp=_e(sz, 2),
                      v_bar=_e(sz, 2), u0=_e(sz), uf=_e(sz))
            st = H.VfsLevel()
            st.H, st.W = sz
            st.ref = b["refs"][k].data_ptr()
            st.other = b["others"][k].data_ptr()
            st.tvf = self.v.translations[k].data_ptr()
            st.mask = self.v.masks[k].data_ptr()
            for name, t in lv.items():
                setattr(st, name, t.data_ptr())
            b["level"].append((lv, st))
            nbytes = L.vfs_adj_save_bytes(sz[0], sz[1], n_warp, n_iter, self.checkpoint)
            b["save"].append(torch.empty(max(nbytes, 256), dtype=torch.uint8, device="cuda"))
            tb = L.vfs_adj_tape_bytes(sz[0], sz[1], n_iter, self.checkpoint, self._tape_warps(k))
            b["tape"].append(torch.empty(tb, dtype=torch.uint8, device="cuda") if tb > 0 else None)
            b["gu"].append(_z(sz))
            b["gub"].append(_z(sz))
            b["gf"].append(_z(sz, 2))
            b["gref"].append(_z(sz))
            b["gpar"].append(_z(sz, 5))
        ws = max(L.vfs_adj_workspace_bytes(sz[0], sz[1], self.checkpoint) for sz in sizes)
        b["ws_bytes"] = ws
        b["ws"] = None if self._shared_ws is not None else torch.empty(ws, dtype=torch.uint8, device="cuda")
        self._b = b
        return b

    def _ws(self):
        """The reverse workspace: this pair's own, or the shared one (grown to this pair's size)."""
        if self._shared_ws is None:
            return self._b["ws"]
        cur = self._shared_ws.get("ws")
        if cur is None or cur.numel() < self._b["ws_bytes"]:
            cur = self._shared_ws["ws"] = torch.empty(self._b["ws_bytes"], dtype=torch.uint8, device="cuda")
        return cur

    def set_workspace(self, workspace):
        """Share the reverse workspace dict `workspace` from now on (None: a workspace of its own)."""
        self._shared_ws = workspace
        if self._b is not None and workspace is None and self._b["ws"] is None:
            self._b["ws"] = torch.empty(self._b["ws_bytes"], dtype=torch.uint8, device="cuda")

    def _reduce(self, x, channels):
        """Device (channels,) float64 sums of an interleaved float32 array."""
        n = x.numel() // channels
        nparts = max(1, min(1024, (n + 255) // 256))
        part = torch.empty(nparts * channels, dtype=torch.float64, device="cuda")
        out = torch.empty(channels, dtype=torch.float64, device="cuda")
        H.call("vfs_reduce_sum", self.s, H.ptr(out), H.ptr(part), nparts, H.ptr(x), n, channels)
        return out

    # -- forward ------------------------------------------------------------------------------
    def forward(self, ref0, img_other):
        """ref0: (H, W) float32 level-0 reference image (u8 / 256); img_other: the other view (uint8).
        Returns the loss (float) and keeps the state the reverse sweep needs."""
        import vfs as V
        self.s = H.stream()
        s = self.s
        p = self.param
        prm = self._prm()
        sizes = self.sizes
        nL = len(sizes)
        n_iter, n_warp, taus, sigmas, mus = self._schedule()
        b = self._buffers()
        ref0 = torch.as_tensor(ref0, device="cuda").to(torch.float32)
        if tuple(ref0.shape) != tuple(sizes[0]):
            raise ValueError(f"ref0 has shape {tuple(ref0.shape)}, expected {tuple(sizes[0])}")
        b["refs"][0].copy_(ref0)
        self.v._image_pyramid_into(s, b["refs"])
        o8 = V._u8_device(img_other)
        if b["other0"] is None or b["other0"].shape != o8.shape:
            b["other0"] = _e(o8.shape)
        other0 = b["other0"]
        H.call("vfs_cast_and_scale", s, H.ptr(other0), H.ptr(o8), other0.shape[0], other0.shape[1])
        H.call("vfs_warp_image", s, H.ptr(b["others"][0]), H.ptr(self.v.calibration), sizes[0][0], sizes[0][1],
               H.ptr(other0), other0.shape[0], other0.shape[1])
        self.v._image_pyramid_into(s, b["others"])
        tp, sp, mp = taus.ctypes.data, sigmas.ctypes.data, mus.ctypes.data
        for k in range(nL - 1, -1, -1):
            hh, ww = sizes[k]
            lv, st = b["level"][k]
            if k == nL - 1:
                lv["flow"].zero_()
                lv["u"].zero_()
                lv["u_bar"].zero_()
            else:
                up, ch = b["level"][k + 1][0], sizes[k + 1]
                H.call("vfs_upsample", s, H.ptr(lv["u"]), hh, ww, H.ptr(up["u"]), ch[0], ch[1], 1, float(p.fScale))
                H.call("vfs_copy", s, H.ptr(lv["u_bar"]), H.ptr(lv["u"]), hh * ww)
                H.call("vfs_upsample", s, H.ptr(lv["flow"]), hh, ww, H.ptr(up["flow"]), ch[0], ch[1], 2,
                       float(p.fScale))
            H.call("vfs_adj_forward_level", s, ctypes.byref(st), ctypes.byref(prm), tp, sp, mp, n_iter, n_warp,
                   self.checkpoint, H.ptr(b["save"][k]), H.ptr(self._ws()), ctypes.byref(self._tiles(k)))
        lv0 = b["level"][0][0]
        self.flow = lv0["flow"].clone()
        H.call("vfs_zero_masked", s, H.ptr(self.flow), H.ptr(self.v.masks[0]), sizes[0][0], sizes[0][1], 2)
        h0, w0 = sizes[0]
        term = _e(sizes[0], 2)
        self.gflow = _e(sizes[0], 2)
        f, u0, v0 = self.cam
        H.call("vfs_depth_loss", s, H.ptr(term), H.ptr(self.gflow), H.ptr(self.flow), H.ptr(self.z_gt),
               H.ptr(self.v.masks[0]), h0, w0, f, u0, v0, float(self.t[0]), float(self.t[2]))
        sums = self._reduce(term, 2).cpu().numpy()
        self.n_valid = max(float(sums[1]), 1.0)
        self.loss = float(sums[0]) / self.n_valid
        self._done_forward = True
        self.generation += 1
        return self.loss

    # -- reverse --------------------------------------------------------------------------------
    def backward(self, probe=None):
        """Returns dict(ref0=(H, W) tensor, Lambda=..., alpha0=..., alpha1=..., beta=..., gamma=...).

        probe: optional dict(level=k, launches=n), a timing hook: the first n temporally
        blocked reverse launches of level k are bracketed by HIP events on the stream; on
        return probe["us"] holds their durations (microseconds), probe["units"] their unit
        counts and probe["tiles"] the launch variant (tile_x, tile_y, units, threads)."""
        if not self._done_forward:
            raise RuntimeError("backward() needs a forward() with the current parameters first")
        s = self.s
        p = self.param
        prm = self._prm()
        sizes = self.sizes
        nL = len(sizes)
        n_iter, n_warp, taus, sigmas, mus = self._schedule()
        b = self._b
        for k in range(nL):
            for name in ("gu", "gub", "gf", "gref", "gpar"):
                b[name][k].zero_()
        torch.div(self.gflow, self.n_valid, out=b["gf"][0])
        tp, sp, mp = taus.ctypes.data, sigmas.ctypes.data, mus.ctypes.data
        # truncated: the finest level's start state is held fixed, so nothing reaches the
        # coarser levels (their reverse would run on zero adjoints)
        n_rev = 1 if self.reverse_warps > 0 else nL
        held = None
        for k in range(n_rev):
            hh, ww = sizes[k]
            lv, st = b["level"][k]
            opts = self._tiles(k)
            if probe is not None and int(probe.get("level", 0)) == k:
                held = self._probe_attach(opts, int(probe["launches"]))
            H.call("vfs_adj_reverse_level", s, ctypes.byref(st), ctypes.byref(prm), tp, sp, mp, n_iter, n_warp,
                   self.checkpoint, H.ptr(b["save"][k]), H.ptr(self._ws()), H.ptr(b["gu"][k]), H.ptr(b["gub"][k]),
                   H.ptr(b["gf"][k]), H.ptr(b["gref"][k]), H.ptr(b["gpar"][k]), ctypes.byref(opts))
            # level start: u = u_bar = fScale up(u_coarse), flow = fScale up(flow_coarse); the
            # coarse level's final u_bar is discarded, so its adjoint stays zero
            if k < n_rev - 1:
                ch, cw = sizes[k + 1]
                b["gu"][k].add_(b["gub"][k])
                H.call("vfs_adj_upsample", s, H.ptr(b["gu"][k + 1]), ch, cw, H.ptr(b["gu"][k]), hh, ww, 1,
                       float(p.fScale))
                H.call("vfs_adj_upsample", s, H.ptr(b["gf"][k + 1]), ch, cw, H.ptr(b["gf"][k]), hh, ww, 2,
                       float(p.fScale))
        # image pyramid: refs[k] = down(refs[k - 1])
        for k in range(nL - 1, 0, -1):
            H.call("vfs_adj_downsample", s, H.ptr(b["gref"][k - 1]), *sizes[k - 1], H.ptr(b["gref"][k]), *sizes[k])
        tot = sum(self._reduce(g, 5) for g in b["gpar"]).cpu().numpy()
        out = {name: float(tot[i]) for i, name in enumerate(PARAMS)}
        out["ref0"] = b["gref"][0].clone()
        if held is not None:
            torch.cuda.synchronize()
            evs, units, count, (opts, _) = held
            n = int(count.value)
            probe["us"] = [1e3 * evs[2 * j].elapsed_time(evs[2 * j + 1]) for j in range(n)]
            probe["units"] = [int(units[j]) for j in range(n)]
            probe["tiles"] = _REV_TILES.get(int(opts.sweep))
        return out

    @staticmethod
    def _probe_attach(opts, n):
        """Fills the timing-probe fields of opts with 2 n timing events (recorded once on
        the current stream so that their HIP handles exist) and host arrays; returns what
        must stay alive until the events are read."""
        evs = [torch.cuda.Event(enable_timing=True) for _ in range(2 * n)]
        for e in evs:
            e.record()
        handles = (ctypes.c_void_p * (2 * n))(*[ctypes.c_void_p(e.cuda_event) for e in evs])
        units = (ctypes.c_int32 * n)()
        count = ctypes.c_int32(0)
        opts.probe_events = ctypes.cast(handles, ctypes.POINTER(ctypes.c_void_p))
        opts.probe_units = ctypes.cast(units, ctypes.POINTER(ctypes.c_int32))
        opts.probe_count = ctypes.pointer(count)
        opts.probe_max = n
        return evs, units, count, (opts, handles)


class DepthLoss(torch.autograd.Function):
    """loss = DepthLoss.apply(ref0, theta, adj, img_other); theta = (lambda, alpha0, alpha1, beta, gamma).

    The reverse sweep reads the state that ``adj`` keeps from its latest forward, so a
    loss can be differentiated only while it is that latest forward of ``adj``: a
    second DepthLoss.apply on the same ``adj`` before the first loss's backward makes
    that backward raise instead of returning the second loss's gradient.  Use one
    VFSDepthAdjoint per loss that is kept for backward."""

    @staticmethod
    def forward(ctx, ref0, theta, adj, img_other):
        adj.set_params(theta.detach().cpu().numpy())
        loss = adj.forward(ref0.detach(), img_other)
        ctx.adj = adj
        ctx.generation = adj.generation
        ctx.theta_meta = (theta.device, theta.dtype)
        ctx.ref0_meta = (ref0.device, ref0.dtype)
        return torch.tensor(loss, dtype=torch.float64, device=theta.device)

    @staticmethod
    def backward(ctx, gl):
        if ctx.adj.generation != ctx.generation:
            raise RuntimeError("DepthLoss.backward: the VFSDepthAdjoint ran another forward after this loss; "
                               "its saved state belongs to that forward")
        g = ctx.adj.backward()
        gl = float(gl)
        dev, dt = ctx.theta_meta
        gtheta = torch.tensor([g[n] * gl for n in PARAMS], dtype=dt, device=dev)
        rdev, rdt = ctx.ref0_meta
        return (g["ref0"] * gl).to(device=rdev, dtype=rdt), gtheta, None, None


_streams = []


def gradients(items, streams=2, workspaces=None):
    """Loss and gradients of several pairs, run concurrently on `streams` HIP streams.

    items: list of (adj, ref0, img_other), adj a VFSDepthAdjoint.  Pair i runs forward()
    then backward() on stream i mod streams, one host thread per stream, so that the
    launches of different pairs overlap on the GPU (their coarse levels and per-warp
    kernels do not fill it alone).  A pair's arithmetic does not depend on the stream:
    the gradients are the same bits as forward(); backward() one pair after another.
    Pairs on different streams must not share a reverse workspace: with `workspaces` (a
    list of `streams` workspace dicts) each pair uses its stream's dict (set_workspace);
    without it every pair keeps the workspace it was built with, which must then be its
    own or shared only within one stream.  Returns [(loss, gradient dict)] in the order
    of items; the streams are synchronised on return.  A worker's exception is re-raised."""
    streams = max(1, int(streams))
    if workspaces is not None and len(workspaces) != streams:
        raise ValueError("gradients: one workspace dict per stream")
    while len(_streams) < streams:
        _streams.append(torch.cuda.Stream())
    out = [None] * len(items)
    errors = []
    caller = torch.cuda.current_stream()  # the inputs are ready in the caller's stream order

    def worker(k):
        try:
            st = _streams[k]
            st.wait_stream(caller)
            with torch.cuda.stream(st):
                for i in range(k, len(items), streams):
                    adj, ref0, other = items[i]
                    if workspaces is not None:
                        adj.set_workspace(workspaces[k])
                    loss = adj.forward(ref0, other)
                    out[i] = (loss, adj.backward())
                st.synchronize()
        except BaseException as e:  # re-raised by the caller's thread
            errors.append(e)

    if streams == 1:
        worker(0)
    else:
        threads = [threading.Thread(target=worker, args=(k,)) for k in range(streams)]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
    if errors:
        raise errors[0]
    for k in range(streams):
        caller.wait_stream(_streams[k])
    for _, g in out:
        g["ref0"].record_stream(caller)
    return out
