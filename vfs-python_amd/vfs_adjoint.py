"""Gradients of a depth loss through the whole VFS solve (hand-written adjoint).

``VFSDepthAdjoint`` runs the coarse-to-fine solve of one pair (Python/vfs.py:460-533,
the reference's kernel decomposition) and keeps what the reverse sweep needs:
the state (u, u_bar, flow) at every warp start and a checkpoint of the
primal-dual state every ``checkpoint`` iterations.  ``backward`` walks the
levels from fine to coarse and the warps in reverse; for each warp it recomputes
the warp's inputs, recomputes each checkpoint segment into a tape, and runs the
reverse kernels of csrc/vfs_adjoint.hip (primal-dual iterations, median, clamp,
derivatives, warp), then the level's tensor and Gaussian, the level transition,
and finally the image pyramid.

The loss is L = mean over valid mask pixels of |log z - log z_gt|, with z the
depth that triangulate_flow (Python/vfs.py:535-559) returns.  Gradients are with
respect to lambda, alpha0, alpha1, beta (the diffusion weight "a"), gamma (the
exponent "b") and the level-0 reference image (u8 / 256 as float32).

``DepthLoss`` exposes this as a torch.autograd.Function of (ref0, theta).
"""

import ctypes
import copy

import numpy as np
import torch

import vfs_hip as H

PARAMS = ("Lambda", "alpha0", "alpha1", "beta", "gamma")


def _z(shape, ch=None, dtype=torch.float32):
    s = tuple(shape) if ch is None else (*shape, ch)
    return torch.zeros(s, dtype=dtype, device="cuda")


def _e(shape, ch=None):
    s = tuple(shape) if ch is None else (*shape, ch)
    return torch.empty(s, dtype=torch.float32, device="cuda")


class VFSDepthAdjoint:
    def __init__(self, param, mask, tvf, cvf, cam, t, z_gt, checkpoint=25):
        import vfs as V
        self.param = copy.copy(param)
        self.checkpoint = int(checkpoint)
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

    # -- helpers ------------------------------------------------------------------------------
    def _prm(self):
        return H.make_params(self.param, self.v.median_filter_kernel_size)

    def set_params(self, theta):
        for name, val in zip(PARAMS, theta):
            setattr(self.param, name, float(val))

    def _warp_inputs(self, k, flow, tv, ow, deriv):
        s = self.s
        hh, ww = self.sizes[k]
        H.call("vfs_find_warping_vector", s, H.ptr(tv), H.ptr(flow), hh, ww, H.ptr(self.v.translations[k]), hh, ww)
        H.call("vfs_warp_image", s, H.ptr(ow), H.ptr(flow), hh, ww, H.ptr(self.others[k]), hh, ww)
        H.call("vfs_compute_derivatives", s, H.ptr(deriv), H.ptr(self.refs[k]), H.ptr(ow),
               H.ptr(self.v.translations[k]), hh, ww)

    def _pd_iter(self, k, st, it, prm, tp):
        """One forward primal-dual iteration in place on st = dict(p, q, u, ub, vb)."""
        s = self.s
        hh, ww = self.sizes[k]
        m = self.v.masks[k]
        L = self.lv[k]
        H.call("vfs_update_dual", s, H.ptr(st["p"]), H.ptr(tp), H.ptr(st["q"]), float(self.sigmas[it]),
               H.ptr(st["ub"]), H.ptr(st["vb"]), H.ptr(L["abc"]), H.ptr(m), hh, ww, ctypes.byref(prm))
        H.call("vfs_l1_thresholding", s, H.ptr(st["u"]), float(self.taus[it]), H.ptr(tp), H.ptr(st["u"]),
               H.ptr(st["ub"]), H.ptr(self.deriv), H.ptr(L["eta"]), H.ptr(m), hh, ww, ctypes.byref(prm))
        H.call("vfs_update_primal", s, H.ptr(st["ub"]), H.ptr(st["vb"]), float(self.taus[it]), float(self.mus[it]),
               H.ptr(st["p"]), H.ptr(st["q"]), H.ptr(st["u"]), H.ptr(st["ub"]), H.ptr(st["vb"]), H.ptr(L["abc"]),
               H.ptr(L["eta"]), H.ptr(m), hh, ww, ctypes.byref(prm))

    @staticmethod
    def _clone(st):
        return {k: v.clone() for k, v in st.items()}

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
        self.taus, self.sigmas, self.mus = H.pd_schedule(int(p.nSolverIters), p.timeStepLambda)
        ref0 = ref0.to(torch.float32).contiguous()
        self.refs = [ref0] + [_e(sz) for sz in sizes[1:]]
        self.v._image_pyramid_into(s, self.refs)
        o8 = V._u8_device(img_other)
        other0 = _e(o8.shape)
        H.call("vfs_cast_and_scale", s, H.ptr(other0), H.ptr(o8), other0.shape[0], other0.shape[1])
        self.others = [_e(sz) for sz in sizes]
        H.call("vfs_warp_image", s, H.ptr(self.others[0]), H.ptr(self.v.calibration), sizes[0][0], sizes[0][1],
               H.ptr(other0), other0.shape[0], other0.shape[1])
        self.v._image_pyramid_into(s, self.others)
        self.lv = [None] * nL
        self.saved = [None] * nL
        flow = u = ub = None
        for k in range(nL - 1, -1, -1):
            hh, ww = sizes[k]
            if k == nL - 1:
                flow, u, ub = _z(sizes[k], 2), _z(sizes[k]), _z(sizes[k])
            else:
                nu = _e(sizes[k])
                H.call("vfs_upsample", s, H.ptr(nu), hh, ww, H.ptr(u), *sizes[k + 1], 1, float(p.fScale))
                nf = _e(sizes[k], 2)
                H.call("vfs_upsample", s, H.ptr(nf), hh, ww, H.ptr(flow), *sizes[k + 1], 2, float(p.fScale))
                flow, u, ub = nf, nu, nu.clone()
            m = self.v.masks[k]
            L = dict(Is=_e(sizes[k]), abc=_e(sizes[k], 3), eta=_e(sizes[k], 3))
            H.call("vfs_gaussian_masked", s, H.ptr(L["Is"]), H.ptr(self.refs[k]), H.ptr(m), hh, ww)
            H.call("vfs_compute_tensor_eta", s, H.ptr(L["abc"]), H.ptr(L["eta"]), H.ptr(L["Is"]), H.ptr(m), hh, ww,
                   ctypes.byref(prm))
            self.lv[k] = L
            self.saved[k] = []
            tv, ow, self.deriv = _e(sizes[k], 2), _e(sizes[k]), _e(sizes[k], 2)
            tp, uf = _e(sizes[k], 2), _e(sizes[k])
            for w in range(int(p.nWarpIters)):
                rec = dict(u=u.clone(), ub=ub.clone(), flow=flow.clone(), ckpt=[])
                self._warp_inputs(k, flow, tv, ow, self.deriv)
                st = dict(p=_z(sizes[k], 2), q=_z(sizes[k], 4), u=u.clone(), ub=ub, vb=_z(sizes[k], 2))
                for it in range(int(p.nSolverIters)):
                    if it % self.checkpoint == 0:
                        rec["ckpt"].append(self._clone(st))
                    self._pd_iter(k, st, it, prm, tp)
                ub = st["ub"]
                H.call("vfs_median_filter", s, H.ptr(uf), H.ptr(st["u"]), H.ptr(m), hh, ww,
                       self.v.median_filter_kernel_size)
                H.call("vfs_optical_flow_with_clamp", s, H.ptr(flow), H.ptr(u), H.ptr(uf), H.ptr(tv), H.ptr(m), hh,
                       ww, float(p.limitRange))
                self.saved[k].append(rec)
        self.flow = flow.clone()
        H.call("vfs_zero_masked", s, H.ptr(self.flow), H.ptr(self.v.masks[0]), sizes[0][0], sizes[0][1], 2)
        h0, w0 = sizes[0]
        term, self.gflow = _e(sizes[0], 2), _e(sizes[0], 2)
        f, u0, v0 = self.cam
        H.call("vfs_depth_loss", s, H.ptr(term), H.ptr(self.gflow), H.ptr(self.flow), H.ptr(self.z_gt),
               H.ptr(self.v.masks[0]), h0, w0, f, u0, v0, float(self.t[0]), float(self.t[2]))
        sums = self._reduce(term, 2)
        self.n_valid = max(float(sums[1]), 1.0)
        self.loss = float(sums[0]) / self.n_valid
        return self.loss

    def _reduce(self, x, channels):
        n = x.numel() // channels
        nparts = max(1, min(1024, (n + 255) // 256))
        part = torch.empty(nparts * channels, dtype=torch.float64, device="cuda")
        out = torch.empty(channels, dtype=torch.float64, device="cuda")
        H.call("vfs_reduce_sum", self.s, H.ptr(out), H.ptr(part), nparts, H.ptr(x), n, channels)
        return out.cpu().numpy()

    # -- reverse --------------------------------------------------------------------------------
    def backward(self):
        """Returns dict(ref0=(H, W) tensor, Lambda=..., alpha0=..., alpha1=..., beta=..., gamma=...)."""
        s = self.s
        p = self.param
        prm = self._prm()
        sizes = self.sizes
        nL = len(sizes)
        n_iter = int(p.nSolverIters)
        ks = self.v.median_filter_kernel_size
        gf = (self.gflow / self.n_valid).contiguous()
        gu = _z(sizes[0])
        gub = _z(sizes[0])
        gref = [_z(sz) for sz in sizes]
        gpar = [_z(sz, 5) for sz in sizes]
        for k in range(nL):
            hh, ww = sizes[k]
            m = self.v.masks[k]
            L = self.lv[k]
            A = dict(gp=_z(sizes[k], 2), gq=_z(sizes[k], 4), um=_z(sizes[k], 3), vm=_z(sizes[k], 6),
                     qmsg=_z(sizes[k], 4), tmsg=_z(sizes[k], 4), gabc=_z(sizes[k], 3), geta=_z(sizes[k], 3),
                     gderiv=_z(sizes[k], 2), gvbB=_z(sizes[k], 2), gubB=_z(sizes[k]), guf=_z(sizes[k]),
                     gtv=_z(sizes[k], 2), gow=_z(sizes[k]), guth=_z(sizes[k]), gvb_fin=_z(sizes[k], 2))
            pos = torch.empty(sizes[k], dtype=torch.int32, device="cuda")
            tv, ow, self.deriv = _e(sizes[k], 2), _e(sizes[k]), _e(sizes[k], 2)
            tp = _e(sizes[k], 2)
            zero = dict(p=_z(sizes[k], 2), q=_z(sizes[k], 4), vb=_z(sizes[k], 2))
            for w in range(int(p.nWarpIters) - 1, -1, -1):
                rec = self.saved[k][w]
                self._warp_inputs(k, rec["flow"], tv, ow, self.deriv)
                # tapes of every checkpoint segment, the last one first (its end is u_th)
                segs = []
                for si, ck in enumerate(rec["ckpt"]):
                    segs.append((si * self.checkpoint, min(n_iter, (si + 1) * self.checkpoint), ck))
                # clamp and median need the final thresholded u: recompute the last segment
                tapes = {}

                def tape_of(seg):
                    a, b, ck = seg
                    st = self._clone(ck)
                    tape = [self._clone(st)]
                    for it in range(a, b):
                        self._pd_iter(k, st, it, prm, tp)
                        tape.append(self._clone(st))
                    return tape

                if segs:
                    tapes[len(segs) - 1] = tape_of(segs[-1])
                    u_th = tapes[len(segs) - 1][-1]["u"]
                else:
                    u_th = rec["u"]
                # clamp: gf unchanged (identity), gu -> clamp share of d/du_w; guf, gtv
                gtv = A["gtv"]
                H.call("vfs_adj_clamp", s, H.ptr(gu), H.ptr(A["guf"]), H.ptr(pos), H.ptr(gtv), H.ptr(gf),
                       H.ptr(u_th), H.ptr(rec["u"]), H.ptr(tv), H.ptr(m), hh, ww, float(p.limitRange), ks)
                H.call("vfs_adj_median", s, H.ptr(A["guth"]), H.ptr(A["guf"]), H.ptr(pos), hh, ww, ks)
                # primal-dual reverse: adjoint state after the last iteration
                g_u_pd = A["guth"].clone()
                A["gubB"].copy_(gub)
                A["gvbB"].zero_()
                A["gp"].zero_()
                A["gq"].zero_()
                A["gderiv"].zero_()
                has_msgs = 0
                for si in range(len(segs) - 1, -1, -1):
                    a, b, _ = segs[si]
                    tape = tapes.pop(si) if si in tapes else tape_of(segs[si])
                    for it in range(b - 1, a - 1, -1):
                        after, before = tape[it - a + 1], tape[it - a]
                        ad = H.VfsAdjPd()
                        ad.H, ad.W = hh, ww
                        for name, t in (("gu", g_u_pd), ("gub", A["gubB"]), ("gvb", A["gvbB"]), ("gp", A["gp"]),
                                        ("gq", A["gq"]), ("um", A["um"]), ("vm", A["vm"]), ("qmsg", A["qmsg"]),
                                        ("tmsg", A["tmsg"]), ("gabc", A["gabc"]), ("geta", A["geta"]),
                                        ("gderiv", A["gderiv"]), ("gpar", gpar[k]), ("p", after["p"]),
                                        ("q", after["q"]), ("u", after["u"]), ("p_prev", before["p"]),
                                        ("q_prev", before["q"]), ("u_prev", before["u"]), ("u_bar_prev", before["ub"]),
                                        ("v_bar_prev", before["vb"]), ("abc", L["abc"]), ("eta", L["eta"]),
                                        ("deriv", self.deriv), ("mask", m)):
                            setattr(ad, name, t.data_ptr())
                        H.call("vfs_adj_pd_iteration", s, ctypes.byref(ad), ctypes.byref(prm), float(self.taus[it]),
                               float(self.sigmas[it]), float(self.mus[it]), has_msgs)
                        has_msgs = 1
                    del tape
                # adjoint of u_bar at warp start (own part + messages); v_bar, p, q start at 0
                if has_msgs:
                    H.call("vfs_adj_pd_finish", s, H.ptr(gub), H.ptr(A["gvb_fin"]), H.ptr(A["gubB"]),
                           H.ptr(A["gvbB"]), H.ptr(A["um"]), H.ptr(A["vm"]), H.ptr(m), hh, ww)
                else:
                    gub.copy_(A["gubB"])
                gu += g_u_pd  # the loop started from u_w
                # derivatives -> reference level image and warped image; warp -> flow
                A["gow"].zero_()
                H.call("vfs_adj_derivatives", s, H.ptr(gref[k]), H.ptr(A["gow"]), H.ptr(A["gderiv"]),
                       H.ptr(self.v.translations[k]), hh, ww)
                H.call("vfs_adj_warp", s, H.ptr(gf), H.ptr(A["gow"]), H.ptr(gtv), H.ptr(rec["flow"]),
                       H.ptr(self.others[k]), H.ptr(self.v.translations[k]), hh, ww)
            # level constants: tensor and Gaussian
            gIs = _z(sizes[k])
            H.call("vfs_adj_tensor", s, H.ptr(gIs), H.ptr(gpar[k]), H.ptr(A["gabc"]), H.ptr(A["geta"]),
                   H.ptr(L["Is"]), H.ptr(m), hh, ww, ctypes.byref(prm))
            H.call("vfs_adj_gaussian", s, H.ptr(gref[k]), H.ptr(gIs), H.ptr(m), hh, ww)
            # level start: u = u_bar = fScale up(u_coarse), flow = fScale up(flow_coarse)
            if k < nL - 1:
                ch, cw = sizes[k + 1]
                gtot = (gu + gub).contiguous()
                ngu, ngf = _z(sizes[k + 1]), _z(sizes[k + 1], 2)
                H.call("vfs_adj_upsample", s, H.ptr(ngu), ch, cw, H.ptr(gtot), hh, ww, 1, float(p.fScale))
                H.call("vfs_adj_upsample", s, H.ptr(ngf), ch, cw, H.ptr(gf), hh, ww, 2, float(p.fScale))
                gu, gf, gub = ngu, ngf, _z(sizes[k + 1])
        # image pyramid: refs[k] = down(refs[k - 1])
        for k in range(nL - 1, 0, -1):
            H.call("vfs_adj_downsample", s, H.ptr(gref[k - 1]), *sizes[k - 1], H.ptr(gref[k]), *sizes[k])
        tot = torch.zeros(5, dtype=torch.float64)
        for g in gpar:
            tot += torch.as_tensor(self._reduce(g, 5))
        out = {name: float(tot[i]) for i, name in enumerate(PARAMS)}
        out["ref0"] = gref[0]
        return out


class DepthLoss(torch.autograd.Function):
    """loss = DepthLoss.apply(ref0, theta, adj, img_other); theta = (lambda, alpha0, alpha1, beta, gamma)."""

    @staticmethod
    def forward(ctx, ref0, theta, adj, img_other):
        adj.set_params(theta.detach().cpu().numpy())
        loss = adj.forward(ref0.detach(), img_other)
        ctx.adj = adj
        return torch.tensor(loss, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gl):
        g = ctx.adj.backward()
        gl = float(gl)
        gtheta = torch.tensor([g[n] * gl for n in PARAMS], dtype=torch.float64)
        return g["ref0"] * gl, gtheta, None, None
