"""Gradient fit of the solver parameters (lambda, alpha0, alpha1, beta, gamma) on
synthetic pairs with known depth, through the hand-written adjoint.

The loss of a pair is the depth loss of ``vfs_adjoint`` (mean |log z - log z_gt|)
over the mask pixels whose ground-truth disparity is at least ``min_disparity``
pixels (near the epipoles the disparity vanishes and triangulation is
ill-conditioned: those pixels would dominate a mean of |log z - log z_gt|).  A
step takes the mean loss and gradient over a chunk of pairs and updates the
logarithms of the five parameters with Adam (they stay positive).  The pairs of a
step run on ``streams`` HIP streams at once (``vfs_adjoint.gradients``), the pairs of
one stream one after another sharing that stream's reverse workspace; at most
``resident`` pairs keep their adjoint buffers on the GPU (least recently used
first out), so a step over many pairs streams them through in chunks.

The gradient is by default the truncated one (``reverse_warps`` = 1: the reverse
of the last warp of the finest level, its start state held fixed).  The exact
gradient of the whole unrolled solve (``reverse_warps`` = 0) is available and is
exact, but over 3 levels x 10 warps x 50 iterations the solve's sensitivity to
the parameters is chaotic: the exact derivative is orders of magnitude larger
than any finite-difference slope down to h = 1e-6 (relative), and it does not
give descent at usable step sizes; the truncated one does.

FIT_FULL is the full configuration (256 pairs at 1024 x 1024, every Adam step over
all of them, 32 resident at a time, 200 steps); FIT_DEMO is the scaled run that
bench.py times.

    python fit.py --spec demo          # on the GPU
"""

import argparse
import collections
import json
import math
import sys
import time

import numpy as np
import torch

import vfs_adjoint as A

PARAMS = A.PARAMS
# The scenes: depth base drawn in [3, 12] m, clipped to [1, 20] m; texture contrast in
# [0.3, 1]; rotation up to 3 deg about a random axis; baseline 0.1-0.4 m mostly along x.
SCENE = dict(depth_base=(3.0, 12.0), depth_clip=(1.0, 20.0), contrast=(0.3, 1.0), rot_deg=3.0, t_norm=(0.1, 0.4))
# Solver schedule of the fit: 3 levels x 10 warps x 50 iterations, checkpoint every 25.
SCHEDULE = dict(nLevel=3, nWarpIters=10, nSolverIters=50, fScale=2.0, checkpoint=25)
# the reference's defaults (Python/vfs_params.py:33-43): the manual starting point
START = dict(Lambda=5.0, alpha0=17.0, alpha1=1.2, beta=9.0, gamma=0.85)
MIN_DISPARITY = 2.0
FIT_FULL = dict(size=(1024, 1024), pairs=256, batch=256, resident=32, steps=200, held_out=32, lr=0.05,
                reverse_warps=1)
FIT_DEMO = dict(size=(512, 512), pairs=16, batch=16, resident=16, steps=12, held_out=8, lr=0.05, reverse_warps=1)


def scene_pairs(n, size, base_seed, device="cuda"):
    """n seeded pairs of the fit's scene distribution with their vector fields (pair i: seed base_seed + i)."""
    import synthetic as S
    import vector_field as VF
    import equidistant_cam as E
    h, w = size
    out = []
    for i in range(n):
        rng = np.random.Generator(np.random.PCG64(base_seed + 20_000 + i))
        d = np.array([1.0, rng.uniform(-0.1, 0.1), rng.uniform(-0.1, 0.1)])
        t = d / np.linalg.norm(d) * rng.uniform(*SCENE["t_norm"])
        pr = S.make_pair(h, w, base_seed + i, t=tuple(t), rot_deg=SCENE["rot_deg"], rot_random=True,
                         n_levels=SCHEDULE["nLevel"], depth_base=rng.uniform(*SCENE["depth_base"]),
                         depth_clip=SCENE["depth_clip"], contrast=rng.uniform(*SCENE["contrast"]), device=device)
        cam = E.EquidistantCam(pr["cam"])
        cam.setup_module()
        tvf, cvf = VF.get_vector_fields(VF.get_grid_zero((h, w)), cam, cam, pr["R"], pr["t"], 1, VF.unit_normalizer)
        pr["tvf"], pr["cvf"], pr["camera"] = S.sanitize_field(tvf), np.asarray(cvf, np.float32), cam
        pr["disparity_gt"] = disparity_gt(pr)
        out.append(pr)
    return out


def disparity_gt(pr):
    """Ground-truth image motion (px) a baseline of |t| causes at each pixel: f |t x d| / D for
    the unit ray d of the equidistant camera and the range D (small-angle form)."""
    f, cx, cy = pr["cam"]
    h, w = pr["img_ref"].shape
    X, Y = np.meshgrid(np.arange(w) - cx, np.arange(h) - cy)
    th, ph = np.hypot(X, Y) / f, np.arctan2(Y, X)
    ray = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1)
    t = np.asarray(pr["t"], dtype=np.float64).reshape(3)
    return f * np.linalg.norm(np.cross(ray, t), axis=-1) / pr["range_gt"]


def loss_target(pr, min_disparity=MIN_DISPARITY):
    """z_gt with the ill-conditioned pixels (disparity below min_disparity) marked -1 (excluded)."""
    z = np.array(pr["z_gt"], dtype=np.float64, copy=True)
    z[pr["disparity_gt"] < min_disparity] = -1.0
    return z


def make_param(img, theta):
    import vfs_params as P
    prm = P.vfs_params(img)
    for k in ("nLevel", "nWarpIters", "nSolverIters", "fScale"):
        setattr(prm, k, SCHEDULE[k])
    for k, v in theta.items():
        setattr(prm, k, float(v))
    return prm


def median_depth_error(pairs, theta, min_disparity=MIN_DISPARITY):
    """Median over pairs of the median relative depth error |z - z_gt| / z_gt over the mask, and
    the mean over pairs of the fit's loss (mean |log z - log z_gt| over the well-conditioned
    mask pixels, z > 0), solved with the blocked path (HIP graph) at the fit's schedule.
    Returns (median error, per-pair errors, mean loss)."""
    import vfs as V
    errs, losses = [], []
    for pr in pairs:
        v = V.VFS(make_param(pr["img_ref"], theta))
        v.solver = "blocked"
        v.setup_module()
        v.compute_pyramid_sizes()
        v.create_mask_pyramid(pr["mask"])
        v.create_vector_field_pyramid(pr["tvf"], pr["cvf"])
        v.solve_stereo(pr["img_ref"], pr["img_other"])
        z = v.triangulate_flow(pr["camera"], pr["t"])[..., 2].cpu().numpy()
        m = (pr["mask"] > 0) & np.isfinite(z) & (pr["z_gt"] > 1e-3)
        errs.append(float(np.median(np.abs(z[m] - pr["z_gt"][m]) / pr["z_gt"][m])))
        zt = loss_target(pr, min_disparity)
        ok = (pr["mask"] > 0) & np.isfinite(z) & (z > 0) & (zt > 0)
        losses.append(float(np.mean(np.abs(np.log(z[ok]) - np.log(zt[ok])))) if ok.any() else float("nan"))
        v.release_plan()
    return float(np.median(errs)), errs, float(np.mean(losses))


class ParameterFit:
    """Adam on log(theta) with the mean depth-loss gradient over chunks of pairs."""

    def __init__(self, pairs, theta0=START, lr=0.05, betas=(0.9, 0.999), eps=1e-8, reverse_warps=1,
                 min_disparity=MIN_DISPARITY, resident=32, streams=2):
        self.pairs = pairs
        self.streams = max(1, int(streams))
        self.ws = [{} for _ in range(self.streams)]  # one shared reverse workspace per stream
        self.theta0 = dict(theta0)
        self.reverse_warps = int(reverse_warps)
        self.min_disparity = float(min_disparity)
        self.resident = max(1, int(resident))
        self._adj = collections.OrderedDict()
        self.phi = np.log(np.array([theta0[k] for k in PARAMS], dtype=np.float64))
        self.m = np.zeros(5)
        self.v = np.zeros(5)
        self.t = 0
        self.lr, self.b1, self.b2, self.eps = lr, betas[0], betas[1], eps
        self.history = []

    def theta(self):
        return {k: float(x) for k, x in zip(PARAMS, np.exp(self.phi))}

    def adjoint(self, i):
        """The adjoint of pair i with its device buffers (created on first use; the least
        recently used one is released beyond `resident`)."""
        if i in self._adj:
            self._adj.move_to_end(i)
            return self._adj[i]
        pr = self.pairs[i]
        a = A.VFSDepthAdjoint(make_param(pr["img_ref"], self.theta0), pr["mask"], pr["tvf"], pr["cvf"], pr["cam"],
                              pr["t"], loss_target(pr, self.min_disparity), checkpoint=SCHEDULE["checkpoint"],
                              workspace=self.ws[0], reverse_warps=self.reverse_warps)
        ref = torch.as_tensor(pr["img_ref"], device="cuda").to(torch.float32) / 256.0
        other = torch.as_tensor(pr["img_other"], device="cuda")
        self._adj[i] = (a, ref, other)
        while len(self._adj) > self.resident:
            self._adj.popitem(last=False)
        return self._adj[i]

    def loss_and_grad(self, idx):
        """Mean loss and d(mean loss)/d(log theta) over the pairs idx."""
        th = np.exp(self.phi)
        tot, g = 0.0, np.zeros(5)
        # chunks of at most `resident` pairs, each chunk's pairs on `streams` streams
        for c in range(0, len(idx), self.resident):
            items = []
            for i in idx[c:c + self.resident]:
                a, ref, other = self.adjoint(i)
                a.set_params(th)
                items.append((a, ref, other))
            for loss, gr in A.gradients(items, streams=self.streams, workspaces=self.ws):
                tot += loss
                g += np.array([gr[k] for k in PARAMS])
        n = len(idx)
        return tot / n, (g / n) * th

    def step(self, idx):
        at = self.theta()  # the parameters the step's loss is measured at
        loss, g = self.loss_and_grad(idx)
        self.t += 1
        self.m = self.b1 * self.m + (1 - self.b1) * g
        self.v = self.b2 * self.v + (1 - self.b2) * g * g
        mh = self.m / (1 - self.b1 ** self.t)
        vh = self.v / (1 - self.b2 ** self.t)
        self.phi = self.phi - self.lr * mh / (np.sqrt(vh) + self.eps)
        self.history.append(dict(step=self.t, loss=loss, at=at, **self.theta()))
        return loss

    def run(self, steps, batch, log=None):
        """`steps` Adam steps, each over `batch` pairs (consecutive, cycling through the pairs)."""
        n = len(self.pairs)
        for s in range(steps):
            start = (s * batch) % n
            idx = [(start + j) % n for j in range(min(batch, n))]
            loss = self.step(idx)
            if log:
                log(f"[fit] step {self.t} loss {loss:.5f} " + " ".join(f"{k} {v:.4g}" for k, v in self.theta().items()))
        return self.history


def run_spec(spec, base_seed=3000, log=None):
    """Generate the pairs, evaluate the start, fit, evaluate again.  Returns a result dict."""
    t0 = time.perf_counter()
    train = scene_pairs(spec["pairs"], spec["size"], base_seed)
    held = scene_pairs(spec["held_out"], spec["size"], base_seed + 100_000)
    t_gen = time.perf_counter() - t0
    err0, _, hl0 = median_depth_error(held, START)
    f = ParameterFit(train, START, lr=spec["lr"], reverse_warps=spec["reverse_warps"], resident=spec["resident"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = f.run(spec["steps"], spec["batch"], log=log)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t1
    # the iterate with the lowest training loss (the truncated gradient is not the loss's exact
    # gradient, so the loss need not keep falling); chosen on the training pairs only
    best = min(f.history, key=lambda h: h["loss"])
    theta = f.theta()
    err1, _, hl1 = median_depth_error(held, theta)
    err_b, _, hl_b = median_depth_error(held, best["at"])
    grads = spec["steps"] * min(spec["batch"], spec["pairs"])
    return dict(spec=spec, start=START, fitted=theta, held_out_median_rel_depth_err_start=err0,
                held_out_median_rel_depth_err_fitted=err1, held_out_loss_start=hl0, held_out_loss_fitted=hl1,
                train_loss=[h["loss"] for h in hist], best_step=best["step"], best_train_loss=best["loss"],
                best_theta=best["at"], held_out_median_rel_depth_err_best=err_b, held_out_loss_best=hl_b,
                fit_seconds=t_fit, pair_gradients=grads, pair_gradients_per_s=grads / t_fit, generate_seconds=t_gen)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spec", default="demo", choices=["demo", "full"])
    ap.add_argument("--steps", type=int, default=None)
    a = ap.parse_args()
    spec = dict(FIT_DEMO if a.spec == "demo" else FIT_FULL)
    if a.steps:
        spec["steps"] = a.steps
    res = run_spec(spec, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(res))


if __name__ == "__main__":
    main()
