"""Benchmark: 800x800 fisheye depth maps per second on one MI355X.

Workload (default): one seeded synthetic pair at 800x800 (f = 800/pi, 180 deg FOV,
baseline 0.3 m along x, rotation <= 2 deg), 5 pyramid levels (ratio 2) x 20 warps x
50 primal-dual iterations, with the reference's default solver parameters
(lambda = 5, alpha0 = 17, alpha1 = 1.2, beta = 9, gamma = 0.85), fp32.  A step is
one VFS.solve_stereo (uint8 images already on the GPU -> flow) followed by
triangulate_flow (-> depth); the timed region covers K steps between device
synchronisations.

Prints one JSON line with the throughput, the roofline of the dominant kernel
(measured with HIP events on the launch stream), and the float64 CPU oracle
timed on the host beside it.  The default run adds three legs to the same line: the
64-pair batched solve ("batched_64") and the depth-loss gradient on two 800x800
pairs ("adjoint_800"), each with its own roofline and CPU baseline, and the scaled
parameter fit ("fit_512").
"""

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "vfs-python_amd")
for _p in (PKG, ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

METRIC = "800² depth maps/s & solver HBM GB/s; batched/adjoint vs plain at same error"
HBM_PEAK_GBS = 8000.0
# the gradient workload's pairs run on this many HIP streams at once (vfs_adjoint.gradients)
ADJ_STREAMS = 2
# Algorithmic bytes per pixel of a level in the per-kernel formulation.
B_DUAL, B_L1, B_PRIMAL = 81, 41, 77
B_ITER = B_DUAL + B_L1 + B_PRIMAL
B_WARP = 150
B_LEVEL = 38


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def setup_solver(name, seed, solver, device=None):
    import torch
    import vfs as V
    import vfs_params as P
    import vector_field as VF
    import equidistant_cam as E
    import synthetic as S

    pair = S.workload_pair(name, seed=seed, device=device)
    cam = E.EquidistantCam(pair["cam"])
    cam.setup_module()
    Hh, Ww = pair["img_ref"].shape
    tvf, cvf = VF.get_vector_fields(VF.get_grid_zero((Hh, Ww)), cam, cam, pair["R"], pair["t"], 1,
                                    VF.unit_normalizer)
    tvf = S.sanitize_field(tvf)
    prm = P.vfs_params(pair["img_ref"])
    S.apply_workload_params(prm, name)
    v = V.VFS(prm)
    v.solver = solver
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(pair["mask"])
    v.create_vector_field_pyramid(tvf, cvf)
    img_ref = torch.as_tensor(pair["img_ref"], device="cuda")
    img_other = torch.as_tensor(pair["img_other"], device="cuda")
    return dict(pair=pair, cam=cam, vfs=v, prm=prm, tvf=tvf, cvf=cvf, img_ref=img_ref, img_other=img_other)


def dual_kernel_roofline(sizes, n_warp, n_iter, prm):
    """Average duration and algorithmic bandwidth of vfs_update_dual over the launch mix of one solve.

    Every level's n_warp * n_iter launches are replayed at that level's size on the
    current stream, bracketed by HIP events (torch.cuda.Event on the same stream).
    """
    import ctypes
    import torch
    import vfs_hip as H
    k = H.make_params(prm)
    g = torch.Generator(device="cuda")
    g.manual_seed(0)
    total_ms, total_launches, total_bytes = 0.0, 0, 0
    per_level = []
    s = H.stream()
    for (hh, ww) in sizes:
        n = n_warp * n_iter
        p = torch.zeros((hh, ww, 2), device="cuda")
        q = torch.zeros((hh, ww, 4), device="cuda")
        tp = torch.empty((hh, ww, 2), device="cuda")
        u = torch.rand((hh, ww), device="cuda", generator=g)
        vb = torch.rand((hh, ww, 2), device="cuda", generator=g) * 0.01
        abc = torch.rand((hh, ww, 3), device="cuda", generator=g)
        mask = (torch.rand((hh, ww), device="cuda", generator=g) > 0.2).to(torch.uint8)
        args = (H.ptr(p), H.ptr(tp), H.ptr(q), 1.0, H.ptr(u), H.ptr(vb), H.ptr(abc), H.ptr(mask), hh, ww,
                ctypes.byref(k))
        fn = H.lib().vfs_update_dual
        for _ in range(5):
            fn(s, *args)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            rc = fn(s, *args)
        e1.record()
        torch.cuda.synchronize()
        H.check(rc, "vfs_update_dual")
        ms = e0.elapsed_time(e1)
        per_level.append(dict(size=[hh, ww], launches=n, avg_us=1e3 * ms / n,
                              gbs=B_DUAL * hh * ww * n / (ms * 1e-3) / 1e9))
        total_ms += ms
        total_launches += n
        total_bytes += B_DUAL * hh * ww * n
    avg_s = total_ms * 1e-3 / total_launches
    bytes_per_launch = total_bytes / total_launches
    return dict(avg_us=avg_s * 1e6, bytes_per_launch=bytes_per_launch,
                achieved_gbs=bytes_per_launch / avg_s / 1e9, per_level=per_level)


def blocked_kernel_roofline(sizes, tiles, n_warp, n_iter, levels=None):
    """Average duration and algorithmic bandwidth of the blocked primal-dual kernel over one solve's launches.

    Bytes per launch (compulsory traffic of the formulation): every pixel of the
    tile extended by k on every side, clipped to the image,
    reads 73 B (u, u_bar, v_bar, p, q = 40 B; abc, eta, deriv = 32 B; mask 1 B;
    the warp's first launch reads 41 B, its v_bar, p, q being zero), and every
    tile pixel writes 40 B.  Each level's warp is recorded into a HIP graph and replayed
    (as in the solve), on the solved pair's own level data when `levels` is given.
    """
    import torch
    import autotune
    total_ms, total_launches, total_bytes = 0.0, 0, 0
    per_level = []
    for (hh, ww), (tx, ty, k, nt) in zip(sizes, tiles):
        nblk = -(-n_iter // k)

        ext = _ext_px(hh, ww, tx, ty, k)
        warp_bytes = (41 * ext + 40 * hh * ww) + (nblk - 1) * (73 * ext + 40 * hh * ww)
        ms = autotune.time_variant(hh, ww, n_iter, (tx, ty, k, nt), reps=n_warp,
                                   level=None if levels is None else levels[len(per_level)])
        per_level.append(dict(size=[hh, ww], tile=[tx, ty, k, nt], launches=nblk * n_warp, avg_us=1e3 * ms / nblk,
                              gbs=warp_bytes / (ms * 1e-3) / 1e9))
        total_ms += ms * n_warp
        total_launches += nblk * n_warp
        total_bytes += warp_bytes * n_warp
    avg_s = total_ms * 1e-3 / total_launches
    bytes_per_launch = total_bytes / total_launches
    return dict(avg_us=avg_s * 1e6, bytes_per_launch=bytes_per_launch,
                achieved_gbs=bytes_per_launch / avg_s / 1e9, per_level=per_level)


def fit_leg():
    """The scaled parameter fit (fit.FIT_DEMO): 16 pairs at 512x512, 12 Adam steps on the
    5 solver parameters from the reference's defaults through the adjoint (truncated
    gradient); loss and median relative depth error of 8 held-out pairs before and after.
    The timed region is the fit itself."""
    import fit
    r = fit.run_spec(dict(fit.FIT_DEMO), log=log)
    return dict(
        value=round(r["pair_gradients_per_s"], 3), unit="pair gradients/s", steps=r["spec"]["steps"],
        ms_per_step=round(1e3 * r["fit_seconds"] / r["spec"]["steps"], 3),
        config={"workload": "fit_512", "size": list(r["spec"]["size"]), "pairs": r["spec"]["pairs"],
                "pairs_per_step": r["spec"]["batch"], "adam_steps": r["spec"]["steps"],
                "held_out": r["spec"]["held_out"], "schedule": fit.SCHEDULE},
        fit_seconds=round(r["fit_seconds"], 3),
        start=r["start"],
        fitted={k: round(v, 5) for k, v in r["fitted"].items()},
        gradient="truncated (last warp of the finest level)" if r["spec"]["reverse_warps"] else "exact",
        held_out_loss={"start": round(r["held_out_loss_start"], 5), "fitted": round(r["held_out_loss_fitted"], 5)},
        held_out_median_rel_depth_err={"start": round(r["held_out_median_rel_depth_err_start"], 5),
                                       "fitted": round(r["held_out_median_rel_depth_err_fitted"], 5)},
        train_loss=[round(x, 5) for x in r["train_loss"]],
        best_by_train_loss={"step": r["best_step"], "theta": {k: round(v, 5) for k, v in r["best_theta"].items()},
                            "held_out_loss": round(r["held_out_loss_best"], 5),
                            "held_out_median_rel_depth_err": round(r["held_out_median_rel_depth_err_best"], 5)},
        # the fit's kernels are the gradient leg's (forward blocked launches, reverse sweep);
        # their roofline and CPU baseline are reported there
        roofline=None, cpu_baseline=None)


def run_fit(a):
    """The fit alone (--workload fit_512; --steps / --warmup do not apply)."""
    r = fit_leg()
    out = {"metric": METRIC, "value": r.pop("value"), "unit": r.pop("unit"), "n_gpus": 1, "steps": r.pop("steps"),
           "warmup": 0, "ms_per_step": r.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": r.pop("config")}
    out.update(r)
    print(json.dumps(out), flush=True)


_COPY_GBS = {}


def stream_copy_gbs(n=1 << 26, reps=5):
    """A measured stream-copy rate of the device: vfs_copy of n floats (read + write,
    8 n bytes), best of `reps` launches timed with HIP events on the launch stream
    (measured once per process)."""
    if n in _COPY_GBS:
        return _COPY_GBS[n]
    import torch
    import vfs_hip as H
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    st = torch.cuda.current_stream()
    best = float("inf")
    for _ in range(reps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        H.call("vfs_copy", H.stream(), H.ptr(dst), H.ptr(src), n)
        e1.record(st)
        e1.synchronize()
        best = min(best, e0.elapsed_time(e1))
    del src, dst
    _COPY_GBS[n] = 8.0 * n / (best * 1e-3) / 1e9
    return _COPY_GBS[n]


def measured_traffic(per_level):
    """HBM bytes per k_pd_block launch, averaged over the launches of one solve like
    `achieved`, from the FETCH_SIZE / WRITE_SIZE runs recorded in
    profiles/pmc_traffic.json (each counter scaled by its reading for a plain copy of
    known size).  None unless every level ran the variant that was measured."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = {(tuple(e["size"]), tuple(e["tile"])): e for e in json.load(f)["levels"]}
    tot_b, tot_n = 0.0, 0
    for lv in per_level:
        e = rec.get((tuple(lv["size"]), tuple(lv["tile"])))
        if e is None:
            return None
        tot_b += lv["launches"] * (e["fetch_bytes"] + e["write_bytes"])
        tot_n += lv["launches"]
    return round(tot_b / tot_n)


def measured_adj_traffic(tile):
    """HBM bytes per k_adj_block launch of the given variant at the 800x800 gradient
    workload: FETCH_SIZE + WRITE_SIZE (each scaled by its reading for a plain copy of
    known size) averaged over the variant's launches of one gradient, from the counter
    table in profiles/kernels_adjoint_800_hbm_lds.md.  None if the table has no single
    row for the variant."""
    path = os.path.join(ROOT, "profiles", "kernels_adjoint_800_hbm_lds.md")
    if not os.path.exists(path):
        return None
    head = "| k_adj_block<" + ", ".join(str(x) for x in tile)
    rows, cols = [], None
    with open(path) as f:
        for line in f:
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if cols is None and "HBM MB / call" in cells:
                cols = cells
            elif line.startswith(head):
                rows.append(cells)
    if cols is None or len(rows) != 1:
        return None
    return round(float(rows[0][cols.index("HBM MB / call")]) * 1e6)


def batched_kernel_roofline(batch):
    """Average duration and algorithmic bandwidth of the batched blocked primal-dual kernel
    over one batch solve's launches.  Per level, one warp of the batch's blocked launches
    (vfs_pd_blocks_batch: every pair's tiles in each launch) is recorded into a HIP graph
    and replayed between HIP events on the launch stream, on the batch's own level data.
    Bytes per launch: the sum over pairs of the single-pair formula of
    blocked_kernel_roofline (73 B per pixel of each extended tile, 41 B in a warp's first
    launch, 40 B written per tile pixel)."""
    import ctypes
    import torch
    import vfs_hip as H
    p = batch.param
    n_iter, n_warp = int(p.nSolverIters), int(p.nWarpIters)
    prm = H.make_params(p, batch.vfs[0].median_filter_kernel_size)
    taus, sigmas, mus = H.pd_schedule(n_iter, p.timeStepLambda)
    sched = [x.ctypes.data_as(ctypes.c_void_p) for x in (taus, sigmas, mus)]
    n = len(batch.vfs)
    s = H.stream()
    total_ms, total_launches, total_bytes = 0.0, 0, 0
    per_level = []
    for lvl, (tx, ty, k, nt) in enumerate(batch._tiles()):
        levels, wss = batch._structs[lvl]
        nblk = -(-n_iter // k)
        ext = tile_px = 0
        for v in batch.vfs:
            hh, ww = v.sizes[lvl]
            tile_px += hh * ww
            for y0 in range(0, hh, ty):
                ey = min(hh, y0 + ty + k) - max(0, y0 - k)
                for x0 in range(0, ww, tx):
                    ext += ey * (min(ww, x0 + tx + k) - max(0, x0 - k))
        warp_bytes = (41 * ext + 40 * tile_px) + (nblk - 1) * (73 * ext + 40 * tile_px)
        args = (H.ptr(batch._ws[lvl]), ctypes.cast(levels, ctypes.c_void_p), ctypes.cast(wss, ctypes.c_void_p), n,
                ctypes.byref(prm), *sched, n_iter, tx, ty, k, nt)
        cs = ctypes.c_void_p()
        H.call("vfs_capture_begin", ctypes.byref(cs))
        try:
            H.call("vfs_pd_blocks_batch", cs, *args)
        finally:
            plan = ctypes.c_void_p()
            H.call("vfs_capture_end", cs, ctypes.byref(plan))
        try:
            H.call("vfs_plan_launch", plan, s)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            reps = max(1, min(n_warp, 5))
            e0.record()
            for _ in range(reps):
                H.call("vfs_plan_launch", plan, s)
            e1.record()
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / reps
        finally:
            H.lib().vfs_plan_destroy(plan)
        per_level.append(dict(level=lvl, pixels=tile_px, tile=[tx, ty, k, nt], launches=nblk * n_warp,
                              avg_us=round(1e3 * ms / nblk, 3), gbs=round(warp_bytes / (ms * 1e-3) / 1e9, 1)))
        total_ms += ms * n_warp
        total_launches += nblk * n_warp
        total_bytes += warp_bytes * n_warp
    avg_s = total_ms * 1e-3 / total_launches
    bytes_per_launch = total_bytes / total_launches
    return dict(avg_us=avg_s * 1e6, bytes_per_launch=bytes_per_launch,
                achieved_gbs=bytes_per_launch / avg_s / 1e9, per_level=per_level)


def batch_leg(n_pairs=64, steps=1, warmup=1, compare=True, roofline=True, cpu_baseline=True):
    """64 pairs of uneven size (16 each of 512^2, 800^2, 1024^2, 1280x960) through VFSBatch:
    pairs/s of the batched solve (+ triangulation), the per-kernel path one pair at a time
    beside it, the batched kernel's roofline and the depth error against the ground truth."""
    import torch
    import synthetic as S
    import vfs as V
    import vfs_batch as B
    import vfs_params as P
    import vector_field as VF
    import equidistant_cam as E
    spec = S.BATCH_64
    n_pairs = min(n_pairs, len(spec["sizes"])) if n_pairs else len(spec["sizes"])
    log(f"[bench] batch: generating {n_pairs} pairs on the GPU")
    pairs = S.batch_pairs(spec, device="cuda", n=n_pairs)
    sizes = [p["img_ref"].shape for p in pairs]
    prm = P.vfs_params()
    S.apply_workload_params(prm, spec)
    batch = B.VFSBatch(prm, sizes)
    cams, fields = [], []
    for i, p in enumerate(pairs):
        cam = E.EquidistantCam(p["cam"])
        cam.setup_module()
        cams.append(cam)
        tvf, cvf = VF.get_vector_fields(VF.get_grid_zero(sizes[i]), cam, cam, p["R"], p["t"], 1, VF.unit_normalizer)
        fields.append((S.sanitize_field(tvf), cvf))
        batch.setup_pair(i, p["mask"], *fields[-1])
    refs = [torch.as_tensor(p["img_ref"], device="cuda") for p in pairs]
    others = [torch.as_tensor(p["img_other"], device="cuda") for p in pairs]
    ts = [p["t"] for p in pairs]

    def step():
        batch.solve(refs, others)
        return batch.triangulate(cams, ts)

    log("[bench] batch: warmup (records the graph)")
    for _ in range(max(1, warmup)):
        step()
    torch.cuda.synchronize()
    log(f"[bench] batch: timing {steps} steps")
    t0 = time.perf_counter()
    for _ in range(steps):
        xyz = step()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    errs = []
    for x, p in zip(xyz, pairs):
        z = x[..., 2].cpu().numpy()
        m = (p["mask"] > 0) & np.isfinite(z) & (p["z_gt"] > 1e-3)
        errs.append(float(np.median(np.abs(z[m] - p["z_gt"][m]) / p["z_gt"][m])))
    px = sum(sum(h * w for h, w in v.sizes) for v in batch.vfs)
    px_iter = px * spec["nWarpIters"] * spec["nSolverIters"]
    bytes_batch = sum(sum(h * w * (spec["nWarpIters"] * (spec["nSolverIters"] * B_ITER + B_WARP) + B_LEVEL)
                          for h, w in v.sizes) for v in batch.vfs)
    value = len(pairs) * steps / dt
    plain = None
    if compare:
        # the per-kernel path, one pair at a time: one pair of each size class timed, scaled by its count
        per_size = {}
        for i, s in enumerate(sizes):
            if s in per_size:
                continue
            v = V.VFS(batch.vfs[i].param)
            v.solver = "graph"
            v.setup_module()
            v.compute_pyramid_sizes()
            v.masks, v.masks_to = batch.vfs[i].masks, batch.vfs[i].masks_to
            v.translations, v.translations_to = batch.vfs[i].translations, batch.vfs[i].translations_to
            v.calibration = batch.vfs[i].calibration
            v.solve_stereo(refs[i], others[i])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            for _ in range(2):
                v.solve_stereo(refs[i], others[i])
                v.triangulate_flow(cams[i], ts[i])
            torch.cuda.synchronize()
            per_size[s] = (time.perf_counter() - t1) / 2
            v.release_plan()
            log(f"[bench] per-kernel path {s}: {1e3 * per_size[s]:.1f} ms/pair")
        tot = sum(per_size[s] for s in sizes)
        plain = dict(value=round(len(pairs) / tot, 3), ms_per_batch=round(1e3 * tot, 1),
                     speedup_of_reported=round(tot / (dt / steps), 2))
    roof = None
    if roofline:
        log("[bench] batch: batched-kernel roofline pass (HIP events)")
        r = batched_kernel_roofline(batch)
        copy_gbs = stream_copy_gbs()
        roof = dict(bound="hbm", kernel="k_pd_block_batched", achieved=round(r["achieved_gbs"], 1),
                    peak=HBM_PEAK_GBS, unit="GB/s", frac=round(r["achieved_gbs"] / HBM_PEAK_GBS, 4), traffic=None,
                    avg_launch_us=round(r["avg_us"], 3), bytes_per_launch=round(r["bytes_per_launch"]),
                    stream_copy_gbs=round(copy_gbs, 1), frac_of_stream_copy=round(r["achieved_gbs"] / copy_gbs, 4),
                    per_level=r["per_level"])
    cpu = None
    if cpu_baseline:
        log("[bench] batch: CPU oracle baseline (one pair per size, 1 warp per level)")
        by_size = {}
        for i, p in enumerate(pairs):
            if sizes[i] not in by_size:
                by_size[sizes[i]] = [p, fields[i][0], np.asarray(fields[i][1], np.float32), 0]
            by_size[sizes[i]][3] += 1
        cpu = cpu_batch_baseline({k: tuple(v) for k, v in by_size.items()}, spec)
    batch.release_plan()
    return dict(value=round(value, 3), unit="depth maps/s", pairs=len(pairs), steps=steps,
                ms_per_batch=round(1e3 * dt / steps, 3), sizes=sorted({f"{w}x{h}" for h, w in sizes}),
                levels=spec["nLevel"], warps_per_level=spec["nWarpIters"], pd_iters_per_warp=spec["nSolverIters"],
                pd_tiles=[list(t) for t in batch._tiles()], plain_per_kernel=plain,
                mpx_iter_per_s=round(px_iter * steps / dt / 1e6, 1),
                solver_effective_gbs=round(bytes_batch * steps / dt / 1e9, 1),
                median_rel_depth_err_vs_gt=round(float(np.median(errs)), 4),
                max_rel_depth_err_vs_gt=round(float(np.max(errs)), 4), roofline=roof, cpu_baseline=cpu)


def run_batch(a):
    """The batched workload alone (--workload batched_64)."""
    r = batch_leg(a.pairs, a.steps, a.warmup, compare=not a.no_compare, roofline=not a.no_roofline,
                  cpu_baseline=not a.no_cpu_baseline)
    out = {
        "metric": METRIC, "value": r["value"], "unit": "depth maps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r["ms_per_batch"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched_64", "pairs": r["pairs"], "sizes": r["sizes"], "levels": r["levels"],
                   "warps_per_level": r["warps_per_level"], "pd_iters_per_warp": r["pd_iters_per_warp"],
                   "solver": "batched blocked (one graph per batch)", "pd_tiles": r["pd_tiles"]},
        "plain_per_kernel": r["plain_per_kernel"],
        "mpx_iter_per_s": r["mpx_iter_per_s"],
        "solver_effective_gbs": r["solver_effective_gbs"],
        "median_rel_depth_err_vs_gt": r["median_rel_depth_err_vs_gt"],
        "roofline": r["roofline"], "cpu_baseline": r["cpu_baseline"],
    }
    print(json.dumps(out), flush=True)


def _ext_px(hh, ww, tx, ty, k):
    """Pixels of every (tx, ty) tile of an hh x ww image extended by k on each side, clipped to the image."""
    tot = 0
    for y0 in range(0, hh, ty):
        ey = min(hh, y0 + ty + k) - max(0, y0 - k)
        for x0 in range(0, ww, tx):
            tot += ey * (min(ww, x0 + tx + k) - max(0, x0 - k))
    return tot


# k_adj_block algorithmic bytes: every pixel of the extended tile (clipped to the image, inside
# the mask or not) reads the mask (1 B), the adjoint state (gu, gub 4 + 4, gvb 8, gp 8, gq 16,
# primal^T messages 8, threshold^T message 4 = 52 B), the level inputs (abc 12, eta 12, deriv 8
# = 32 B) and per unit the forward tape (p 8, q 16, u_bar 4, v_bar 8, u and u_bar of the
# previous iteration 4 + 4 = 44 B); every tile pixel reads and writes its accumulators (abc,
# eta, deriv shares 32 B, lambda / alpha0 / alpha1 shares 12 B: 44 B each way) and writes the
# adjoint state (52 B)
ADJ_EXT_B, ADJ_UNIT_B, ADJ_TILE_B = 85, 44, 140


def adjoint_roofline(adj, ref0, img_other, n_launches):
    """Average duration and algorithmic bandwidth of the temporally blocked reverse-sweep
    kernel (k_adj_block) over the finest level's launches of one gradient, each bracketed
    by HIP events on the stream it is launched on (VFSDepthAdjoint.backward's probe)."""
    adj.forward(ref0, img_other)
    probe = dict(level=0, launches=n_launches)
    adj.backward(probe=probe)
    tx, ty, kr, nt = probe["tiles"]
    hh, ww = adj.sizes[0]
    ext = _ext_px(hh, ww, tx, ty, kr)
    nbytes = [ext * (ADJ_EXT_B + ADJ_UNIT_B * u) + ADJ_TILE_B * hh * ww for u in probe["units"]]
    us = probe["us"]
    avg_s = 1e-6 * sum(us) / len(us)
    bpl = sum(nbytes) / len(nbytes)
    return dict(avg_us=1e6 * avg_s, bytes_per_launch=bpl, achieved_gbs=bpl / avg_s / 1e9, launches=len(us),
                tile=[tx, ty, kr, nt], size=[hh, ww], ext_px=ext, units=sorted(set(probe["units"])))


def cpu_adjoint_baseline(pair, tvf, cvf, spec):
    """The float64 restatement of the solve (oracle/vfs_torch_ref.py) differentiated by torch
    autograd on the host: after an untimed warm-up run, four bounded runs of one pair at full
    size with 1 or 2 warps per level and 2 or 6 iterations per warp are fitted (least squares)
    to T(w, n) = A + w (B + n D), which is extrapolated to the workload's warps and
    iterations (labelled extrapolated)."""
    import torch
    from oracle import vfs_oracle as O
    from oracle import vfs_torch_ref as R
    host = _host_info()
    nt = int(host["affinity"] or 1)
    prev = torch.get_num_threads()
    torch.set_num_threads(nt)
    try:
        hh, ww = pair["img_ref"].shape
        sizes = O.compute_pyramid_sizes(hh, ww, spec["nLevel"], spec.get("fScale", 2.0))
        masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pair["mask"], sizes)]
        trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
        t = torch.as_tensor(np.asarray(pair["t"], np.float64).reshape(3))
        z_gt = torch.as_tensor(pair["z_gt"])
        theta = [float(spec[k]) for k in ("Lambda", "alpha0", "alpha1", "beta", "gamma")]

        def run(n_warp, n_iter):
            t0 = time.perf_counter()
            r0 = torch.as_tensor(O.cast_and_scale(pair["img_ref"])).requires_grad_(True)
            ow = torch.as_tensor(O.warping_image(O.cast_and_scale(pair["img_other"]), cvf))
            th = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in theta]
            flow = R.solve_stereo(r0, ow, masks, trs, sizes, *th, n_warp, n_iter, f_scale=spec.get("fScale", 2.0))
            R.depth_loss(flow, pair["cam"], t, masks[0], z_gt).backward()
            return time.perf_counter() - t0

        run(1, 1)
        pts = [(wn, n, run(wn, n)) for wn, n in ((1, 2), (1, 6), (2, 2), (2, 6))]
    finally:
        torch.set_num_threads(prev)
    A_ = np.array([[1.0, wn, wn * n] for wn, n, _ in pts])
    a_, b, d = np.linalg.lstsq(A_, np.array([t_ for _, _, t_ in pts]), rcond=None)[0]
    full = float(a_ + spec["nWarpIters"] * (b + spec["nSolverIters"] * d))
    runs = ", ".join(f"{wn} x {n}: {t_:.1f} s" for wn, n, t_ in pts)
    return dict(value=1.0 / full, unit="pair gradients/s", cores=nt, kind="port",
                sample=(f"float64 restatement of the solve (oracle/vfs_torch_ref.py) with torch autograd on {nt} "
                        f"threads, one {hh}x{ww} pair, {spec['nLevel']} levels, warps x iterations per level "
                        f"({runs}), fitted to A + w (B + n D) (A {a_:.2f} s, B {b:.2f} s, D {d:.3f} s) and "
                        f"extrapolated to {spec['nWarpIters']} warps x {spec['nSolverIters']} iterations: {full:.0f} s"),
                host=host)


def adjoint_leg(n_pairs, steps, warmup, seed=0, compare=True, roofline=True, cpu_baseline=True):
    """Gradient of the depth loss (lambda, alpha0, alpha1, beta, gamma, reference image) for
    `n_pairs` seeded 800x800 pairs, 3 levels x 10 warps x 50 iterations, checkpoint every 25.
    A step is forward + backward of every pair.  The plain (per-kernel, graph) and blocked
    forward solves of the same schedule are timed beside it for the cost ratio."""
    import torch
    import synthetic as S
    import vfs_adjoint as A

    name = "adjoint_800"
    w = S.WORKLOADS[name]
    n_pairs = min(n_pairs, w["pairs"]) if n_pairs else w["pairs"]
    log(f"[bench] adjoint: {n_pairs} pairs")
    ctxs, adjs = [], []
    for i in range(n_pairs):
        ctx = setup_solver(name, seed + i, "graph", device="cuda")
        pr = ctx["pair"]
        adjs.append(A.VFSDepthAdjoint(ctx["prm"], pr["mask"], ctx["tvf"], ctx["cvf"], pr["cam"], pr["t"],
                                      pr["z_gt"], checkpoint=w["checkpoint"]))
        ctxs.append(ctx)
    refs0 = [c["img_ref"].to(torch.float32) / 256.0 for c in ctxs]

    def clock(fn, reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    v = ctxs[0]["vfs"]
    res = {}
    t_fwd = t_bwd = None
    if compare:
        for solver in ("graph", "blocked"):
            v.solver = solver
            for _ in range(warmup):
                v.solve_stereo(ctxs[0]["img_ref"], ctxs[0]["img_other"])
            res[solver] = clock(lambda: v.solve_stereo(ctxs[0]["img_ref"], ctxs[0]["img_other"]), max(steps, 3))
            log(f"[bench] adjoint: forward solve ({solver}): {1e3 * res[solver]:.2f} ms")
        v.release_plan()
        adj = adjs[0]
        for _ in range(warmup):
            adj.forward(refs0[0], ctxs[0]["img_other"])
            adj.backward()
        t_fwd = clock(lambda: adj.forward(refs0[0], ctxs[0]["img_other"]), max(steps, 3))
        t_bwd = clock(adj.backward, max(steps, 3))
        log(f"[bench] adjoint: forward (with saves) {1e3 * t_fwd:.2f} ms, backward {1e3 * t_bwd:.2f} ms")

    losses, grads = [], []
    items = [(adjs[i], refs0[i], ctxs[i]["img_other"]) for i in range(n_pairs)]

    def step():
        losses.clear()
        grads.clear()
        for loss, g in A.gradients(items, streams=ADJ_STREAMS):
            losses.append(loss)
            grads.append([g[k] for k in A.PARAMS])

    log(f"[bench] adjoint: timing {steps} steps of {n_pairs} pair gradients on {ADJ_STREAMS} streams")
    for _ in range(warmup):
        step()
    dt = clock(step, steps)
    gmean = np.mean(np.asarray(grads), axis=0)
    roof = None
    if roofline:
        log("[bench] adjoint: reverse-sweep kernel roofline pass (HIP events)")
        r = adjoint_roofline(adjs[0], refs0[0], ctxs[0]["img_other"], 4096)
        roof = dict(bound="hbm", kernel="k_adj_block", achieved=round(r["achieved_gbs"], 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(r["achieved_gbs"] / HBM_PEAK_GBS, 4),
                    traffic=measured_adj_traffic(r["tile"]) if r["size"] == [800, 800] else None,
                    avg_launch_us=round(r["avg_us"], 3), bytes_per_launch=round(r["bytes_per_launch"]),
                    launches=r["launches"], tile=r["tile"], size=r["size"], units_per_launch=r["units"],
                    stream_copy_gbs=round(stream_copy_gbs(), 1))
    cpu = None
    if cpu_baseline:
        log("[bench] adjoint: CPU baseline (float64 restatement with autograd)")
        pr = ctxs[0]["pair"]
        cpu = cpu_adjoint_baseline(pr, ctxs[0]["tvf"], np.asarray(ctxs[0]["cvf"], np.float32), w)
    for c in ctxs:
        c["vfs"].release_plan()
    return dict(value=round(n_pairs / dt, 3), unit="pair gradients/s", pairs=n_pairs, steps=steps,
                ms_per_step=round(1e3 * dt, 3), streams=ADJ_STREAMS,
                config={"workload": name, "size": [w["H"], w["W"]], "levels": w["nLevel"],
                        "warps_per_level": w["nWarpIters"], "pd_iters_per_warp": w["nSolverIters"],
                        "checkpoint": w["checkpoint"], "pairs": n_pairs},
                forward_solve_ms={k: round(1e3 * t, 3) for k, t in res.items()},
                adjoint_forward_ms=None if t_fwd is None else round(1e3 * t_fwd, 3),
                adjoint_backward_ms=None if t_bwd is None else round(1e3 * t_bwd, 3),
                backward_vs_plain_forward=None if t_bwd is None else round(t_bwd / res["graph"], 3),
                gradient_vs_plain_forward=None if t_bwd is None else round((t_fwd + t_bwd) / res["graph"], 3),
                gradient_vs_blocked_forward=None if t_bwd is None else round((t_fwd + t_bwd) / res["blocked"], 3),
                mean_loss=round(float(np.mean(losses)), 6),
                mean_gradient={k: float(x) for k, x in zip(A.PARAMS, gmean)},
                roofline=roof, cpu_baseline=cpu)


def run_adjoint(a):
    """The gradient workload alone (--workload adjoint_800)."""
    r = adjoint_leg(a.pairs, a.steps, a.warmup, seed=a.seed, compare=not a.no_compare,
                    roofline=not a.no_roofline, cpu_baseline=not a.no_cpu_baseline)
    out = {"metric": METRIC, "value": r.pop("value"), "unit": r.pop("unit"), "n_gpus": 1, "steps": a.steps,
           "warmup": a.warmup, "ms_per_step": r.pop("ms_per_step"), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic", "config": r.pop("config")}
    r.pop("pairs")
    r.pop("steps")
    out.update(r)
    print(json.dumps(out), flush=True)


def run_probe(a):
    """Launches for the hardware-counter runs (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE):
    three plain copies of 2^26 floats (256 MiB read, 256 MiB written: the counters'
    scale), then, per pyramid level of the workload, two warps of the blocked
    primal-dual kernel with the level's autotuned variant (the first warp warms up)."""
    import torch
    import autotune
    import synthetic as S
    import vfs_hip as H
    n = 1 << 26
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        H.call("vfs_copy", H.stream(), H.ptr(dst), H.ptr(src), n)
    torch.cuda.synchronize()
    w = S.WORKLOADS[a.workload]
    sizes = [(w["H"], w["W"])]
    for _ in range(w["nLevel"] - 1):
        sizes.append((int(sizes[-1][0] / 2.0), int(sizes[-1][1] / 2.0)))
    for hh, ww in sizes:
        tile = autotune.tile_for(hh, ww, w["nSolverIters"])
        autotune.time_variant(hh, ww, w["nSolverIters"], tile, reps=1)
        log(f"[probe] {hh}x{ww} tile {tile}")
    torch.cuda.synchronize()
    print(json.dumps({"probe": a.workload, "copy_floats": n, "sizes": sizes}), flush=True)


def _host_info():
    try:
        aff = len(os.sched_getaffinity(0))
    except Exception:
        aff = os.cpu_count()
    cpu = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(cpu_count=os.cpu_count(), affinity=aff, cpu_model=cpu)


def _oracle_params(name_or_spec, hh, ww, n_warp=None):
    from oracle import vfs_oracle as O
    import synthetic as S
    w = S.WORKLOADS[name_or_spec] if isinstance(name_or_spec, str) else name_or_spec
    prm = O.OracleParams(hh, ww)
    S.apply_workload_params(prm, w)
    if n_warp is not None:
        prm.nWarpIters = n_warp
    return prm


def cpu_oracle_baseline(ctx, name, gpu_flow=None, gpu_z=None):
    """The float64 C restatement of the oracle (oracle/vfs_oracle.c) on the host: one whole
    solve of the same pair and schedule plus triangulation, on every core of the process's
    affinity (OpenMP over image rows; the flow does not depend on the thread count), and a
    bounded single-thread sample (1 warp per level, extrapolated).  Returns the baseline and,
    given the timed GPU flow and depth, their deviation from the oracle's over the mask
    (|du| in pixels; relative depth difference where both depths are finite and positive)."""
    from oracle import vfs_oracle as O
    from oracle import vfs_oracle_c as C
    import synthetic as S
    w = S.WORKLOADS[name]
    pair = ctx["pair"]
    hh, ww = pair["img_ref"].shape
    host = _host_info()
    nt = int(host["affinity"] or 1)
    args = (pair["img_ref"], pair["img_other"], pair["mask"], ctx["tvf"], np.asarray(ctx["cvf"], np.float32))
    t0 = time.perf_counter()
    flow = C.solve_stereo(*args, _oracle_params(name, hh, ww), threads=nt)
    xyz = O.triangulate_flow(flow, pair["cam"], pair["t"].reshape(3), pair["mask"])
    dt = time.perf_counter() - t0
    t1 = time.perf_counter()
    C.solve_stereo(*args, _oracle_params(name, hh, ww, n_warp=1), threads=1)
    dt1 = (time.perf_counter() - t1) * w["nWarpIters"]
    vs = None
    if gpu_flow is not None:
        m = pair["mask"] > 0
        d = np.linalg.norm(np.asarray(gpu_flow, np.float64) - flow, axis=-1)[m]
        vs = dict(median_abs_du_px=float(np.median(d)), p99_abs_du_px=float(np.percentile(d, 99)),
                  max_abs_du_px=float(d.max()))
        if gpu_z is not None:
            zo, zg = xyz[..., 2][m], np.asarray(gpu_z, np.float64)[m]
            ok = np.isfinite(zo) & np.isfinite(zg) & (zo > 0) & (zg > 0)
            vs["median_rel_depth_diff"] = float(np.median(np.abs(zg[ok] - zo[ok]) / zo[ok]))
            vs["oracle_median_rel_depth_err_vs_gt"] = float(np.median(
                np.abs(zo[ok] - pair["z_gt"][m][ok]) / pair["z_gt"][m][ok]))
    cpu = dict(value=1.0 / dt, unit="depth maps/s", cores=nt, kind="port",
               sample=(f"float64 C restatement of the oracle (oracle/vfs_oracle.c), one whole {hh}x{ww} solve "
                       f"({w['nLevel']} levels x {w['nWarpIters']} warps x {w['nSolverIters']} iterations) and "
                       f"triangulation on {nt} threads: {dt:.1f} s"),
               single_thread=dict(value=1.0 / dt1, unit="depth maps/s", cores=1,
                                  sample=(f"1 thread, 1 of {w['nWarpIters']} warps per level: {dt1 / w['nWarpIters']:.1f} s, "
                                          f"extrapolated x{w['nWarpIters']}")),
               host=host)
    return cpu, vs


def cpu_batch_baseline(pairs_by_size, spec):
    """The C oracle on the batched workload: one pair of each size class with 1 warp per
    level on every core, extrapolated to the whole schedule and to the 16 pairs of each
    class (labelled extrapolated)."""
    from oracle import vfs_oracle as O
    from oracle import vfs_oracle_c as C
    host = _host_info()
    nt = int(host["affinity"] or 1)
    total, parts = 0.0, []
    for (hh, ww), (p, tvf, cvf, count) in pairs_by_size.items():
        t0 = time.perf_counter()
        flow = C.solve_stereo(p["img_ref"], p["img_other"], p["mask"], tvf, cvf,
                              _oracle_params(spec, hh, ww, n_warp=1), threads=nt)
        O.triangulate_flow(flow, p["cam"], np.asarray(p["t"]).reshape(3), p["mask"])
        dt = time.perf_counter() - t0
        total += dt * spec["nWarpIters"] * count
        parts.append(f"{ww}x{hh} {dt:.2f} s")
    n = sum(c for (_, _, _, c) in pairs_by_size.values())
    return dict(value=n / total, unit="depth maps/s", cores=nt, kind="port",
                sample=(f"float64 C restatement of the oracle on {nt} threads, one pair of each size with 1 of "
                        f"{spec['nWarpIters']} warps per level ({', '.join(parts)}), extrapolated x{spec['nWarpIters']} "
                        f"and to {n} pairs"), host=host)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain_full_800",
                    choices=["plain_full_800", "plain_small_256", "default_800", "batched_64", "adjoint_800",
                             "fit_512"])
    ap.add_argument("--pairs", type=int, default=None,
                    help="batched_64 / adjoint_800: number of pairs (the first n of the workload's 64 / 8)")
    ap.add_argument("--solver", default="blocked", choices=["graph", "eager", "blocked", "blocked_eager"])
    ap.add_argument("--no-compare", action="store_true", help="skip timing the plain per-kernel path beside it")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-pair batched leg of the default run")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient leg of the default run")
    ap.add_argument("--no-fit", action="store_true", help="skip the parameter-fit leg of the default run")
    ap.add_argument("--save-flow", default=None, help="write the last step's flow (H, W, 2) to this .npy file")
    ap.add_argument("--probe", action="store_true",
                    help="only the launches of the hardware-counter runs (copy kernel + blocked kernel per level)")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: every pair is independent and this build runs on a single MI355X; "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch
    import synthetic as S
    torch.cuda.set_device(0)
    if a.workload == "batched_64":
        run_batch(a)
        return
    if a.workload == "adjoint_800":
        run_adjoint(a)
        return
    if a.workload == "fit_512":
        run_fit(a)
        return
    if a.probe:
        run_probe(a)
        return
    w = S.WORKLOADS[a.workload]
    log(f"[bench] workload {a.workload}: generating the pair and vector fields")
    ctx = setup_solver(a.workload, a.seed, a.solver)
    v = ctx["vfs"]
    sizes = v.sizes
    px = sum(h * ww for h, ww in sizes)
    px_iter = px * w["nWarpIters"] * w["nSolverIters"]
    bytes_map = sum(h * ww * (w["nWarpIters"] * (w["nSolverIters"] * B_ITER + B_WARP) + B_LEVEL) for h, ww in sizes)

    def step():
        v.solve_stereo(ctx["img_ref"], ctx["img_other"])
        return v.triangulate_flow(ctx["cam"], ctx["pair"]["t"])

    def timed(label):
        log(f"[bench] {label}: warmup {a.warmup}")
        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        log(f"[bench] {label}: timing {a.steps} steps")
        t0 = time.perf_counter()
        for _ in range(a.steps):
            out = step()
        torch.cuda.synchronize()
        return time.perf_counter() - t0, out

    dt, xyz = timed(a.solver)
    maps_s = a.steps / dt
    launch_mode = v.launch_mode
    flow_main = v.warpUV.cpu().numpy()
    plain = None
    if not a.no_compare and a.solver not in ("graph", "eager"):
        v.solver = "graph"
        dtp, _ = timed("plain per-kernel")
        same = bool(np.array_equal(v.warpUV.cpu().numpy(), flow_main))
        plain = dict(value=round(a.steps / dtp, 3), ms_per_step=round(1e3 * dtp / a.steps, 3),
                     speedup_of_reported=round(dtp / dt, 2), flow_bitwise_equal=same,
                     launches=v.launch_count(), launch_mode=v.launch_mode)
        v.solver = a.solver

    # accuracy of the timed result against the synthetic ground truth (not a parity check)
    pair = ctx["pair"]
    m = pair["mask"] > 0
    z = xyz[..., 2].cpu().numpy()[m]
    zg = pair["z_gt"][m]
    ok = np.isfinite(z) & (zg > 1e-3)
    depth_err = float(np.median(np.abs(z[ok] - zg[ok]) / zg[ok]))

    roof = None
    if not a.no_roofline:
        if a.solver in ("blocked", "blocked_eager"):
            log("[bench] blocked-kernel roofline pass (HIP events)")
            b = v._buffers()
            lv_data = [dict(mask=v.masks[i], abc=b["levels"][i]["abc"], eta=b["levels"][i]["eta"],
                            deriv=b["levels"][i]["deriv"]) for i in range(len(sizes))]
            r = blocked_kernel_roofline(sizes, v._tiles(), w["nWarpIters"], w["nSolverIters"], lv_data)
            kname = "k_pd_block"
        else:
            log("[bench] dual-kernel roofline pass (HIP events)")
            r = dual_kernel_roofline(sizes, w["nWarpIters"], w["nSolverIters"], v.param)
            kname = "k_update_dual"
        traffic = measured_traffic(r["per_level"]) if kname == "k_pd_block" else None
        copy_gbs = stream_copy_gbs()
        roof = dict(bound="hbm", kernel=kname, achieved=round(r["achieved_gbs"], 1), peak=HBM_PEAK_GBS,
                    unit="GB/s", frac=round(r["achieved_gbs"] / HBM_PEAK_GBS, 4), traffic=traffic,
                    avg_launch_us=round(r["avg_us"], 3), bytes_per_launch=round(r["bytes_per_launch"]),
                    stream_copy_gbs=round(copy_gbs, 1), frac_of_stream_copy=round(r["achieved_gbs"] / copy_gbs, 4),
                    per_level=r["per_level"])

    cpu = vs_oracle = None
    if not a.no_cpu_baseline:
        log("[bench] CPU oracle baseline")
        cpu, vs_oracle = cpu_oracle_baseline(ctx, a.workload, flow_main, xyz[..., 2].cpu().numpy())

    # the batched workload beside the headline: one timed batch of the 64 pairs; and the
    # gradient workload on 2 of its 8 pairs
    batched = adjoint = None
    if a.workload == "plain_full_800" and not a.no_batch:
        v.release_plan()
        batched = batch_leg(64, steps=1, warmup=1, compare=not a.no_compare, roofline=not a.no_roofline,
                            cpu_baseline=not a.no_cpu_baseline)
    if a.workload == "plain_full_800" and not a.no_adjoint:
        v.release_plan()
        torch.cuda.empty_cache()
        adjoint = adjoint_leg(2, steps=2, warmup=1, seed=a.seed, compare=not a.no_compare,
                              roofline=not a.no_roofline, cpu_baseline=not a.no_cpu_baseline)
    fitted = None
    if a.workload == "plain_full_800" and not a.no_fit:
        torch.cuda.empty_cache()
        fitted = fit_leg()

    out = {
        "metric": METRIC,
        "value": round(maps_s, 3),
        "unit": "depth maps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(1e3 * dt / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": a.workload, "size": [w["H"], w["W"]], "levels": w["nLevel"],
                   "level_ratio": w.get("fScale", 2.0),
                   "warps_per_level": w["nWarpIters"], "pd_iters_per_warp": w["nSolverIters"],
                   "solver": a.solver, "launch_mode": launch_mode, "launches_per_map": v.launch_count(),
                   "pairs": 1,
                   "pd_tiles": [list(t) for t in v._tiles()] if a.solver.startswith("blocked") else None},
        "plain_per_kernel": plain,
        "mpx_iter_per_s": round(px_iter * maps_s / 1e6, 1),
        "solver_effective_gbs": round(bytes_map * maps_s / 1e9, 1),
        "median_rel_depth_err_vs_gt": round(depth_err, 4),
        "vs_oracle": vs_oracle,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched_64": batched,
        "adjoint_800": adjoint,
        "fit_512": fitted,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
