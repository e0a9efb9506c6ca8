"""Autotuner of the temporally blocked primal-dual kernel (tile shape x k).

For each pyramid-level size the tuner times every compiled variant
(``vfs_hip.pd_variants()``: tile {16, 32, 64} x {16, 32} x k {1, 2, 4, 8}) on one
warp's worth of iterations (ceil(n_iter / k) launches, ping-ponged state, HIP
events on the launch stream) and keeps the fastest.  Results are stored in a
JSON table next to this file (``autotune_gfx950.json``) and read back by
``tile_for``; sizes missing from the table use the entry with the nearest pixel
count, or a fixed default before any tuning.

Usage on the GPU:  python autotune.py --sizes 800x800,400x400,... --iters 50
"""

import argparse
import ctypes
import json
import math
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, "autotune_gfx950.json")

_table = None
_batch_table = None


def _load():
    global _table, _batch_table
    if _table is None:
        _table, _batch_table = {}, {}
        if os.path.exists(TABLE):
            with open(TABLE) as f:
                doc = json.load(f)
            for row in doc.get("entries", []):
                _table[(int(row["h"]), int(row["w"]), int(row["iters"]))] = variant(row["best"])
            for row in doc.get("batch_entries", []):
                _batch_table[(int(row["pixels"]), int(row["iters"]))] = variant(row["best"])
    return _table


def variant(t):
    """(tile_x, tile_y, k, threads) from a 3- or 4-tuple; threads 0 selects the first
    compiled thread count of that tile and k."""
    t = tuple(int(x) for x in t)
    if len(t) == 3:
        return t + (0,)
    if len(t) != 4:
        raise ValueError(f"a blocked variant is (tile_x, tile_y, k[, threads]), got {t}")
    return t


def default_tile(h, w):
    """Choice before any tuning: larger k where launches dominate (small levels)."""
    n = h * w
    if n >= 250_000:
        return (32, 32, 4, 1024)
    if n >= 40_000:
        return (32, 16, 8, 1024)
    return (16, 16, 8, 1024)


def tile_for(h, w, n_iter):
    """(tile_x, tile_y, k, threads) for a level of h x w pixels and n_iter iterations per warp."""
    t = _load()
    if (h, w, n_iter) in t:
        return t[(h, w, n_iter)]
    if t:
        best = min(t.items(), key=lambda kv: (abs(kv[0][0] * kv[0][1] - h * w), abs(kv[0][2] - n_iter)))
        return best[1]
    return default_tile(h, w)


def tile_for_pixels(n, n_iter):
    """The variant for one launch over n pixels spread across many pairs: the batch
    entry (measured on whole batched solves) whose pixel count is nearest to n on a log
    scale, for n within the batch entries' range; else the single-level entry nearest
    to n."""
    t = _load()
    if _batch_table and n >= min(k[0] for k in _batch_table) // 2:
        best = min(_batch_table.items(), key=lambda kv: (abs(math.log(kv[0][0]) - math.log(max(n, 1))),
                                                         abs(kv[0][1] - n_iter)))
        return best[1]
    if not t:
        side = int(math.sqrt(n))
        return default_tile(side, side)
    best = min(t.items(), key=lambda kv: (abs(math.log(kv[0][0] * kv[0][1]) - math.log(max(n, 1))),
                                          abs(kv[0][2] - n_iter)))
    return best[1]


def time_variant(h, w, n_iter, var, reps=5, seed=0, graph=True, level=None):
    """Mean time (ms) of one warp's n_iter blocked iterations at h x w with the variant ``var``.

    With ``graph`` (as the solver runs them) the warp's launches are recorded once into
    a HIP graph and the graph is replayed between events on the launch stream, so the
    time is that of the kernels, not of issuing them from Python.  ``level`` (a dict of
    mask (h, w) uint8, abc, eta (h, w, 3), deriv (h, w, 2) device tensors) replaces the
    seeded random level data, e.g. with a solved pair's own."""
    import numpy as np
    import torch
    import vfs_hip as H

    class P:
        beta, gamma, alpha0, alpha1, Lambda, limitRange, fScale = 10.0, 0.5, 5.0, 1.0, 50.0, 0.1, 2.0

    prm = H.make_params(P)
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    dev = dict(device="cuda")
    if level is not None:
        mask, abc, eta, deriv = (level[n].contiguous() for n in ("mask", "abc", "eta", "deriv"))
    else:
        mask = (torch.rand((h, w), generator=g, **dev) > 0.1).to(torch.uint8)
        abc = torch.rand((h, w, 3), generator=g, **dev)
        eta = torch.rand((h, w, 3), generator=g, **dev) * 20 + 1
        deriv = torch.randn((h, w, 2), generator=g, **dev) * 0.1
    sets = []
    for _ in range(2):
        sets.append(dict(u=torch.zeros((h, w), **dev), ub=torch.zeros((h, w), **dev), vb=torch.zeros((h, w, 2), **dev),
                         p=torch.zeros((h, w, 2), **dev), q=torch.zeros((h, w, 4), **dev)))
    u0 = torch.zeros((h, w), **dev)
    taus, sigmas, mus = H.pd_schedule(n_iter, 1.0)
    tx, ty, k, nt = variant(var)
    s = H.stream()
    fn = H.lib().vfs_pd_blocks

    def run_warp(s=s):
        cur, done, first = 0, 0, True
        while done < n_iter:
            kc = min(k, n_iter - done)
            a, b = sets[cur], sets[1 - cur]
            sl = slice(done, done + kc)
            rc = fn(s, h, w, H.ptr(mask), H.ptr(abc), H.ptr(eta), H.ptr(deriv), H.ptr(u0 if first else a["u"]),
                    H.ptr(a["ub"]), H.ptr(a["vb"]), H.ptr(a["p"]), H.ptr(a["q"]), H.ptr(b["u"]), H.ptr(b["ub"]),
                    H.ptr(b["vb"]), H.ptr(b["p"]), H.ptr(b["q"]), 1 if first else 0,
                    np.ascontiguousarray(taus[sl]).ctypes.data_as(ctypes.c_void_p),
                    np.ascontiguousarray(sigmas[sl]).ctypes.data_as(ctypes.c_void_p),
                    np.ascontiguousarray(mus[sl]).ctypes.data_as(ctypes.c_void_p), kc, tx, ty, k, nt,
                    ctypes.byref(prm))
            H.check(rc, "vfs_pd_blocks")
            cur, done, first = 1 - cur, done + kc, False

    plan = None
    if graph:
        cs = ctypes.c_void_p()
        H.call("vfs_capture_begin", ctypes.byref(cs))
        run_warp(cs)
        plan = ctypes.c_void_p()
        H.call("vfs_capture_end", cs, ctypes.byref(plan))

        def step():
            H.call("vfs_plan_launch", plan, s)
    else:
        step = run_warp
    try:
        step()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            step()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps
    finally:
        if plan is not None:
            H.lib().vfs_plan_destroy(plan)


def tune(sizes, n_iter, reps=5, rounds=3, log=print):
    """Per size, `rounds` passes over every variant (interleaved, so clock or thermal
    drift affects all variants alike), each timing `reps` warps; a variant's time is
    its best pass."""
    import vfs_hip as H
    variants = H.pd_variants()
    entries = []
    for (h, w) in sizes:
        best_ms = {}
        for _ in range(rounds):
            for v in variants:
                ms = time_variant(h, w, n_iter, v, reps=reps)
                best_ms[v] = min(ms, best_ms.get(v, float("inf")))
        results = sorted((ms, v) for v, ms in best_ms.items())
        for ms, v in results:
            log(f"[autotune] {h}x{w} iters {n_iter} tile {v[0]}x{v[1]} k {v[2]} threads {v[3]}: {ms * 1e3:.1f} us/warp")
        best = results[0][1]
        entries.append(dict(h=h, w=w, iters=n_iter, best=list(best), us_per_warp=round(results[0][0] * 1e3, 2),
                            all=[dict(tile=list(v), us=round(ms * 1e3, 2)) for ms, v in results]))
        log(f"[autotune] {h}x{w}: best tile {best[0]}x{best[1]} k {best[2]} ({results[0][0] * 1e3:.1f} us/warp)")
    return entries


def save(entries, path=TABLE):
    doc = {"arch": "gfx950"}
    if os.path.exists(path):
        with open(path) as f:
            doc = json.load(f)
    keep = {(e["h"], e["w"], e["iters"]): e for e in doc.get("entries", [])}
    for e in entries:
        keep[(e["h"], e["w"], e["iters"])] = e
    doc["entries"] = sorted(keep.values(), key=lambda e: -e["h"] * e["w"])
    with open(path, "w") as f:
        json.dump(doc, f, indent=1)
    global _table
    _table = None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="800x800,400x400,200x200,100x100,50x50")
    ap.add_argument("--iters", type=int, default=50)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--out", default=TABLE)
    a = ap.parse_args()
    sizes = [tuple(int(x) for x in s.split("x")) for s in a.sizes.split(",")]
    entries = tune(sizes, a.iters, a.reps, a.rounds, log=lambda *m: print(*m, file=sys.stderr, flush=True))
    save(entries, a.out)
    print(json.dumps([dict(h=e["h"], w=e["w"], best=e["best"], us_per_warp=e["us_per_warp"]) for e in entries]))


if __name__ == "__main__":
    if HERE not in sys.path:
        sys.path.insert(0, HERE)
    main()
