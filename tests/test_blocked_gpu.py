"""The temporally blocked primal-dual loop against the per-kernel loop and the oracle.

Both evaluate the same per-pixel arithmetic (csrc/vfs_pd.h), so k blocked
iterations must reproduce k iterations of the three per-kernel launches to the
last bit, for every compiled tile variant, on ragged masks, image sizes that
are not multiples of the tile, and masks that reach the image border.
"""

import ctypes

import numpy as np
import pytest

from oracle import vfs_oracle as O
from conftest import flow_diff_stats

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


class _Prm:
    beta, gamma, alpha0, alpha1, Lambda, limitRange, fScale = 10.0, 0.5, 5.0, 1.0, 50.0, 0.1, 2.0


def _state(T, rng, Hh, Ww, border_mask):
    dev = lambda a: T.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    m = (rng.random((Hh, Ww)) > 0.12).astype(np.uint8) * 255
    if not border_mask:
        m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = 0
    f32 = np.float32
    return dict(
        mask=dev(m),
        abc=dev(rng.random((Hh, Ww, 3)).astype(f32)),
        eta=dev((rng.random((Hh, Ww, 3)) * 20 + 1).astype(f32)),
        deriv=dev((rng.normal(size=(Hh, Ww, 2)) * 0.2).astype(f32)),
        u=dev(rng.normal(size=(Hh, Ww)).astype(f32)),
        ub=dev(rng.normal(size=(Hh, Ww)).astype(f32)),
        vb=dev((rng.normal(size=(Hh, Ww, 2)) * 0.1).astype(f32)),
        p=dev((rng.normal(size=(Hh, Ww, 2)) * 0.3).astype(f32)),
        q=dev((rng.normal(size=(Hh, Ww, 4)) * 0.3).astype(f32)),
    )


def _plain_iters(T, st, n, first, taus, sigmas, mus, prm):
    import vfs_hip as H
    Hh, Ww = st["u"].shape
    u, ub, vb = st["u"].clone(), st["ub"].clone(), st["vb"].clone()
    p = T.zeros_like(st["p"]) if first else st["p"].clone()
    q = T.zeros_like(st["q"]) if first else st["q"].clone()
    if first:
        vb.zero_()
    tp = T.empty_like(p)
    s = H.stream()
    for i in range(n):
        H.call("vfs_update_dual", s, H.ptr(p), H.ptr(tp), H.ptr(q), float(sigmas[i]), H.ptr(ub), H.ptr(vb),
               H.ptr(st["abc"]), H.ptr(st["mask"]), Hh, Ww, ctypes.byref(prm))
        H.call("vfs_l1_thresholding", s, H.ptr(u), float(taus[i]), H.ptr(tp), H.ptr(u), H.ptr(ub), H.ptr(st["deriv"]),
               H.ptr(st["eta"]), H.ptr(st["mask"]), Hh, Ww, ctypes.byref(prm))
        H.call("vfs_update_primal", s, H.ptr(ub), H.ptr(vb), float(taus[i]), float(mus[i]), H.ptr(p), H.ptr(q),
               H.ptr(u), H.ptr(ub), H.ptr(vb), H.ptr(st["abc"]), H.ptr(st["eta"]), H.ptr(st["mask"]), Hh, Ww,
               ctypes.byref(prm))
    return dict(u=u, ub=ub, vb=vb, p=p, q=q)


def _blocked_iters(T, st, n, first, taus, sigmas, mus, prm, variant):
    import vfs_hip as H
    Hh, Ww = st["u"].shape
    out = {k: T.full_like(st[k], float("nan")) for k in ("u", "ub", "vb", "p", "q")}
    fp = lambda a: np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    t_, s_, m_ = fp(taus[:n]), fp(sigmas[:n]), fp(mus[:n])
    H.call("vfs_pd_blocks", H.stream(), Hh, Ww, H.ptr(st["mask"]), H.ptr(st["abc"]), H.ptr(st["eta"]),
           H.ptr(st["deriv"]), H.ptr(st["u"]), H.ptr(st["ub"]), H.ptr(st["vb"]), H.ptr(st["p"]), H.ptr(st["q"]),
           H.ptr(out["u"]), H.ptr(out["ub"]), H.ptr(out["vb"]), H.ptr(out["p"]), H.ptr(out["q"]), int(first),
           t_, s_, m_, n, variant[0], variant[1], variant[2], variant[3] if len(variant) > 3 else 0,
           ctypes.byref(prm))
    return out


def _variants():
    import sys
    import os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vfs-python_amd"))
    import vfs_hip as H
    try:
        return H.pd_variants()
    except RuntimeError:
        return []


@pytest.mark.parametrize("variant", _variants(), ids=lambda v: f"{v[0]}x{v[1]}k{v[2]}t{v[3]}")
@pytest.mark.parametrize("shape,border,first", [((70, 53), False, 0), ((97, 131), True, 1), ((8, 9), True, 0)])
def test_blocked_iterations_equal_per_kernel_iterations(T, variant, shape, border, first):
    import vfs_hip as H
    rng = np.random.default_rng(hash((variant, shape)) % 2**32)
    st = _state(T, rng, *shape, border)
    prm = H.make_params(_Prm)
    taus, sigmas, mus = H.pd_schedule(20, 1.0)
    k = variant[2]
    for n in sorted({1, k}):
        ref = _plain_iters(T, st, n, first, taus, sigmas, mus, prm)
        got = _blocked_iters(T, st, n, first, taus, sigmas, mus, prm, variant)
        T.cuda.synchronize()
        for name in ("u", "ub", "vb", "p", "q"):
            a, b = got[name].cpu().numpy(), ref[name].cpu().numpy()
            assert np.isfinite(a).all(), (name, n)
            np.testing.assert_array_equal(a, b, err_msg=f"{name} after {n} iterations")


def _solver(small_case, mode, tiles=None):
    import vfs as V
    import vfs_params as P
    import synthetic as S
    pair = small_case["pair"]
    prm = P.vfs_params(pair["img_ref"])
    S.apply_workload_params(prm, "plain_small_256")
    v = V.VFS(prm)
    v.solver = mode
    v.pd_tiles = tiles
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(pair["mask"])
    v.create_vector_field_pyramid(small_case["tvf"], small_case["cvf"])
    return v


@pytest.mark.parametrize("tiles", [None, {0: (32, 16, 8), 1: (16, 16, 1), 2: (64, 16, 2)}])
def test_blocked_solve_equals_plain_solve(T, small_case, small_oracle, tiles):
    pair = small_case["pair"]
    plain = _solver(small_case, "graph").solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()
    for mode in ("blocked", "blocked_eager"):
        f = _solver(small_case, mode, tiles).solve_stereo(pair["img_ref"], pair["img_other"]).cpu().numpy()
        np.testing.assert_array_equal(f, plain)
    st = flow_diff_stats(f, small_oracle["flow"], pair["mask"])
    print("blocked solve vs oracle:", st)
    assert st["median"] <= 1e-4
