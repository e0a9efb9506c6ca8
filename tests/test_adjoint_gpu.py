"""Hand-written adjoint (csrc/vfs_adjoint.hip, vfs_adjoint.py) against two references:
autograd through the unrolled PyTorch float64 restatement (oracle/vfs_torch_ref.py),
and central finite differences of the float64 NumPy oracle.

Tolerances (fp32 state, fp32 adjoint accumulation, against float64 references):
  * N reversed primal-dual iterations: every input adjoint within 1e-4 of the
    autograd gradient relative to the largest entry (same fp32 inputs);
  * full solve (48 x 48, 2 levels x 2 warps x 6 iterations), parameters and reference
    image: relative error <= 1e-3 per parameter (measured 5e-4 to 8.7e-4) and cosine
    similarity >= 0.999 against autograd through the float64 solve;
  * the differentiation workload's schedule (3 levels x 10 warps x 50 iterations) on
    a 128 x 128 crop of its 800 x 800 pair: against central differences of the
    float64 oracle at h = 1e-5 (relative), the gradient vector within 1e-3 relative
    (measured 7e-4); see the test for the finite differences' own convergence;
  * finite differences (central, h = 1e-4 relative) of the float64 oracle loss:
    per parameter within max(2e-2 relative, 3x the deviation of the same gradient
    taken by autograd through the restatement run in float32).  The depth loss
    is ill-conditioned in fp32 on small crops (the float32 restatement is 6% off
    for beta on the 40 x 40 case), the same allowance the flow tolerances make.
"""

import ctypes

import numpy as np
import pytest
import torch

from oracle import vfs_oracle as O
from oracle import vfs_torch_ref as R

pytestmark = pytest.mark.gpu

THETA = dict(Lambda=50.0, alpha0=5.0, alpha1=1.0, beta=10.0, gamma=0.5)


@pytest.fixture(scope="module")
def T():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _rel(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


@pytest.mark.parametrize("shape", [(70, 150), (33, 64), (5, 3)])
def test_derivative_taps_transpose_matches_autograd(T, shape):
    """vfs_adj_derivatives (the taps' shares summed in int64 fixed point per block) against
    torch autograd of the float32 restatement of computeDerivatives
    (oracle/vfs_torch_ref.compute_derivatives) on random images, directions and output
    adjoints, over several 64 x 16 blocks, the image border (mirrored taps) and an image
    smaller than a tap's reach.  The tap coordinates and bilinear weights are the same
    float32 values on both sides (up to the last bit of the normalised direction), so the
    two differ by the rounding of the shares' products and of autograd's float32 sums:
    every entry within 2e-6 of the largest."""
    import vfs_hip as H
    hh, ww = shape
    rng = np.random.default_rng(hh * 1000 + ww)
    ref = rng.random((hh, ww), dtype=np.float32)
    ow = rng.random((hh, ww), dtype=np.float32)
    ang = rng.uniform(0.0, 2.0 * np.pi, (hh, ww))
    mag = rng.uniform(0.5, 2.0, (hh, ww, 1))
    tvf = (np.stack([np.cos(ang), np.sin(ang)], -1) * mag).astype(np.float32)
    gd = rng.normal(size=(hh, ww, 2)).astype(np.float32)
    r = torch.tensor(ref, requires_grad=True)
    o = torch.tensor(ow, requires_grad=True)
    (R.compute_derivatives(o, r, torch.tensor(tvf)) * torch.tensor(gd)).sum().backward()
    gref = T.zeros((hh, ww), device="cuda")
    gow = T.zeros((hh, ww), device="cuda")
    dgd, dtvf = T.tensor(gd, device="cuda"), T.tensor(tvf, device="cuda")
    H.call("vfs_adj_derivatives", H.stream(), H.ptr(gref), H.ptr(gow), H.ptr(dgd), H.ptr(dtvf), hh, ww)
    T.cuda.synchronize()
    for got, ref_grad in ((gref, r.grad), (gow, o.grad)):
        g, e = got.cpu().numpy(), ref_grad.numpy()
        assert np.isfinite(g).all()
        assert _rel(g, e) <= 2e-6, _rel(g, e)


@pytest.mark.parametrize("n_iter", [1, 4])
def test_pd_iterations_adjoint_matches_autograd(T, n_iter):
    import vfs_hip as H
    rng = np.random.default_rng(7 + n_iter)
    Hh, Ww = 23, 31
    f32 = np.float32
    m = (rng.random((Hh, Ww)) > 0.15)
    m[0, :] = False
    st0 = dict(p=rng.normal(size=(Hh, Ww, 2)) * 0.3, q=rng.normal(size=(Hh, Ww, 4)) * 0.3, u=rng.normal(size=(Hh, Ww)),
               ub=rng.normal(size=(Hh, Ww)), vb=rng.normal(size=(Hh, Ww, 2)) * 0.1)
    abc = rng.random((Hh, Ww, 3))
    eta = rng.random((Hh, Ww, 3)) * 20 + 1
    deriv = rng.normal(size=(Hh, Ww, 2)) * 0.2
    st0 = {k: v.astype(f32) for k, v in st0.items()}
    abc, eta, deriv = abc.astype(f32), eta.astype(f32), deriv.astype(f32)
    taus, sigmas, mus = H.pd_schedule(n_iter, 1.0)
    wts = {k: rng.normal(size=v.shape) for k, v in st0.items()}
    lam, a0, a1 = 50.0, 5.0, 1.0

    # float64 autograd
    tt = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in st0.items()}
    tabc = torch.tensor(abc, dtype=torch.float64, requires_grad=True)
    teta = torch.tensor(eta, dtype=torch.float64, requires_grad=True)
    tder = torch.tensor(deriv, dtype=torch.float64, requires_grad=True)
    tpar = [torch.tensor(v, dtype=torch.float64, requires_grad=True) for v in (lam, a0, a1)]
    tm = torch.tensor(m)
    p, q, u, ub, vb = tt["p"], tt["q"], tt["u"], tt["ub"], tt["vb"]
    for i in range(n_iter):
        p, tp, q = R.update_dual(p, q, float(sigmas[i]), ub, vb, tabc, tm, tpar[1], tpar[2])
        u = R.l1_thresholding(float(taus[i]), tp, u, ub, tder, teta, tm, tpar[0])
        ub, vb = R.update_primal(float(taus[i]), float(mus[i]), p, q, u, ub, vb, tabc, teta, tm, tpar[1], tpar[2])
    out = dict(p=p, q=q, u=u, ub=ub, vb=vb)
    loss = sum((out[k] * torch.tensor(wts[k])).sum() for k in out)
    loss.backward()

    # HIP forward (tape) + reverse
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a), device="cuda")  # noqa: E731
    s = H.stream()

    class P:
        beta, gamma, alpha0, alpha1, Lambda, limitRange, fScale = 10.0, 0.5, a0, a1, lam, 0.1, 2.0

    prm = H.make_params(P)
    mask = dev(m.astype(np.uint8) * 255)
    cabc, ceta, cder = dev(abc), dev(eta), dev(deriv)
    st = {k: dev(v) for k, v in st0.items()}
    tape = [{k: v.clone() for k, v in st.items()}]
    tpb = torch.empty_like(st["p"])
    for i in range(n_iter):
        H.call("vfs_update_dual", s, H.ptr(st["p"]), H.ptr(tpb), H.ptr(st["q"]), float(sigmas[i]), H.ptr(st["ub"]),
               H.ptr(st["vb"]), H.ptr(cabc), H.ptr(mask), Hh, Ww, ctypes.byref(prm))
        H.call("vfs_l1_thresholding", s, H.ptr(st["u"]), float(taus[i]), H.ptr(tpb), H.ptr(st["u"]), H.ptr(st["ub"]),
               H.ptr(cder), H.ptr(ceta), H.ptr(mask), Hh, Ww, ctypes.byref(prm))
        H.call("vfs_update_primal", s, H.ptr(st["ub"]), H.ptr(st["vb"]), float(taus[i]), float(mus[i]), H.ptr(st["p"]),
               H.ptr(st["q"]), H.ptr(st["u"]), H.ptr(st["ub"]), H.ptr(st["vb"]), H.ptr(cabc), H.ptr(ceta), H.ptr(mask),
               Hh, Ww, ctypes.byref(prm))
        tape.append({k: v.clone() for k, v in st.items()})
    z = lambda *c: torch.zeros((Hh, Ww, *c), device="cuda")  # noqa: E731
    g = dict(gu=dev(wts["u"].astype(f32)), gub=dev(wts["ub"].astype(f32)), gvb=dev(wts["vb"].astype(f32)),
             gp=dev(wts["p"].astype(f32)), gq=dev(wts["q"].astype(f32)), um=z(3), vm=z(6), qmsg=z(4), tmsg=z(4),
             gabc=z(3), geta=z(3), gderiv=z(2), gpar=z(5))
    has = 0
    for i in range(n_iter - 1, -1, -1):
        ad = H.VfsAdjPd()
        ad.H, ad.W = Hh, Ww
        for name, t in g.items():
            setattr(ad, name, t.data_ptr())
        after, before = tape[i + 1], tape[i]
        for name, t in (("p", after["p"]), ("q", after["q"]), ("u", after["u"]), ("p_prev", before["p"]),
                        ("q_prev", before["q"]), ("u_prev", before["u"]), ("u_bar_prev", before["ub"]),
                        ("v_bar_prev", before["vb"]), ("abc", cabc), ("eta", ceta), ("deriv", cder), ("mask", mask)):
            setattr(ad, name, t.data_ptr())
        H.call("vfs_adj_pd_iteration", s, ctypes.byref(ad), ctypes.byref(prm), float(taus[i]), float(sigmas[i]),
               float(mus[i]), has)
        has = 1
    gub_fin, gvb_fin = z(), z(2)
    H.call("vfs_adj_pd_finish", s, H.ptr(gub_fin), H.ptr(gvb_fin), H.ptr(g["gub"]), H.ptr(g["gvb"]), H.ptr(g["um"]),
           H.ptr(g["vm"]), H.ptr(mask), Hh, Ww)
    torch.cuda.synchronize()
    got = dict(p=g["gp"], q=g["gq"], u=g["gu"], ub=gub_fin, vb=gvb_fin)
    for k in got:
        r = _rel(got[k].cpu().numpy(), tt[k].grad.numpy())
        print(k, r)
        assert r <= 1e-4, k
    assert _rel(g["gabc"].cpu().numpy(), tabc.grad.numpy()) <= 1e-4
    assert _rel(g["geta"].cpu().numpy(), teta.grad.numpy()) <= 1e-4
    assert _rel(g["gderiv"].cpu().numpy(), tder.grad.numpy()) <= 1e-4
    par = g["gpar"].double().sum(dim=(0, 1)).cpu().numpy()
    ref = [t.grad.item() for t in tpar]
    for gi, ri in zip(par[:3], ref):
        assert abs(gi - ri) <= 1e-4 * max(abs(ri), 1e-3) + 1e-3 * max(abs(x) for x in ref)


def _small_case(n=48, seed=5, levels=2):
    import synthetic as S
    pr = S.make_pair(n, n, seed, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=levels)
    cam = pr["cam"]
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((n, n)), cam, cam, pr["R"], pr["t"], 1)
    return pr, S.sanitize_field(tvf), cvf.astype(np.float32)


def _adjoint(pr, tvf, cvf, n_warp, n_iter, ckpt, levels=2, **kw):
    import vfs_adjoint as A
    import vfs_params as P
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = levels, 2.0, n_warp, n_iter
    for k, v in THETA.items():
        setattr(prm, k, v)
    return A.VFSDepthAdjoint(prm, pr["mask"], tvf, cvf, pr["cam"], pr["t"], pr["z_gt"], checkpoint=ckpt, **kw)


def test_solve_adjoint_matches_autograd_through_unrolled_torch(T):
    pr, tvf, cvf = _small_case()
    n, n_warp, n_iter = 48, 2, 6
    adj = _adjoint(pr, tvf, cvf, n_warp, n_iter, ckpt=4)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    loss = adj.forward(ref0, pr["img_other"])
    g = adj.backward()
    # float64 autograd through the same solve
    sizes = O.compute_pyramid_sizes(n, n, 2, 2.0)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pr["mask"], sizes)]
    trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
    r0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"])).requires_grad_(True)
    ow = torch.as_tensor(O.warping_image(O.cast_and_scale(pr["img_other"]), cvf))
    th = [torch.tensor(THETA[k], dtype=torch.float64, requires_grad=True) for k in
          ("Lambda", "alpha0", "alpha1", "beta", "gamma")]
    flow = R.solve_stereo(r0, ow, masks, trs, sizes, *th, n_warp, n_iter)
    lt = R.depth_loss(flow, pr["cam"], torch.as_tensor(pr["t"].reshape(3)), masks[0], torch.as_tensor(pr["z_gt"]))
    lt.backward()
    print("loss hip", loss, "torch", float(lt.detach()))
    assert np.isfinite(float(lt.detach()))
    assert abs(loss - float(lt.detach())) <= 1e-3 * abs(float(lt.detach()))
    for i, k in enumerate(("Lambda", "alpha0", "alpha1", "beta", "gamma")):
        ref = th[i].grad.item()
        print(k, g[k], ref, "relative", abs(g[k] - ref) / abs(ref))
        # measured 5.0e-4 to 8.7e-4 relative
        assert abs(g[k] - ref) <= 1e-3 * abs(ref), k
    gr, rr = g["ref0"].cpu().numpy().ravel(), r0.grad.numpy().ravel()
    cos = float(np.dot(gr, rr) / (np.linalg.norm(gr) * np.linalg.norm(rr)))
    print("ref0 cosine", cos, "rel", np.linalg.norm(gr - rr) / np.linalg.norm(rr))
    assert cos >= 0.999


def _oracle_loss(pr, tvf, cvf, theta, n_warp, n_iter, levels=2):
    n = pr["img_ref"].shape[0]
    prm = O.OracleParams(n, n, nLevel=levels, fScale=2.0, nWarpIters=n_warp, nSolverIters=n_iter, **theta)
    flow = O.solve_stereo(pr["img_ref"], pr["img_other"], pr["mask"], tvf, cvf, prm)
    z = O.triangulate_flow(flow, pr["cam"], pr["t"].reshape(3), pr["mask"])[..., 2]
    m = (pr["mask"] > 0) & (z > 0) & np.isfinite(z) & (pr["z_gt"] > 0)
    return float(np.mean(np.abs(np.log(z[m]) - np.log(pr["z_gt"][m]))))


def _torch_fp32_gradients(pr, tvf, cvf, n_warp, n_iter):
    """Parameter gradients of the same loss by autograd through the restatement run in
    float32 (on the CPU): the spread that fp32 arithmetic alone puts on the gradient."""
    n = pr["img_ref"].shape[0]
    f = torch.float32
    sizes = O.compute_pyramid_sizes(n, n, 2, 2.0)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pr["mask"], sizes)]
    trs = [torch.as_tensor(t).to(f) for t in O.create_vector_field_pyramid(tvf, sizes)]
    r0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"])).to(f)
    ow = torch.as_tensor(O.warping_image(O.cast_and_scale(pr["img_other"]), cvf)).to(f)
    th = [torch.tensor(THETA[k], dtype=f, requires_grad=True) for k in ("Lambda", "alpha0", "alpha1", "beta", "gamma")]
    flow = R.solve_stereo(r0, ow, masks, trs, sizes, *th, n_warp, n_iter)
    R.depth_loss(flow, pr["cam"], torch.as_tensor(pr["t"].reshape(3)).to(f), masks[0],
                 torch.as_tensor(pr["z_gt"]).to(f)).backward()
    return [x.grad.item() for x in th]


def test_solve_adjoint_matches_central_differences_of_the_oracle(T):
    pr, tvf, cvf = _small_case(40, 9)
    n_warp, n_iter = 2, 5
    adj = _adjoint(pr, tvf, cvf, n_warp, n_iter, ckpt=2)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    adj.forward(ref0, pr["img_other"])
    g = adj.backward()
    g32 = _torch_fp32_gradients(pr, tvf, cvf, n_warp, n_iter)
    for i, k in enumerate(("Lambda", "alpha0", "alpha1", "beta", "gamma")):
        h = 1e-4 * THETA[k]
        tp, tm = dict(THETA), dict(THETA)
        tp[k] += h
        tm[k] -= h
        fd = (_oracle_loss(pr, tvf, cvf, tp, n_warp, n_iter) - _oracle_loss(pr, tvf, cvf, tm, n_warp, n_iter)) / (2 * h)
        tol = max(2e-2 * abs(fd), 3.0 * abs(g32[i] - fd)) + 1e-5
        print(k, "adjoint", g[k], "central difference", fd, "fp32 restatement", g32[i])
        assert abs(g[k] - fd) <= tol, k


@pytest.mark.parametrize("blocked", [False, True])
def test_forward_level_driver_matches_plain_solver(T, blocked):
    """vfs_adj_forward_level (per-kernel or blocked warp loop + saves) gives the plain solver's
    flow bit for bit."""
    import vfs as V
    import vfs_params as P
    pr, tvf, cvf = _small_case(64, 3)
    adj = _adjoint(pr, tvf, cvf, 3, 7, ckpt=3, blocked=blocked)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    adj.forward(ref0, pr["img_other"])
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 2, 2.0, 3, 7
    for k, v in THETA.items():
        setattr(prm, k, v)
    v = V.VFS(prm)
    v.solver = "eager"
    v.setup_module()
    v.compute_pyramid_sizes()
    v.create_mask_pyramid(pr["mask"])
    v.create_vector_field_pyramid(tvf, cvf)
    flow = v.solve_stereo(pr["img_ref"], pr["img_other"])
    torch.cuda.synchronize()
    assert torch.equal(flow, adj.flow)


def _assert_same_bits(g0, g1):
    names = ("Lambda", "alpha0", "alpha1", "beta", "gamma")
    for k in names:
        print(k, g0[k], g1[k], "same" if g0[k] == g1[k] else "DIFFERENT")
    d = float((g1["ref0"] - g0["ref0"]).abs().max())
    print("ref0 max diff", d, "max", float(g0["ref0"].abs().max()))
    assert [g1[k] for k in names] == [g0[k] for k in names]
    assert torch.equal(g1["ref0"], g0["ref0"])


def test_gradients_do_not_depend_on_checkpoints_or_kernel_choice(T):
    """Checkpoint interval, per-kernel or blocked forward, per-kernel or blocked (k = 1)
    tape recompute: the states are the same bits, every reverse kernel sums in a fixed
    order (no atomics) and each pixel accumulates its parameter shares in the order of
    the reversed steps, so the gradients are the same bits."""
    pr, tvf, cvf = _small_case(48, 4)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    res = []
    for ck, blocked, rec in ((1, False, None), (4, True, (16, 16, 1)), (9, True, (32, 32, 4)),
                             (4, False, (16, 16, 2)), (9, True, None), (9, True, (16, 16, 8)), (5, True, "auto")):
        adj = _adjoint(pr, tvf, cvf, 2, 9, ckpt=ck, blocked=blocked, rec_tile=rec)
        loss = adj.forward(ref0, pr["img_other"])
        res.append((loss, adj.backward()))
    l0, g0 = res[0]
    for loss, g in res[1:]:
        assert loss == l0
        _assert_same_bits(g0, g)


@pytest.mark.parametrize("rw", [0, 1])
def test_kept_last_segment_tape_gives_the_same_gradients(T, rw):
    """The forward's tape of each reversed warp's last checkpoint segment (the default)
    against recomputing that segment: the same loss and gradient bits, for the exact and
    the truncated gradient, with segments of 4 and 9 iterations (a last segment shorter
    than the checkpoint interval, and one that is the whole warp, written from the
    warp's first launch)."""
    pr, tvf, cvf = _small_case(48, 4)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    for ck in (4, 9):
        res = []
        for keep in (False, True):
            adj = _adjoint(pr, tvf, cvf, 2, 9, ckpt=ck, reverse_warps=rw, tape_last_segment=keep)
            assert adj._tape_warps(0) == (0 if not keep else (rw or 2))
            loss = adj.forward(ref0, pr["img_other"])
            res.append((loss, adj.backward()))
        (l0, g0), (l1, g1) = res
        assert l0 == l1
        _assert_same_bits(g0, g1)


@pytest.mark.parametrize("shape,levels,ratio", [((17, 23), 2, 2.0), ((64, 64), 2, 2.0), ((44, 52), 4, 1.2)])
def test_gradients_are_bitwise_reproducible(T, shape, levels, ratio):
    """Two gradients of the same pair, back to back on fresh objects, are the same bits:
    no reverse kernel adds floats with atomics (the transposes of the tensor, Gaussian,
    pyramid and upsampling are gathers in a fixed order; the derivative taps' shares are
    summed in int64 fixed point, whose sum does not depend on the order)."""
    import synthetic as S
    import vfs_adjoint as A
    import vfs_params as P
    hh, ww = shape
    pr = S.make_pair(hh, ww, 11, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=levels, level_ratio=ratio)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((hh, ww)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
    tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = levels, ratio, 3, 9
    for k, v in THETA.items():
        setattr(prm, k, v)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    res = []
    for _ in range(2):
        adj = A.VFSDepthAdjoint(prm, pr["mask"], tvf, cvf, pr["cam"], pr["t"], pr["z_gt"], checkpoint=4)
        loss = adj.forward(ref0, pr["img_other"])
        res.append((loss, adj.backward()))
    (l0, g0), (l1, g1) = res
    assert l0 == l1
    assert np.isfinite([g0[k] for k in ("Lambda", "alpha0", "alpha1", "beta", "gamma")]).all()
    _assert_same_bits(g0, g1)


@pytest.mark.parametrize("shared", [False, True])
def test_concurrent_gradients_equal_sequential_ones(T, shared):
    """vfs_adjoint.gradients runs three pairs on two streams (two host threads): the
    losses and gradients equal forward(); backward() run one pair after another, bit for
    bit, with each pair's own workspace and with one shared workspace per stream."""
    import synthetic as S
    import vfs_adjoint as A
    import vfs_params as P
    items, seq = [], []
    for seed, (hh, ww) in zip((21, 22, 23), ((40, 48), (33, 29), (48, 40))):
        pr = S.make_pair(hh, ww, seed, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=2)
        tvf, cvf = O.get_vector_fields(O.get_grid_zero((hh, ww)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
        tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
        prm = P.vfs_params(pr["img_ref"])
        prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 2, 2.0, 3, 9
        for k, v in THETA.items():
            setattr(prm, k, v)
        ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
        make = lambda: A.VFSDepthAdjoint(prm, pr["mask"], tvf, cvf, pr["cam"], pr["t"], pr["z_gt"],  # noqa: E731
                                         checkpoint=4, workspace={} if shared else None)
        a = make()
        seq.append((a.forward(ref0, pr["img_other"]), a.backward()))
        items.append((make(), ref0, pr["img_other"]))
    ws = [{}, {}] if shared else None
    for _ in range(2):
        got = A.gradients(items, streams=2, workspaces=ws)
        for (l0, g0), (l1, g1) in zip(seq, got):
            assert l0 == l1
            _assert_same_bits(g0, g1)


def test_reverse_probe_times_the_blocked_launches_and_leaves_gradients_unchanged(T):
    """The timing probe of backward(): events around each temporally blocked reverse
    launch of the probed level, one unit count per launch (n_iter - 1 units per warp:
    the warp's first and last units run in the fused kernel), the same gradient bits
    as without the probe, and a cap on the number of launches probed."""
    pr, tvf, cvf = _small_case()
    n_warp, n_iter = 2, 9
    adj = _adjoint(pr, tvf, cvf, n_warp, n_iter, ckpt=4)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    adj.forward(ref0, pr["img_other"])
    g0 = adj.backward()
    adj.forward(ref0, pr["img_other"])
    probe = dict(level=0, launches=1000)
    g1 = adj.backward(probe=probe)
    _assert_same_bits(g0, g1)
    tx, ty, k, nt = probe["tiles"]
    print("probe", probe["tiles"], probe["units"], [round(x, 2) for x in probe["us"]])
    assert len(probe["us"]) == len(probe["units"]) > 0
    assert all(x > 0 for x in probe["us"])
    assert all(1 <= u <= k for u in probe["units"])
    assert sum(probe["units"]) == n_warp * (n_iter - 1)
    adj.forward(ref0, pr["img_other"])
    capped = dict(level=0, launches=2)
    g2 = adj.backward(probe=capped)
    _assert_same_bits(g0, g2)
    assert capped["units"] == probe["units"][:2]


@pytest.mark.parametrize("shape,ckpt,rows", [((54, 70), 4, None), ((48, 48), 9, None), ((54, 70), 4, 1),
                                             ((54, 70), 5, 2), ((54, 70), 3, 4), ((100, 68), 4, 4), ((17, 23), 4, 1)])
def test_fused_sweep_matches_one_kernel_per_step(T, shape, ckpt, rows):
    """The fused reverse sweep (dual^T + primal^T + threshold^T per launch, ring messages in
    LDS) against the one-kernel-per-step sweep, bit for bit: the per-pixel arithmetic is
    shared and each pixel adds its parameter shares in the same order; partial tiles at
    the right / bottom edges (17 x 23: every tile partial, levels of 17 x 23 and 8 x 11)."""
    import synthetic as S
    import vfs_adjoint as A
    import vfs_params as P
    hh, ww = shape
    pr = S.make_pair(hh, ww, 8, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=2)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((hh, ww)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
    tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 2, 2.0, 3, 9
    for k, v in THETA.items():
        setattr(prm, k, v)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    res = []
    for fused in (False, True):
        adj = A.VFSDepthAdjoint(prm, pr["mask"], tvf, cvf, pr["cam"], pr["t"], pr["z_gt"], checkpoint=ckpt,
                                fused=fused, rows=rows, reverse_block=None)
        loss = adj.forward(ref0, pr["img_other"])
        res.append((loss, adj.backward()))
    (l0, g0), (l1, g1) = res
    assert l0 == l1
    _assert_same_bits(g0, g1)


@pytest.mark.parametrize("shape,ckpt,rb", [((54, 70), 4, 2), ((48, 48), 9, 3), ((54, 70), 5, 4), ((61, 47), 4, 1),
                                           ((54, 70), 9, (2, 512)), ((37, 90), 4, (1, 512)), ((61, 47), 7, 5),
                                           ((54, 70), 9, 6), ((37, 90), 9, (3, 512)), ((17, 23), 4, 3),
                                           ((17, 23), 9, (2, 512)), ((54, 70), 9, (3, 512, "lean")),
                                           ((37, 90), 4, (1, 512, "lean")), ((17, 23), 9, (2, 512, "lean"))])
def test_blocked_sweep_matches_fused_sweep(T, shape, ckpt, rb):
    """The temporally blocked reverse sweep (K units per launch, adjoint state and parameter
    shares in registers, messages in LDS) against the fused sweep, bit for bit: the
    per-pixel adjoints are the same arithmetic in the same order.  Segments of 4, 5 and 9
    iterations give launches of fewer units than K at every segment boundary."""
    import synthetic as S
    import vfs_adjoint as A
    import vfs_params as P
    hh, ww = shape
    pr = S.make_pair(hh, ww, 8, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=2)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((hh, ww)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
    tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.fScale, prm.nWarpIters, prm.nSolverIters = 2, 2.0, 3, 9
    for k, v in THETA.items():
        setattr(prm, k, v)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    res = []
    for blk in (None, rb):
        adj = A.VFSDepthAdjoint(prm, pr["mask"], tvf, cvf, pr["cam"], pr["t"], pr["z_gt"], checkpoint=ckpt,
                                reverse_block=blk)
        loss = adj.forward(ref0, pr["img_other"])
        res.append((loss, adj.backward()))
    (l0, g0), (l1, g1) = res
    assert l0 == l1
    _assert_same_bits(g0, g1)


def test_adjoint_at_default_parameters_matches_autograd(T):
    """The gradient at the reference's default parameters and pyramid ratio (fScale 1.2,
    lambda 5, alpha0 17, alpha1 1.2, beta 9, gamma 0.85) on a non-square 44 x 52 pair with
    four levels down to 25 x 29, so the up- and downsampling adjoints run at a general
    ratio: against float64 autograd through the unrolled restatement.  With these
    parameters float32 arithmetic alone moves the parameter gradients by up to 6 % (the
    same restatement under float32 autograd: lambda -0.0655 against -0.0688), so each is
    held to max(2e-2 relative, 3x that float32 deviation); the image gradient to cosine
    >= 0.999."""
    import synthetic as S
    import vfs_adjoint as A
    import vfs_params as P
    hh, ww, levels, n_warp, n_iter = 44, 52, 4, 2, 6
    pr = S.make_pair(hh, ww, 6, t=(0.2, 0.0, 0.02), rot_deg=1.0, n_levels=levels, level_ratio=1.2)
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((hh, ww)), pr["cam"], pr["cam"], pr["R"], pr["t"], 1)
    tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
    prm = P.vfs_params(pr["img_ref"])
    prm.nLevel, prm.nWarpIters, prm.nSolverIters = levels, n_warp, n_iter
    assert prm.fScale == 1.2
    theta = {k: float(getattr(prm, k)) for k in ("Lambda", "alpha0", "alpha1", "beta", "gamma")}
    assert theta == dict(Lambda=5.0, alpha0=17.0, alpha1=1.2, beta=9.0, gamma=0.85)
    adj = A.VFSDepthAdjoint(prm, pr["mask"], tvf, cvf, pr["cam"], pr["t"], pr["z_gt"], checkpoint=4)
    ref0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"]).astype(np.float32), device="cuda")
    loss = adj.forward(ref0, pr["img_other"])
    g = adj.backward()
    sizes = O.compute_pyramid_sizes(hh, ww, levels, 1.2)
    assert sizes[-1] == (25, 29)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(pr["mask"], sizes)]
    trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
    r0 = torch.as_tensor(O.cast_and_scale(pr["img_ref"])).requires_grad_(True)
    ow = torch.as_tensor(O.warping_image(O.cast_and_scale(pr["img_other"]), cvf))
    th = [torch.tensor(theta[k], dtype=torch.float64, requires_grad=True) for k in
          ("Lambda", "alpha0", "alpha1", "beta", "gamma")]
    flow = R.solve_stereo(r0, ow, masks, trs, sizes, *th, n_warp, n_iter, f_scale=1.2)
    lt = R.depth_loss(flow, pr["cam"], torch.as_tensor(pr["t"].reshape(3)), masks[0], torch.as_tensor(pr["z_gt"]))
    lt.backward()
    # the same restatement under float32 autograd
    f = torch.float32
    th32 = [torch.tensor(theta[k], dtype=f, requires_grad=True) for k in ("Lambda", "alpha0", "alpha1", "beta", "gamma")]
    flow32 = R.solve_stereo(torch.as_tensor(O.cast_and_scale(pr["img_ref"])).to(f), ow.to(f), masks,
                            [t.to(f) for t in trs], sizes, *th32, n_warp, n_iter, f_scale=1.2)
    R.depth_loss(flow32, pr["cam"], torch.as_tensor(pr["t"].reshape(3)).to(f), masks[0],
                 torch.as_tensor(pr["z_gt"]).to(f)).backward()
    print("loss hip", loss, "torch", float(lt))
    assert abs(loss - float(lt)) <= 1e-3 * abs(float(lt))
    for i, k in enumerate(("Lambda", "alpha0", "alpha1", "beta", "gamma")):
        ref, r32 = th[i].grad.item(), th32[i].grad.item()
        print(k, g[k], ref, "fp32 restatement", r32)
        assert abs(g[k] - ref) <= max(2e-2 * abs(ref), 3 * abs(r32 - ref)) + 1e-6, k
    gr, rr = g["ref0"].cpu().numpy().ravel(), r0.grad.numpy().ravel()
    cos = float(np.dot(gr, rr) / (np.linalg.norm(gr) * np.linalg.norm(rr)))
    print("ref0 cosine", cos, "rel", np.linalg.norm(gr - rr) / np.linalg.norm(rr))
    assert cos >= 0.999


@pytest.mark.parametrize("theta_device", ["cpu", "cuda"])
def test_depth_loss_autograd_function(T, theta_device):
    """DepthLoss.apply as the integration guide writes it: theta.grad and ref0.grad equal
    adj.backward()'s gradients bit for bit (theta on the CPU or the GPU), and a second
    forward on the same adjoint object before backward makes the first loss's backward
    raise instead of returning the second loss's gradient."""
    import vfs_adjoint as A
    pr, tvf, cvf = _small_case(40, 13)
    adj = _adjoint(pr, tvf, cvf, 2, 5, ckpt=3)
    ref_np = O.cast_and_scale(pr["img_ref"]).astype(np.float32)
    ref0 = torch.as_tensor(ref_np, device="cuda").requires_grad_()
    theta = torch.tensor([THETA[k] for k in A.PARAMS], dtype=torch.float64, device=theta_device, requires_grad=True)
    loss = A.DepthLoss.apply(ref0, theta, adj, pr["img_other"])
    assert loss.device == theta.device
    loss.backward()
    assert theta.grad.device == theta.device and theta.grad.dtype == theta.dtype
    assert ref0.grad.device == ref0.device and ref0.grad.dtype == ref0.dtype
    # the same gradient straight from the adjoint object
    adj2 = _adjoint(pr, tvf, cvf, 2, 5, ckpt=3)
    l2 = adj2.forward(torch.as_tensor(ref_np, device="cuda"), pr["img_other"])
    g2 = adj2.backward()
    assert float(loss) == l2
    assert theta.grad.cpu().tolist() == [g2[k] for k in A.PARAMS]
    assert torch.equal(ref0.grad, g2["ref0"])
    # two losses on one adjoint object: the first one's backward refuses to run
    r1 = torch.as_tensor(ref_np, device="cuda").requires_grad_()
    la = A.DepthLoss.apply(r1, theta.detach().clone().requires_grad_(), adj, pr["img_other"])
    lb = A.DepthLoss.apply(r1, theta.detach().clone().requires_grad_(), adj, pr["img_other"])
    with pytest.raises(RuntimeError, match="another forward"):
        la.backward()
    lb.backward()
    assert torch.isfinite(r1.grad).all()


def test_adjoint_matches_central_differences_on_a_crop_of_the_800_workload(T):
    """The differentiation workload's schedule (3 levels x 10 warps x 50 iterations,
    checkpoint 25, the reference's default parameters) on the 128 x 128 centre crop of
    its seed-0 800 x 800 pair (camera principal point moved with the crop, mask with a
    3-pixel zero border).

    Central differences of the float64 oracle's loss converge as h shrinks, since the
    solve is piecewise smooth (threshold branches, projections, median): relative to
    the float64 autograd of the restatement they are off by up to 1.9 % at h = 1e-3,
    1.2 % at 1e-4 and 0.14 % at 1e-5 (measured).  At h = 1e-5:
      * the differences agree with the float64 autograd: gradient vector within 1e-3
        relative (measured 4e-4) -- the checker is exact;
      * the hand adjoint (fp32) agrees with the differences: gradient vector within
        1e-3 relative (measured 7e-4; per parameter lambda 2e-4, alpha1 7e-4, beta
        9e-5, gamma 6e-4, and alpha0, whose derivative is 6e-5 of the vector's norm,
        4e-3)."""
    import synthetic as S
    import vfs_adjoint as A
    import vfs_params as P
    pr = S.workload_pair("adjoint_800", seed=0)
    c0, n = 400 - 64, 128
    sl = (slice(c0, c0 + n), slice(c0, c0 + n))
    f, cx, cy = pr["cam"]
    cam = (f, cx - c0, cy - c0)
    ref, oth, zg = pr["img_ref"][sl].copy(), pr["img_other"][sl].copy(), pr["z_gt"][sl].copy()
    mask = np.zeros((n, n), np.uint8)
    mask[3:-3, 3:-3] = 255
    tvf, cvf = O.get_vector_fields(O.get_grid_zero((n, n)), cam, cam, pr["R"], pr["t"], 1)
    tvf, cvf = S.sanitize_field(tvf), cvf.astype(np.float32)
    w = S.WORKLOADS["adjoint_800"]
    nl, nw, ni = w["nLevel"], w["nWarpIters"], w["nSolverIters"]
    theta = {k: float(w[k]) for k in A.PARAMS}
    prm = P.vfs_params(ref)
    S.apply_workload_params(prm, w)
    adj = A.VFSDepthAdjoint(prm, mask, tvf, cvf, cam, pr["t"], zg, checkpoint=w["checkpoint"])
    adj.forward(torch.as_tensor(O.cast_and_scale(ref).astype(np.float32), device="cuda"), oth)
    g = adj.backward()
    ga = np.array([g[k] for k in A.PARAMS])
    # float64 autograd of the restatement
    sizes = O.compute_pyramid_sizes(n, n, nl, 2.0)
    masks = [torch.as_tensor(m > 0) for m in O.create_mask_pyramid(mask, sizes)]
    trs = [torch.as_tensor(t) for t in O.create_vector_field_pyramid(tvf, sizes)]
    ow = torch.as_tensor(O.warping_image(O.cast_and_scale(oth), cvf))
    th = [torch.tensor(theta[k], dtype=torch.float64, requires_grad=True) for k in A.PARAMS]
    flow = R.solve_stereo(torch.as_tensor(O.cast_and_scale(ref)), ow, masks, trs, sizes, *th, nw, ni)
    R.depth_loss(flow, cam, torch.as_tensor(pr["t"].reshape(3)), masks[0], torch.as_tensor(zg)).backward()
    g64 = np.array([x.grad.item() for x in th])

    def oloss(tt):
        p = O.OracleParams(n, n, nLevel=nl, fScale=2.0, nWarpIters=nw, nSolverIters=ni, **tt)
        fl = O.solve_stereo(ref, oth, mask, tvf, cvf, p)
        z = O.triangulate_flow(fl, cam, pr["t"].reshape(3), mask)[..., 2]
        m = (mask > 0) & (z > 0) & np.isfinite(z) & (zg > 0)
        return float(np.mean(np.abs(np.log(z[m]) - np.log(zg[m]))))

    h = 1e-5
    fd = []
    for k in A.PARAMS:
        tp, tm = dict(theta), dict(theta)
        tp[k] *= 1 + h
        tm[k] *= 1 - h
        fd.append((oloss(tp) - oloss(tm)) / (2 * h * theta[k]))
    fd = np.array(fd)
    for k, a_, c_, d_ in zip(A.PARAMS, ga, g64, fd):
        print(f"{k}: adjoint {a_:.8g} float64 autograd {c_:.8g} central difference {d_:.8g}")
    r_fd = np.linalg.norm(fd - g64) / np.linalg.norm(g64)
    r_adj = np.linalg.norm(ga - fd) / np.linalg.norm(fd)
    print(f"|fd - autograd| / |autograd| = {r_fd:.2e}; |adjoint - fd| / |fd| = {r_adj:.2e}")
    assert r_fd <= 1e-3
    assert r_adj <= 1e-3
