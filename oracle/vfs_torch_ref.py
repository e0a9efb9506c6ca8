"""Plain unrolled PyTorch restatement of the VFS forward solve, differentiable by autograd.

TEST INFRASTRUCTURE ONLY (like oracle/vfs_oracle.py): it checks the hand-written
adjoint of the HIP library.  Every function follows the same reference kernel
as its NumPy counterpart in vfs_oracle.py (file:line cited there) and computes
in the dtype of its inputs (float64 in the tests).  Solver parameters (lambda,
alpha0, alpha1, beta, gamma) may be 0-d tensors that require grad.

Non-smooth points take these one-sided derivatives (the same as the
hand-written adjoint): max(|p|, 1) and max(exp(.), 1e-8) pass the gradient to
the larger argument (to the first on a tie); min(d, limit) passes it to d when
d < limit; the threshold uses the branch the forward took; the median routes
the gradient to the selected element; bilinear weights are differentiated on
the cell the forward used; norms have zero derivative at 0.
"""

import torch

from oracle import vfs_oracle as O

EPS = O.EPS


def _mirror(i, n):
    m = torch.remainder(i, 2 * n)
    return torch.where(m >= n, 2 * n - 1 - m, m)


def sample_linear(img, x, y):
    H, W = img.shape[:2]
    xc = torch.clamp(torch.nan_to_num(x.detach(), nan=0.0), -1e7, 1e7)
    yc = torch.clamp(torch.nan_to_num(y.detach(), nan=0.0), -1e7, 1e7)
    x0 = torch.floor(xc)
    y0 = torch.floor(yc)
    ax = x - x0
    ay = y - y0
    ix0, ix1 = _mirror(x0.long(), W), _mirror(x0.long() + 1, W)
    iy0, iy1 = _mirror(y0.long(), H), _mirror(y0.long() + 1, H)
    flat = img.reshape(H * W, -1)
    t00, t10 = flat[iy0 * W + ix0], flat[iy0 * W + ix1]
    t01, t11 = flat[iy1 * W + ix0], flat[iy1 * W + ix1]
    ax, ay = ax[..., None], ay[..., None]
    r0 = t00 + ax * (t10 - t00)
    r1 = t01 + ax * (t11 - t01)
    out = r0 + ay * (r1 - r0)
    return out[..., 0] if img.dim() == 2 else out


def grid(H, W, ref):
    Y, X = torch.meshgrid(torch.arange(H, dtype=ref.dtype, device=ref.device),
                          torch.arange(W, dtype=ref.dtype, device=ref.device), indexing="ij")
    return X, Y


def _nbr(a, dx, dy):
    out = a
    if dx == 1:
        out = torch.cat([out[:, 1:], out[:, -1:]], 1)
    elif dx == -1:
        out = torch.cat([out[:, :1], out[:, :-1]], 1)
    if dy == 1:
        out = torch.cat([out[1:], out[-1:]], 0)
    elif dy == -1:
        out = torch.cat([out[:1], out[:-1]], 0)
    return out


def safe_norm(*comps):
    """sqrt(sum c^2) with zero derivative at 0."""
    z = comps[0] == 0
    for c in comps[1:]:
        z = z & (c == 0)
    s = sum(torch.where(z, torch.ones_like(c), c) ** 2 if i == 0 else torch.where(z, torch.zeros_like(c), c) ** 2
            for i, c in enumerate(comps))
    return torch.where(z, torch.zeros_like(s), torch.sqrt(s))


def downsample_bilinear(src, out_shape):
    Hs, Ws = src.shape[:2]
    Hd, Wd = out_shape
    X, Y = grid(Hd, Wd, src)
    rx, ry = Ws / Wd, Hs / Hd
    xc, yc = (X + 0.5) * rx - 0.5, (Y + 0.5) * ry - 0.5
    xm, xp = (X + 0.25) * rx - 0.5, (X + 0.75) * rx - 0.5
    ym, yp = (Y + 0.25) * ry - 0.5, (Y + 0.75) * ry - 0.5
    return 0.25 * (sample_linear(src, xm, yc) + sample_linear(src, xp, yc)
                   + sample_linear(src, xc, ym) + sample_linear(src, xc, yp))


def warping_image(src, warp):
    H, W = warp.shape[:2]
    X, Y = grid(H, W, warp)
    return sample_linear(src, X + warp[..., 0], Y + warp[..., 1])


def gaussian_masked(img, m):
    v1 = _nbr(img, -1, -1) + _nbr(img, -1, 1) + _nbr(img, 1, -1) + _nbr(img, 1, 1)
    v2 = _nbr(img, -1, 0) + _nbr(img, 1, 0) + _nbr(img, 0, -1) + _nbr(img, 0, 1)
    out = 0.25 * img + 0.0625 * v1 + 0.125 * v2
    return torch.where(m, out, torch.zeros_like(out))


def compute_tensor_eta(img, m, beta, gamma, alpha0, alpha1):
    i11 = _nbr(img, 1, 1)
    gx = i11 - _nbr(img, 0, 1)
    gy = i11 - _nbr(img, 1, 0)
    ng = safe_norm(gx, gy)
    small = ng < O.MIN_NORMALIZED_GRAD
    den = torch.where(small, torch.ones_like(ng), ng)
    nx = torch.where(small, torch.ones_like(gx), gx / den)
    ny = torch.where(small, torch.zeros_like(gy), gy / den)
    tx, ty = ny, -nx
    pw = torch.where(ng > 0, torch.where(ng > 0, ng, torch.ones_like(ng)) ** gamma, torch.zeros_like(ng))
    ex = torch.exp(-beta * pw)
    w = torch.where(ex >= O.MIN_TENSOR_VALUE, ex, torch.full_like(ex, O.MIN_TENSOR_VALUE))
    a = w * nx * nx + tx * tx
    b = w * ny * ny + ty * ty
    c = w * nx * ny + tx * ty
    a11, a004 = alpha1 * alpha1, 4.0 * alpha0 * alpha0
    e0 = a11 * (a * a + b * b + 2.0 * c * c) + a11 * ((a + c) ** 2 + (b + c) ** 2)
    e1 = a11 * (b * b + c * c) + a004
    e2 = a11 * (a * a + c * c) + a004
    z = torch.zeros_like(a)
    abc = torch.stack([torch.where(m, a, z), torch.where(m, b, z), torch.where(m, c, z)], -1)
    eta = torch.stack([torch.where(m, e0, z), torch.where(m, e1, z), torch.where(m, e2, z)], -1)
    return abc, eta


def compute_derivatives(ow, ref, tvf):
    H, W = ref.shape
    X, Y = grid(H, W, ref)
    nrm = safe_norm(tvf[..., 0], tvf[..., 1])
    ex, ey = tvf[..., 0] / nrm, tvf[..., 1] / nrm

    def taps(img):
        v = sample_linear(img, X - 2 * ex, Y - 2 * ey)
        v = v - 8.0 * sample_linear(img, X - ex, Y - ey)
        v = v + 8.0 * sample_linear(img, X + ex, Y + ey)
        return v - sample_linear(img, X + 2 * ex, Y + 2 * ey)

    return torch.stack([(taps(ref) + taps(ow)) / 24.0, ow - ref], -1)


def update_dual(p, q, sigma, u_, v_, abc, m, alpha0, alpha1):
    mE, mS = _nbr(m, 1, 0), _nbr(m, 0, 1)
    uE, uW, uS, uN = _nbr(u_, 1, 0), _nbr(u_, -1, 0), _nbr(u_, 0, 1), _nbr(u_, 0, -1)
    vE, vW, vS, vN = _nbr(v_, 1, 0), _nbr(v_, -1, 0), _nbr(v_, 0, 1), _nbr(v_, 0, -1)
    vx, vy = v_[..., 0], v_[..., 1]
    gux = torch.where(mE, uE - u_ - vx, u_ - uW - vx)
    gvx = torch.where(mE, vE[..., 0] - vx, vx - vW[..., 0])
    gvw = torch.where(mE, vE[..., 1] - vy, vy - vW[..., 1])
    guy = torch.where(mS, uS - u_ - vy, u_ - uN - vy)
    gvy = torch.where(mS, vS[..., 1] - vy, vy - vN[..., 1])
    gvz = torch.where(mS, vS[..., 0] - vx, vx - vN[..., 0])
    a, b, c = abc[..., 0], abc[..., 1], abc[..., 2]
    s = alpha1 * sigma / O.ETA_P
    px = p[..., 0] + s * (a * gux + c * guy)
    py = p[..., 1] + s * (c * gux + b * guy)
    n = safe_norm(px, py)
    n = torch.where(n > 1.0, n, torch.ones_like(n))
    px, py = px / n, py / n
    s = alpha0 * sigma / O.ETA_Q
    qn = [q[..., 0] + s * gvx, q[..., 1] + s * gvy, q[..., 2] + s * gvz, q[..., 3] + s * gvw]
    n = safe_norm(*qn)
    n = torch.where(n > 1.0, n, torch.ones_like(n))
    z = torch.zeros_like(px)
    P = torch.stack([torch.where(m, px, z), torch.where(m, py, z)], -1)
    TP = torch.stack([torch.where(m, a * px + c * py, z), torch.where(m, c * px + b * py, z)], -1)
    Q = torch.stack([torch.where(m, x / n, z) for x in qn], -1)
    return P, TP, Q


def l1_thresholding(tau, tp, u, u_, deriv, eta, m, lam):
    mE, mW, mS, mN = _nbr(m, 1, 0), _nbr(m, -1, 0), _nbr(m, 0, 1), _nbr(m, 0, -1)
    tpx, tpy = tp[..., 0], tp[..., 1]
    tpWx, tpNy = _nbr(tp, -1, 0)[..., 0], _nbr(tp, 0, -1)[..., 1]
    gx = torch.where(mE & mW, tpx - tpWx, torch.where(~mE, -tpWx, tpx))
    gy = torch.where(mS & mN, tpy - tpNy, torch.where(~mS, -tpNy, tpy))
    iu, iz, e0 = deriv[..., 0], deriv[..., 1], eta[..., 0]
    te = torch.where(e0.abs() < EPS, tau * torch.ones_like(e0), tau / torch.where(e0 == 0, torch.ones_like(e0), e0))
    ut = u_ + te * (gx + gy)
    rho = iu * (ut - u) + iz
    bound = lam * te * iu * iu
    d0 = ut - u
    small = iu.abs() < EPS
    du_in = torch.where(small, d0, d0 - rho / torch.where(small, torch.ones_like(iu), iu))
    inside = (rho <= bound) & (rho >= -bound)
    du = torch.where(inside, du_in, torch.where(rho < -bound, d0 + lam * te * iu,
                                                torch.where(rho > bound, d0 - lam * te * iu, torch.zeros_like(d0))))
    return torch.where(m, u + du, torch.zeros_like(u))


def update_primal(tau, mu, p, q, u, u_, v_, abc, eta, m, alpha0, alpha1):
    mE, mW, mS, mN = _nbr(m, 1, 0), _nbr(m, -1, 0), _nbr(m, 0, 1), _nbr(m, 0, -1)
    qW, qN = _nbr(q, -1, 0), _nbr(q, 0, -1)
    z = torch.zeros_like(u)
    gqx = torch.where(mE, q[..., 0], z) - torch.where(mW, qW[..., 0], z)
    gqw = torch.where(mE, q[..., 3], z) - torch.where(mW, qW[..., 3], z)
    gqy = torch.where(mS, q[..., 1], z) - torch.where(mN, qN[..., 1], z)
    gqz = torch.where(mS, q[..., 2], z) - torch.where(mN, qN[..., 2], z)
    a, b, c = abc[..., 0], abc[..., 1], abc[..., 2]
    tux, tuy = a * p[..., 0] + c * p[..., 1], c * p[..., 0] + b * p[..., 1]
    e1, e2 = eta[..., 1], eta[..., 2]
    one = torch.ones_like(e1)
    t1 = torch.where(e1.abs() < EPS, tau * one, tau / torch.where(e1 == 0, one, e1))
    t2 = torch.where(e2.abs() < EPS, tau * one, tau / torch.where(e2 == 0, one, e2))
    vx = t1 * (alpha1 * tux + alpha0 * (gqx + gqz)) + v_[..., 0]
    vy = t2 * (alpha1 * tuy + alpha0 * (gqw + gqy)) + v_[..., 1]
    un = mu * (u - u_) + u
    V = torch.stack([torch.where(m, mu * (vx - v_[..., 0]) + vx, z), torch.where(m, mu * (vy - v_[..., 1]) + vy, z)],
                    -1)
    return torch.where(m, un, z), V


def median_filter(u, m, k=5):
    H, W = u.shape
    h = k // 2
    pad = torch.cat([u[:, :1].expand(H, h), u, u[:, -1:].expand(H, h)], 1)
    pad = torch.cat([pad[:1].expand(h, W + 2 * h), pad, pad[-1:].expand(h, W + 2 * h)], 0)
    win = pad.unfold(0, k, 1).unfold(1, k, 1).reshape(H, W, k * k)
    med = torch.kthvalue(win, (k * k) // 2 + 1, dim=-1).values
    return torch.where(m, med, u)


def flow_clamp(flow, u0, uf, tv, m, limit):
    d = uf - u0
    delta = torch.where(d < limit, d, torch.full_like(d, limit))
    delta = torch.where(m, delta, torch.zeros_like(delta))
    return flow + delta[..., None] * tv, u0 + delta


def upsampling(src, out_shape, f_scale):
    Hs, Ws = src.shape[:2]
    Hd, Wd = out_shape
    X, Y = grid(Hd, Wd, src)
    return f_scale * sample_linear(src, (X + 0.5) * (Ws / Wd) - 0.5, (Y + 0.5) * (Hs / Hd) - 0.5)


def unproject_xy(uv, f, u0, v0):
    """x / z of the unit ray through uv (the quantity triangulation uses)."""
    x = uv[..., 0] - u0
    y = uv[..., 1] - v0
    r = safe_norm(x, y)
    th = r / f
    tr = torch.where(r > 0, torch.tan(th) / torch.where(r > 0, r, torch.ones_like(r)), torch.full_like(r, 1.0 / f))
    return x * tr, y * tr


def triangulate_z(flow, cam, t, m):
    H, W = flow.shape[:2]
    X, Y = grid(H, W, flow)
    g = torch.stack([X, Y], -1)
    xr, _ = unproject_xy(g, *cam)
    xo, _ = unproject_xy(g + flow, *cam)
    z = (t[0] - t[2] * xo) / (xo - xr)
    return torch.where(m, z, torch.full_like(z, -1.0))


def solve_stereo(ref0, other0_warped, masks, trs, sizes, lam, alpha0, alpha1, beta, gamma, n_warp, n_iter,
                 limit=0.1, f_scale=2.0, time_step_lambda=1.0):
    """Python/vfs.py:460-533 from the float level-0 images (ref0 = u8/256; the other view already
    de-rotated).  masks: bool level masks; trs: tvf pyramid.  Returns the masked flow."""
    refs = [ref0]
    others = [other0_warped]
    for sz in sizes[1:]:
        refs.append(downsample_bilinear(refs[-1], sz))
        others.append(downsample_bilinear(others[-1], sz))
    taus, sigmas, mus = O.pd_schedule(n_iter, time_step_lambda)
    dt = ref0.dtype
    dev = ref0.device
    warp = u = u_ = None
    for k in range(len(sizes), 0, -1):
        s = sizes[k - 1]
        m = masks[k - 1]
        if k == len(sizes):
            warp = torch.zeros((*s, 2), dtype=dt, device=dev)
            u = torch.zeros(s, dtype=dt, device=dev)
            u_ = torch.zeros(s, dtype=dt, device=dev)
        ref_s = gaussian_masked(refs[k - 1], m)
        abc, eta = compute_tensor_eta(ref_s, m, beta, gamma, alpha0, alpha1)
        for _ in range(n_warp):
            tv = warping_image(trs[k - 1], warp)
            ow = warping_image(others[k - 1], warp)
            deriv = compute_derivatives(ow, refs[k - 1], trs[k - 1])
            p = torch.zeros((*s, 2), dtype=dt, device=dev)
            q = torch.zeros((*s, 4), dtype=dt, device=dev)
            v_ = torch.zeros((*s, 2), dtype=dt, device=dev)
            u0 = u
            for ks in range(n_iter):
                p, tp, q = update_dual(p, q, float(sigmas[ks]), u_, v_, abc, m, alpha0, alpha1)
                u = l1_thresholding(float(taus[ks]), tp, u, u_, deriv, eta, m, lam)
                u_, v_ = update_primal(float(taus[ks]), float(mus[ks]), p, q, u, u_, v_, abc, eta, m, alpha0, alpha1)
            uf = median_filter(u, m)
            warp, u = flow_clamp(warp, u0, uf, tv, m, limit)
        if k > 1:
            u_ = upsampling(u, sizes[k - 2], f_scale)
            u = u_
            warp = upsampling(warp, sizes[k - 2], f_scale)
    return torch.where(masks[0][..., None], warp, torch.zeros_like(warp))


def depth_loss(flow, cam, t, m, z_gt):
    """mean |log z - log z_gt| over the valid mask pixels: z finite and > 0, z_gt > 0
    (the pixels k_depth_loss counts)."""
    z = triangulate_z(flow, cam, t, m)
    ok = m & (z > 0) & torch.isfinite(z) & (z_gt > 0)
    zs = torch.where(ok, z, torch.ones_like(z))
    zg = torch.where(ok, z_gt, torch.ones_like(z_gt))
    return torch.sum(torch.where(ok, torch.abs(torch.log(zs) - torch.log(zg)), torch.zeros_like(z))) / ok.sum()
