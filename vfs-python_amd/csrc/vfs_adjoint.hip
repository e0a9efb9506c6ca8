// Hand-written adjoint (reverse mode) of the VFS solve, gfx950.
//
// Each forward kernel of the solve has a reverse kernel here.  The primal-dual
// iteration is reversed step by step (primal, threshold, dual), re-reading the
// forward state of the iteration from a tape that the host recomputes from
// checkpoints.  Contributions to a neighbour's adjoint are written as per-pixel
// messages and collected by the receiving pixel ("pull"), so these sums are
// deterministic; the bilinear transposes (warp, derivative taps, pyramid,
// upsampling) scatter with float atomics.  Gradients of the scalar parameters
// (lambda, alpha0, alpha1, beta, gamma) accumulate per pixel in a (H, W, 5)
// plane and are reduced once at the end (vfs_adj_reduce).
//
// Non-smooth points follow one convention throughout: projections are active
// when the norm exceeds 1; max(exp, 1e-8) differentiates exp when exp >= 1e-8;
// min(d, limit) passes the gradient to d when d < limit; the threshold uses
// the branch its forward took; the median routes to the selected element.
#include "vfs_adj_pixel.h"
#include "vfs_median.h"

using namespace vfs;

namespace {

#define VFS_ADJ_PIXEL(H, W)                              \
    const int X = blockIdx.x * blockDim.x + threadIdx.x; \
    const int Y = blockIdx.y * blockDim.y + threadIdx.y; \
    if (X >= (W) || Y >= (H)) return;                    \
    const int idx = Y * (W) + X;

// dual^T messages stored per pixel: um (H,W,3), vm (H,W,6)
struct GlobalMsgs {
    const float* um;
    const float* vm;
    int W;
    __device__ __forceinline__ DualMsg operator()(int x, int y) const
    {
        const int i = y * W + x;
        DualMsg m;
        for (int k = 0; k < 3; ++k) m.um[k] = um[3 * i + k];
        for (int k = 0; k < 6; ++k) m.vm[k] = vm[6 * i + k];
        return m;
    }
};

// ---------------------------------------------------------------------------
// primal-dual iteration, reverse (one kernel per step)
// ---------------------------------------------------------------------------

// Reverse of updatePrimalVariables at iteration t.
// In:  gubB / gvbB hold the own-pixel part of the adjoint of u_bar_t / v_bar_t; the
//      dual messages um / vm of the following iteration complete it (has_msgs = 0
//      when there is none).
// Out: gubB / gvbB := own-pixel part of the adjoint of u_bar_{t-1} / v_bar_{t-1};
//      gp += , gq += (own), qmsg = (to W: x, w; to N: y, z), gu += (adjoint of u_t);
//      per-pixel accumulators of abc, eta and the parameters.
__global__ void __launch_bounds__(256) k_adj_primal(float* __restrict__ gubB, float2* __restrict__ gvbB,
                                                    const float* __restrict__ um, const float* __restrict__ vm,
                                                    int has_msgs, float2* __restrict__ gp, float4* __restrict__ gq,
                                                    float4* __restrict__ qmsg, float* __restrict__ gu,
                                                    float* __restrict__ gabc, float* __restrict__ geta,
                                                    float* __restrict__ gpar, const float2* __restrict__ p,
                                                    const float4* __restrict__ q, const float* __restrict__ abc,
                                                    const float* __restrict__ eta, const uint8_t* __restrict__ mask,
                                                    int H, int W, float tau, float mu, float alpha0, float alpha1)
{
    VFS_ADJ_PIXEL(H, W)
    float g_ubo = gubB[idx];
    float2 g_vbo = gvbB[idx];
    if (has_msgs) gather_msgs(mask, GlobalMsgs{um, vm, W}, X, Y, H, W, g_ubo, g_vbo);
    const Nbr n = nbr(mask, X, Y, H, W);
    if (!n.m) {
        // outputs are constant zero: nothing flows back
        gubB[idx] = 0.0f;
        gvbB[idx] = make_float2(0.0f, 0.0f);
        qmsg[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float ubB, uu = gu[idx];
    float2 vbB, pp = gp[idx];
    float4 qq = gq[idx], qm;
    float* gpp = gpar + GP_N * idx;
    float ga[3] = {gabc[3 * idx], gabc[3 * idx + 1], gabc[3 * idx + 2]};
    float e1 = geta[3 * idx + 1], e2 = geta[3 * idx + 2];
    float pa0 = gpp[GP_ALPHA0], pa1 = gpp[GP_ALPHA1];
    float2 unused;
    primal_adj_pixel(n, q[idx], n.mW ? q[n.iW] : zero4, n.mN ? q[n.iN] : zero4, abc[3 * idx], abc[3 * idx + 1],
                     abc[3 * idx + 2], p[idx], eta[3 * idx + 1], eta[3 * idx + 2], tau, mu, alpha0, alpha1, g_ubo,
                     g_vbo, ubB, vbB, uu, pp, qq, qm, ga, e1, e2, pa0, pa1, unused);
    gubB[idx] = ubB;
    gvbB[idx] = vbB;
    gu[idx] = uu;
    gpp[GP_ALPHA1] = pa1;
    gpp[GP_ALPHA0] = pa0;
    geta[3 * idx + 1] = e1;
    geta[3 * idx + 2] = e2;
    gp[idx] = pp;
    for (int k = 0; k < 3; ++k) gabc[3 * idx + k] = ga[k];
    gq[idx] = qq;
    qmsg[idx] = qm;
}

// Reverse of l1Thresholding at iteration t.
// In:  gu = adjoint of u_t (complete).  State: p_t (tp_t = T p_t is recomputed),
//      u_{t-1}, u_bar_{t-1}.
// Out: gu := adjoint of u_{t-1}; tmsg = (own tp.x, own tp.y, to W (x), to N (y));
//      gubB += ; accumulators of deriv, eta.x and lambda.
__global__ void __launch_bounds__(256) k_adj_l1(float* __restrict__ gu, float4* __restrict__ tmsg,
                                                float* __restrict__ gubB, float* __restrict__ gderiv,
                                                float* __restrict__ geta, float* __restrict__ gpar,
                                                const float2* __restrict__ p, const float* __restrict__ abc,
                                                const float* __restrict__ u_prev, const float* __restrict__ u_bar_prev,
                                                const float2* __restrict__ deriv, const float* __restrict__ eta,
                                                const uint8_t* __restrict__ mask, int H, int W, float tau,
                                                float lambda)
{
    VFS_ADJ_PIXEL(H, W)
    const Nbr n = nbr(mask, X, Y, H, W);
    if (!n.m) {
        gu[idx] = 0.0f;
        tmsg[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    float4 tm;
    float ubB = gubB[idx], e0 = geta[3 * idx], lam = gpar[GP_N * idx + GP_LAMBDA];
    float2 gd = make_float2(gderiv[2 * idx], gderiv[2 * idx + 1]);
    float unused;
    const float g = l1_adj_pixel(n, gu[idx], tp_at(p, abc, mask, idx), tp_at(p, abc, mask, n.iW).x,
                                 tp_at(p, abc, mask, n.iN).y, deriv[idx], eta[3 * idx], u_prev[idx], u_bar_prev[idx],
                                 tau, lambda, tm, ubB, e0, gd, lam, unused);
    gu[idx] = g;
    gubB[idx] = ubB;
    geta[3 * idx] = e0;
    gderiv[2 * idx] = gd.x;
    gderiv[2 * idx + 1] = gd.y;
    gpar[GP_N * idx + GP_LAMBDA] = lam;
    tmsg[idx] = tm;
}

// Reverse of updateDualVariables at iteration t.
// In:  gp, gq: adjoints of p_t, q_t (own parts; the primal messages qmsg of this
//      iteration complete gq); tmsg: threshold messages (complete gtp_t).
//      State: p_{t-1}, q_{t-1}, u_bar_{t-1}, v_bar_{t-1}.
// Out: gp, gq := adjoints of p_{t-1}, q_{t-1}; um, vm: messages for u_bar_{t-1},
//      v_bar_{t-1}; accumulators of abc, alpha0, alpha1.
__global__ void __launch_bounds__(256) k_adj_dual(float2* __restrict__ gp, float4* __restrict__ gq,
                                                  const float4* __restrict__ qmsg, const float4* __restrict__ tmsg,
                                                  float* __restrict__ um, float* __restrict__ vm,
                                                  float* __restrict__ gabc, float* __restrict__ gpar,
                                                  const float2* __restrict__ p_prev, const float4* __restrict__ q_prev,
                                                  const float* __restrict__ u_bar_prev,
                                                  const float2* __restrict__ v_bar_prev,
                                                  const float* __restrict__ abc, const uint8_t* __restrict__ mask,
                                                  int H, int W, DualSteps st)
{
    VFS_ADJ_PIXEL(H, W)
    const Nbr n = nbr(mask, X, Y, H, W);
    float4 gqn = gq[idx];
    float2 gtp;
    collect_qt(qmsg, tmsg, n, X, Y, H, W, idx, gqn, gtp);
    if (!n.m) {
        gp[idx] = make_float2(0.0f, 0.0f);
        gq[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 3; ++k) um[3 * idx + k] = 0.0f;
        for (int k = 0; k < 6; ++k) vm[6 * idx + k] = 0.0f;
        return;
    }
    const bool fx = n.mE, fy = n.mS;
    const int ix = fx ? n.iE : n.iW, iy = fy ? n.iS : n.iN;
    float2 gpo;
    float4 gqo;
    DualMsg msg;
    float ga[3] = {gabc[3 * idx], gabc[3 * idx + 1], gabc[3 * idx + 2]};
    float* gpp = gpar + GP_N * idx;
    float pa0 = gpp[GP_ALPHA0], pa1 = gpp[GP_ALPHA1];
    dual_adj_pixel(fx, fy, u_bar_prev[idx], v_bar_prev[idx], u_bar_prev[ix], v_bar_prev[ix], u_bar_prev[iy],
                   v_bar_prev[iy], abc[3 * idx], abc[3 * idx + 1], abc[3 * idx + 2], p_prev[idx], q_prev[idx], st,
                   gp[idx], gqn, gtp, gpo, gqo, msg, ga, pa0, pa1);
    gp[idx] = gpo;
    gq[idx] = gqo;
    gpp[GP_ALPHA1] = pa1;
    gpp[GP_ALPHA0] = pa0;
    for (int k = 0; k < 3; ++k) gabc[3 * idx + k] = ga[k];
    for (int k = 0; k < 3; ++k) um[3 * idx + k] = msg.um[k];
    for (int k = 0; k < 6; ++k) vm[6 * idx + k] = msg.vm[k];
}

// Adjoints of u_bar and v_bar at the start of the primal-dual loop: own parts +
// the dual messages of the first iteration.
__global__ void __launch_bounds__(256) k_adj_pd_finish(float* __restrict__ gub, float2* __restrict__ gvb,
                                                       const float* __restrict__ gubB, const float2* __restrict__ gvbB,
                                                       const float* __restrict__ um, const float* __restrict__ vm,
                                                       const uint8_t* __restrict__ mask, int H, int W)
{
    VFS_ADJ_PIXEL(H, W)
    float g = gubB[idx];
    float2 gv = gvbB[idx];
    gather_msgs(mask, GlobalMsgs{um, vm, W}, X, Y, H, W, g, gv);
    gub[idx] = g;
    gvb[idx] = gv;
}

// ---------------------------------------------------------------------------
// per-warp reverse: clamp, median, derivatives, warp
// ---------------------------------------------------------------------------

// Reverse of computeOpticalFlowWithClamp (Python/vfs.cu:641-666) after medianFilter
// (:583-631).  The median of the thresholded u and the window position it came
// from are recomputed (ties: the first position in row-major order holding the
// median value).  In: gf (adjoint of flow_{w+1}, unchanged: d flow_{w+1}/d flow_w = I),
// gu (adjoint of u_{w+1}; becomes the clamp's share of the adjoint of u_w).
// Out: guf (adjoint of the median output), pos (window position, -1 where none),
// gtv (adjoint of the warping vector).
template <int KM>
__global__ void __launch_bounds__(256) k_adj_clamp(float* __restrict__ gu, float* __restrict__ guf,
                                                   int* __restrict__ pos, float2* __restrict__ gtv,
                                                   const float2* __restrict__ gf, const float* __restrict__ u_th,
                                                   const float* __restrict__ u0, const float2* __restrict__ tv,
                                                   const uint8_t* __restrict__ mask, int H, int W, float limit)
{
    VFS_ADJ_PIXEL(H, W)
    if (mask[idx] == 0) {
        guf[idx] = 0.0f;
        pos[idx] = -1;
        gtv[idx] = make_float2(0.0f, 0.0f);
        return;
    }
    constexpr int N = KM * KM;
    float v[N], w[N];
#pragma unroll
    for (int j = 0; j < KM; ++j) {
        const int yy = clamp_idx(Y + j - KM / 2, H);
#pragma unroll
        for (int i = 0; i < KM; ++i) w[j * KM + i] = v[j * KM + i] = u_th[yy * W + clamp_idx(X + i - KM / 2, W)];
    }
    const float med = median_select<KM>(v);
    int k = 0;
#pragma unroll
    for (int i = N - 1; i >= 0; --i)
        if (w[i] == med) k = i;
    const float base = u0[idx];
    const float d = med - base;
    const bool pass = d < limit;
    const float delta = pass ? d : limit;
    const float2 g = gf[idx];
    const float2 t = tv[idx];
    const float gdelta = g.x * t.x + g.y * t.y + gu[idx];
    gtv[idx] = make_float2(g.x * delta, g.y * delta);
    guf[idx] = pass ? gdelta : 0.0f;
    gu[idx] = gu[idx] - (pass ? gdelta : 0.0f);
    pos[idx] = k;
}

// Reverse of medianFilter: the thresholded u at pixel (X, Y) receives the adjoint
// of every median whose selected window position maps onto it.
template <int KM>
__global__ void __launch_bounds__(256) k_adj_median(float* __restrict__ gu_th, const float* __restrict__ guf,
                                                    const int* __restrict__ pos, int H, int W)
{
    VFS_ADJ_PIXEL(H, W)
    float acc = 0.0f;
    for (int dy = -(KM / 2); dy <= KM / 2; ++dy) {
        const int ys = Y + dy;
        if (ys < 0 || ys >= H) continue;
        for (int dx = -(KM / 2); dx <= KM / 2; ++dx) {
            const int xs = X + dx;
            if (xs < 0 || xs >= W) continue;
            const int i = ys * W + xs;
            const int k = pos[i];
            if (k < 0) continue;
            const int tx = clamp_idx(xs + k % KM - KM / 2, W), ty = clamp_idx(ys + k / KM - KM / 2, H);
            if (tx == X && ty == Y) acc += guf[i];
        }
    }
    gu_th[idx] = acc;
}

__device__ __forceinline__ void scatter_bilinear(float* __restrict__ g, int H, int W, float x, float y, float v)
{
    if (v == 0.0f || !(x == x) || !(y == y)) return;
    const Tap t = make_tap(x, y, H, W);
    atomicAdd(&g[t.i00], v * (1.0f - t.ax) * (1.0f - t.ay));
    atomicAdd(&g[t.i10], v * t.ax * (1.0f - t.ay));
    atomicAdd(&g[t.i01], v * (1.0f - t.ax) * t.ay);
    atomicAdd(&g[t.i11], v * t.ax * t.ay);
}

// Reverse of computeDerivatives (Python/vfs.cu:120-142): scatters the adjoint of
// (Iu, Iz) onto the reference level image and the warped image (atomics).
__global__ void __launch_bounds__(256) k_adj_derivatives(float* __restrict__ gref, float* __restrict__ gow,
                                                         const float* __restrict__ gderiv,
                                                         const float2* __restrict__ tvf, int H, int W)
{
    VFS_ADJ_PIXEL(H, W)
    const float giu = gderiv[2 * idx] / 24.0f, giz = gderiv[2 * idx + 1];
    if (giz != 0.0f) {
        atomicAdd(&gow[idx], giz);
        atomicAdd(&gref[idx], -giz);
    }
    if (giu == 0.0f) return;
    const float2 vec = tvf[idx];
    const float nrm = hypotf(vec.x, vec.y);
    const float ex = vec.x / nrm, ey = vec.y / nrm;
    const float x = static_cast<float>(X), y = static_cast<float>(Y);
    const float xs[4] = {fmaf(-2.0f, ex, x), x - ex, x + ex, fmaf(2.0f, ex, x)};
    const float ys[4] = {fmaf(-2.0f, ey, y), y - ey, y + ey, fmaf(2.0f, ey, y)};
    const float wt[4] = {1.0f, -8.0f, 8.0f, -1.0f};
    for (int k = 0; k < 4; ++k) {
        scatter_bilinear(gref, H, W, xs[k], ys[k], giu * wt[k]);
        scatter_bilinear(gow, H, W, xs[k], ys[k], giu * wt[k]);
    }
}

// derivative of a bilinear sample with respect to its coordinates
__device__ __forceinline__ float2 dsample1(const float* __restrict__ img, int H, int W, float x, float y)
{
    const Tap t = make_tap(x, y, H, W);
    const float t00 = img[t.i00], t10 = img[t.i10], t01 = img[t.i01], t11 = img[t.i11];
    const float r0 = fmaf(t.ax, t10 - t00, t00), r1 = fmaf(t.ax, t11 - t01, t01);
    return make_float2((t10 - t00) + t.ay * ((t11 - t01) - (t10 - t00)), r1 - r0);
}

// Reverse of warpingImage / findWarpingVector at flow_w (Python/vfs.cu:186-202,283-302):
// gf += gow * dS(other)/dx + gtv . dS(tvf)/dx.
__global__ void __launch_bounds__(256) k_adj_warp(float2* __restrict__ gf, const float* __restrict__ gow,
                                                  const float2* __restrict__ gtv, const float2* __restrict__ flow,
                                                  const float* __restrict__ other, const float2* __restrict__ tvf,
                                                  int H, int W)
{
    VFS_ADJ_PIXEL(H, W)
    const float2 f = flow[idx];
    const float xs = X + f.x, ys = Y + f.y;
    float2 g = gf[idx];
    const float go = gow[idx];
    if (go != 0.0f) {
        const float2 d = dsample1(other, H, W, xs, ys);
        g.x += go * d.x;
        g.y += go * d.y;
    }
    const float2 gt = gtv[idx];
    if (gt.x != 0.0f || gt.y != 0.0f) {
        const Tap t = make_tap(xs, ys, H, W);
        const float2 a = tvf[t.i00], b = tvf[t.i10], c = tvf[t.i01], d = tvf[t.i11];
        // channel x
        {
            const float r0 = fmaf(t.ax, b.x - a.x, a.x), r1 = fmaf(t.ax, d.x - c.x, c.x);
            g.x += gt.x * ((b.x - a.x) + t.ay * ((d.x - c.x) - (b.x - a.x)));
            g.y += gt.x * (r1 - r0);
        }
        {
            const float r0 = fmaf(t.ax, b.y - a.y, a.y), r1 = fmaf(t.ax, d.y - c.y, c.y);
            g.x += gt.y * ((b.y - a.y) + t.ay * ((d.y - c.y) - (b.y - a.y)));
            g.y += gt.y * (r1 - r0);
        }
    }
    gf[idx] = g;
}

// ---------------------------------------------------------------------------
// per-level reverse: tensor, Gaussian, pyramid, upsampling
// ---------------------------------------------------------------------------

// Reverse of computeTensorEtaMasked (Python/vfs.cu:228-276): the adjoints of abc and
// eta go to the smoothed image (atomics on its three taps) and to beta, gamma,
// alpha0, alpha1 (per-pixel accumulators).
__global__ void __launch_bounds__(256) k_adj_tensor(float* __restrict__ gimg, float* __restrict__ gpar,
                                                    const float* __restrict__ gabc, const float* __restrict__ geta,
                                                    const float* __restrict__ img, const uint8_t* __restrict__ mask,
                                                    int H, int W, float beta, float gamma, float alpha0, float alpha1)
{
    VFS_ADJ_PIXEL(H, W)
    if (mask[idx] == 0) return;
    const int xp = clamp_idx(X + 1, W), yp = clamp_idx(Y + 1, H);
    const int i11 = yp * W + xp, i01 = yp * W + X, i10 = Y * W + xp;
    const float gx = img[i11] - img[i01];
    const float gy = img[i11] - img[i10];
    const float ng = hypotf(gx, gy);
    const bool small = ng < 1e-8f;
    const float nx = small ? 1.0f : gx / ng, ny = small ? 0.0f : gy / ng;
    const float P = ng > 0.0f ? powf(ng, gamma) : 0.0f;
    const float ex = expf(-beta * P);
    const bool wact = ex >= 1e-8f;
    const float w = wact ? ex : 1e-8f;
    const float a = w * nx * nx + ny * ny;
    const float b = w * ny * ny + nx * nx;
    const float c = w * nx * ny - ny * nx;
    const float a11 = alpha1 * alpha1;
    float ga = gabc[3 * idx], gb = gabc[3 * idx + 1], gc = gabc[3 * idx + 2];
    const float ge0 = geta[3 * idx], ge1 = geta[3 * idx + 1], ge2 = geta[3 * idx + 2];
    ga += ge0 * a11 * (2.0f * a + 2.0f * (a + c)) + ge2 * 2.0f * a11 * a;
    gb += ge0 * a11 * (2.0f * b + 2.0f * (b + c)) + ge1 * 2.0f * a11 * b;
    gc += ge0 * a11 * (4.0f * c + 2.0f * (a + c) + 2.0f * (b + c)) + ge1 * 2.0f * a11 * c + ge2 * 2.0f * a11 * c;
    const float g_a11 = ge0 * (a * a + b * b + 2.0f * c * c + (a + c) * (a + c) + (b + c) * (b + c)) +
                        ge1 * (b * b + c * c) + ge2 * (a * a + c * c);
    float* gpp = gpar + GP_N * idx;
    gpp[GP_ALPHA1] += g_a11 * 2.0f * alpha1;
    gpp[GP_ALPHA0] += (ge1 + ge2) * 8.0f * alpha0;
    const float g_w = ga * nx * nx + gb * ny * ny + gc * nx * ny;
    float g_nx = ga * 2.0f * w * nx + gb * 2.0f * nx + gc * (w - 1.0f) * ny;
    float g_ny = ga * 2.0f * ny + gb * 2.0f * w * ny + gc * (w - 1.0f) * nx;
    float g_ng = 0.0f;
    if (wact) {
        const float g_P = -g_w * ex * beta;
        gpp[GP_BETA] += -g_w * ex * P;
        if (ng > 0.0f) {
            g_ng += g_P * gamma * P / ng;
            gpp[GP_GAMMA] += g_P * P * logf(ng);
        }
    }
    float g_gx = 0.0f, g_gy = 0.0f;
    if (!small) {
        const float dn = nx * g_nx + ny * g_ny;
        g_gx = (g_nx - nx * dn) / ng;
        g_gy = (g_ny - ny * dn) / ng;
    }
    if (ng > 0.0f) {
        g_gx += g_ng * gx / ng;
        g_gy += g_ng * gy / ng;
    }
    atomicAdd(&gimg[i11], g_gx + g_gy);
    atomicAdd(&gimg[i01], -g_gx);
    atomicAdd(&gimg[i10], -g_gy);
}

// Reverse of gaussianMasked (Python/vfs.cu:81-102,204-226).
__global__ void __launch_bounds__(256) k_adj_gaussian(float* __restrict__ gimg, const float* __restrict__ gout,
                                                      const uint8_t* __restrict__ mask, int H, int W)
{
    VFS_ADJ_PIXEL(H, W)
    if (mask[idx] == 0) return;
    const float g = gout[idx];
    if (g == 0.0f) return;
    const int xm = clamp_idx(X - 1, W), xp = clamp_idx(X + 1, W);
    const int ym = clamp_idx(Y - 1, H), yp = clamp_idx(Y + 1, H);
    atomicAdd(&gimg[idx], 0.25f * g);
    const float g1 = 0.0625f * g, g2 = 0.125f * g;
    atomicAdd(&gimg[ym * W + xm], g1);
    atomicAdd(&gimg[yp * W + xm], g1);
    atomicAdd(&gimg[ym * W + xp], g1);
    atomicAdd(&gimg[yp * W + xp], g1);
    atomicAdd(&gimg[Y * W + xm], g2);
    atomicAdd(&gimg[Y * W + xp], g2);
    atomicAdd(&gimg[ym * W + X], g2);
    atomicAdd(&gimg[yp * W + X], g2);
}

// Reverse of downSamplingImage (four bilinear taps, Python/vfs.cu:35-49,171-174).
__global__ void __launch_bounds__(256) k_adj_downsample(float* __restrict__ gsrc, int Hs, int Ws,
                                                        const float* __restrict__ gout, int Hd, int Wd, float rx,
                                                        float ry)
{
    VFS_ADJ_PIXEL(Hd, Wd)
    const float g = 0.25f * gout[idx];
    if (g == 0.0f) return;
    const float xc = fmaf(X + 0.5f, rx, -0.5f), yc = fmaf(Y + 0.5f, ry, -0.5f);
    const float xm = fmaf(X + 0.25f, rx, -0.5f), xp = fmaf(X + 0.75f, rx, -0.5f);
    const float ym = fmaf(Y + 0.25f, ry, -0.5f), yp = fmaf(Y + 0.75f, ry, -0.5f);
    scatter_bilinear(gsrc, Hs, Ws, xm, yc, g);
    scatter_bilinear(gsrc, Hs, Ws, xp, yc, g);
    scatter_bilinear(gsrc, Hs, Ws, xc, ym, g);
    scatter_bilinear(gsrc, Hs, Ws, xc, yp, g);
}

// Reverse of upSamplingVectorField(1) (Python/vfs.cu:69-79,678-705) for `channels`
// interleaved channels: gsrc += scale * S^T gout.
__global__ void __launch_bounds__(256) k_adj_upsample(float* __restrict__ gsrc, int Hs, int Ws,
                                                      const float* __restrict__ gout, int Hd, int Wd, int channels,
                                                      float rx, float ry, float scale)
{
    VFS_ADJ_PIXEL(Hd, Wd)
    const float xs = fmaf(X + 0.5f, rx, -0.5f), ys = fmaf(Y + 0.5f, ry, -0.5f);
    const Tap t = make_tap(xs, ys, Hs, Ws);
    for (int ch = 0; ch < channels; ++ch) {
        const float g = scale * gout[channels * idx + ch];
        if (g == 0.0f) continue;
        atomicAdd(&gsrc[channels * t.i00 + ch], g * (1.0f - t.ax) * (1.0f - t.ay));
        atomicAdd(&gsrc[channels * t.i10 + ch], g * t.ax * (1.0f - t.ay));
        atomicAdd(&gsrc[channels * t.i01 + ch], g * (1.0f - t.ax) * t.ay);
        atomicAdd(&gsrc[channels * t.i11 + ch], g * t.ax * t.ay);
    }
}

// ---------------------------------------------------------------------------
// depth loss: L = mean over valid mask pixels of |log z - log z_gt|, with z the
// depth of triangulateRays (Python/vfs.cu:707-735) from the final flow
// ---------------------------------------------------------------------------

// x / z of the equidistant ray through (u, v) and its derivative in (u, v).
__device__ __forceinline__ float ray_x(float du, float dv, float f, float& dxdu, float& dxdv)
{
    const float r = hypotf(du, dv);
    if (r == 0.0f) {
        dxdu = 1.0f / f;
        dxdv = 0.0f;
        return 0.0f;
    }
    const float th = r / f;
    const float T = tanf(th) / r;
    const float ct = cosf(th);
    const float dT = (r / (f * ct * ct) - tanf(th)) / (r * r);  // dT/dr
    dxdu = T + du * dT * du / r;
    dxdv = du * dT * dv / r;
    return du * T;
}

// Per pixel: loss term, valid flag and the flow adjoint before normalisation.
__global__ void __launch_bounds__(256) k_depth_loss(float* __restrict__ term, float2* __restrict__ gflow,
                                                    const float2* __restrict__ flow, const float* __restrict__ z_gt,
                                                    const uint8_t* __restrict__ mask, int H, int W, float f, float u0,
                                                    float v0, float tx, float tz)
{
    VFS_ADJ_PIXEL(H, W)
    term[2 * idx] = 0.0f;
    term[2 * idx + 1] = 0.0f;
    gflow[idx] = make_float2(0.0f, 0.0f);
    if (mask[idx] == 0) return;
    float a, b, c, d;
    const float xr = ray_x(X - u0, Y - v0, f, a, b);
    const float2 fl = flow[idx];
    const float xo = ray_x(X + fl.x - u0, Y + fl.y - v0, f, c, d);
    const float den = xo - xr;
    const float z = (tx - tz * xo) / den;
    const float zg = z_gt[idx];
    if (!(z > 0.0f) || !(zg > 0.0f) || !isfinite(z)) return;
    const float e = logf(z) - logf(zg);
    term[2 * idx] = fabsf(e);
    term[2 * idx + 1] = 1.0f;
    const float s = e > 0.0f ? 1.0f : (e < 0.0f ? -1.0f : 0.0f);
    const float dz_dxo = (tz * xr - tx) / (den * den);
    const float g = s / z * dz_dxo;
    gflow[idx] = make_float2(g * c, g * d);
}

// Two-stage deterministic sums: per-block partials in double, then one block.
__global__ void __launch_bounds__(256) k_reduce_partial(double* __restrict__ part, const float* __restrict__ in,
                                                        int64_t n, int channels)
{
    __shared__ double sm[256];
    for (int ch = 0; ch < channels; ++ch) {
        double acc = 0.0;
        for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * blockDim.x)
            acc += static_cast<double>(in[i * channels + ch]);
        // wavefront reduction, then across the four wavefronts of the block
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += sm[w];
            part[blockIdx.x * channels + ch] = t;
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_reduce_final(double* __restrict__ out, const double* __restrict__ part,
                                                      int nparts, int channels)
{
    __shared__ double sm[256];
    for (int ch = 0; ch < channels; ++ch) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < nparts; i += blockDim.x) acc += part[i * channels + ch];
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
        __syncthreads();
        if (threadIdx.x == 0) {
            double t = 0.0;
            for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) t += sm[w];
            out[ch] = t;
        }
        __syncthreads();
    }
}

// Per-block partials of the masked residual: sum of (a - b)^2 over `channels`
// interleaved channels and the number of mask pixels, in double.
__global__ void __launch_bounds__(256) k_residual_partial(double* __restrict__ part, const float* __restrict__ a,
                                                          const float* __restrict__ b,
                                                          const uint8_t* __restrict__ mask, int64_t n, int channels)
{
    __shared__ double sm[2][4];
    double ss = 0.0, cnt = 0.0;
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (mask != nullptr && mask[i] == 0) continue;
        float acc = 0.0f;
        for (int c = 0; c < channels; ++c) {
            const float d = a[i * channels + c] - b[i * channels + c];
            acc = fmaf(d, d, acc);
        }
        ss += static_cast<double>(acc);
        cnt += 1.0;
    }
    for (int off = 32; off > 0; off >>= 1) {
        ss += __shfl_down(ss, off, 64);
        cnt += __shfl_down(cnt, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sm[0][threadIdx.x >> 6] = ss;
        sm[1][threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double t0 = 0.0, t1 = 0.0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            t0 += sm[0][w];
            t1 += sm[1][w];
        }
        part[2 * blockIdx.x] = t0;
        part[2 * blockIdx.x + 1] = t1;
    }
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int vfs_adj_pd_iteration(void* stream, const vfs_adj_pd* a, const vfs_params* prm, float tau, float sigma,
                         float mu, int has_msgs)
{
    if (a == nullptr || prm == nullptr) return VFS_EINVAL;
    VFS_CHECK_DIMS(a->H, a->W);
    const void* ptrs[] = {a->gu, a->gub, a->gvb, a->gp, a->gq, a->um, a->vm, a->qmsg, a->tmsg, a->gabc, a->geta,
                          a->gderiv, a->gpar, a->p, a->q, a->u, a->u_bar_prev, a->v_bar_prev, a->p_prev, a->q_prev,
                          a->u_prev, a->abc, a->eta, a->deriv, a->mask};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    const int H = a->H, W = a->W;
    hipStream_t s = S(stream);
    hipLaunchKernelGGL(k_adj_primal, grid2d(H, W), block2d(), 0, s, a->gub, reinterpret_cast<float2*>(a->gvb), a->um,
                       a->vm, has_msgs, reinterpret_cast<float2*>(a->gp), reinterpret_cast<float4*>(a->gq),
                       reinterpret_cast<float4*>(a->qmsg), a->gu, a->gabc, a->geta, a->gpar,
                       reinterpret_cast<const float2*>(a->p), reinterpret_cast<const float4*>(a->q), a->abc, a->eta,
                       a->mask, H, W, tau, mu, prm->alpha0, prm->alpha1);
    hipLaunchKernelGGL(k_adj_l1, grid2d(H, W), block2d(), 0, s, a->gu, reinterpret_cast<float4*>(a->tmsg), a->gub,
                       a->gderiv, a->geta, a->gpar, reinterpret_cast<const float2*>(a->p), a->abc, a->u_prev,
                       a->u_bar_prev, reinterpret_cast<const float2*>(a->deriv), a->eta, a->mask, H, W, tau,
                       prm->lambda);
    hipLaunchKernelGGL(k_adj_dual, grid2d(H, W), block2d(), 0, s, reinterpret_cast<float2*>(a->gp),
                       reinterpret_cast<float4*>(a->gq), reinterpret_cast<const float4*>(a->qmsg),
                       reinterpret_cast<const float4*>(a->tmsg), a->um, a->vm, a->gabc, a->gpar,
                       reinterpret_cast<const float2*>(a->p_prev), reinterpret_cast<const float4*>(a->q_prev),
                       a->u_bar_prev, reinterpret_cast<const float2*>(a->v_bar_prev), a->abc, a->mask, H, W,
                       dual_steps(sigma, prm->alpha0, prm->alpha1, prm->eta_p, prm->eta_q));
    VFS_LAUNCH_STATUS();
}

int vfs_adj_clamp(void* stream, float* gu, float* guf, int* pos, float* gtv, const float* gf, const float* u_th,
                  const float* u0, const float* tv, const uint8_t* mask, int H, int W, float limit, int ksize)
{
    VFS_CHECK_DIMS(H, W);
    const void* ptrs[] = {gu, guf, pos, gtv, gf, u_th, u0, tv, mask};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    if (ksize == 5)
        hipLaunchKernelGGL(k_adj_clamp<5>, grid2d(H, W), block2d(), 0, S(stream), gu, guf, pos,
                           reinterpret_cast<float2*>(gtv), reinterpret_cast<const float2*>(gf), u_th, u0,
                           reinterpret_cast<const float2*>(tv), mask, H, W, limit);
    else if (ksize == 3)
        hipLaunchKernelGGL(k_adj_clamp<3>, grid2d(H, W), block2d(), 0, S(stream), gu, guf, pos,
                           reinterpret_cast<float2*>(gtv), reinterpret_cast<const float2*>(gf), u_th, u0,
                           reinterpret_cast<const float2*>(tv), mask, H, W, limit);
    else
        return VFS_EINVAL;
    VFS_LAUNCH_STATUS();
}

int vfs_adj_median(void* stream, float* gu_th, const float* guf, const int* pos, int H, int W, int ksize)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(gu_th);
    VFS_CHECK_PTR(guf);
    VFS_CHECK_PTR(pos);
    if (ksize == 5)
        hipLaunchKernelGGL(k_adj_median<5>, grid2d(H, W), block2d(), 0, S(stream), gu_th, guf, pos, H, W);
    else if (ksize == 3)
        hipLaunchKernelGGL(k_adj_median<3>, grid2d(H, W), block2d(), 0, S(stream), gu_th, guf, pos, H, W);
    else
        return VFS_EINVAL;
    VFS_LAUNCH_STATUS();
}

int vfs_adj_derivatives(void* stream, float* gref, float* gow, const float* gderiv, const float* tvf, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(gref);
    VFS_CHECK_PTR(gow);
    VFS_CHECK_PTR(gderiv);
    VFS_CHECK_PTR(tvf);
    hipLaunchKernelGGL(k_adj_derivatives, grid2d(H, W), block2d(), 0, S(stream), gref, gow, gderiv,
                       reinterpret_cast<const float2*>(tvf), H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_adj_warp(void* stream, float* gf, const float* gow, const float* gtv, const float* flow, const float* other,
                 const float* tvf, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    const void* ptrs[] = {gf, gow, gtv, flow, other, tvf};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    hipLaunchKernelGGL(k_adj_warp, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(gf), gow,
                       reinterpret_cast<const float2*>(gtv), reinterpret_cast<const float2*>(flow), other,
                       reinterpret_cast<const float2*>(tvf), H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_adj_tensor(void* stream, float* gimg, float* gpar, const float* gabc, const float* geta, const float* img,
                   const uint8_t* mask, int H, int W, const vfs_params* prm)
{
    VFS_CHECK_DIMS(H, W);
    const void* ptrs[] = {gimg, gpar, gabc, geta, img, mask, prm};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    hipLaunchKernelGGL(k_adj_tensor, grid2d(H, W), block2d(), 0, S(stream), gimg, gpar, gabc, geta, img, mask, H, W,
                       prm->beta, prm->gamma, prm->alpha0, prm->alpha1);
    VFS_LAUNCH_STATUS();
}

int vfs_adj_gaussian(void* stream, float* gimg, const float* gout, const uint8_t* mask, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(gimg);
    VFS_CHECK_PTR(gout);
    VFS_CHECK_PTR(mask);
    hipLaunchKernelGGL(k_adj_gaussian, grid2d(H, W), block2d(), 0, S(stream), gimg, gout, mask, H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_adj_downsample(void* stream, float* gsrc, int Hs, int Ws, const float* gout, int Hd, int Wd)
{
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_DIMS(Hd, Wd);
    VFS_CHECK_PTR(gsrc);
    VFS_CHECK_PTR(gout);
    hipLaunchKernelGGL(k_adj_downsample, grid2d(Hd, Wd), block2d(), 0, S(stream), gsrc, Hs, Ws, gout, Hd, Wd,
                       static_cast<float>(Ws) / static_cast<float>(Wd), static_cast<float>(Hs) / static_cast<float>(Hd));
    VFS_LAUNCH_STATUS();
}

int vfs_adj_upsample(void* stream, float* gsrc, int Hs, int Ws, const float* gout, int Hd, int Wd, int channels,
                     float f_scale)
{
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_DIMS(Hd, Wd);
    VFS_CHECK_PTR(gsrc);
    VFS_CHECK_PTR(gout);
    if (channels < 1 || channels > 4) return VFS_EINVAL;
    hipLaunchKernelGGL(k_adj_upsample, grid2d(Hd, Wd), block2d(), 0, S(stream), gsrc, Hs, Ws, gout, Hd, Wd, channels,
                       static_cast<float>(Ws) / static_cast<float>(Wd), static_cast<float>(Hs) / static_cast<float>(Hd),
                       f_scale);
    VFS_LAUNCH_STATUS();
}

int vfs_depth_loss(void* stream, float* term, float* gflow, const float* flow, const float* z_gt, const uint8_t* mask,
                   int H, int W, float f, float u0, float v0, float tx, float tz)
{
    VFS_CHECK_DIMS(H, W);
    const void* ptrs[] = {term, gflow, flow, z_gt, mask};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    hipLaunchKernelGGL(k_depth_loss, grid2d(H, W), block2d(), 0, S(stream), term, reinterpret_cast<float2*>(gflow),
                       reinterpret_cast<const float2*>(flow), z_gt, mask, H, W, f, u0, v0, tx, tz);
    VFS_LAUNCH_STATUS();
}

int vfs_reduce_sum(void* stream, double* out, double* partial, int nparts, const float* in, int64_t n, int channels)
{
    if (out == nullptr || partial == nullptr || in == nullptr || n < 0 || channels < 1 || nparts < 1) return VFS_EINVAL;
    hipLaunchKernelGGL(k_reduce_partial, dim3(nparts), dim3(256), 0, S(stream), partial, in, n, channels);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, S(stream), out, partial, nparts, channels);
    VFS_LAUNCH_STATUS();
}

int vfs_residual_norm(void* stream, double* out, double* partial, int nparts, const float* a, const float* b,
                      const uint8_t* mask, int64_t n, int channels)
{
    if (out == nullptr || partial == nullptr || a == nullptr || b == nullptr || n < 0 || channels < 1 ||
        nparts < 1 || nparts > 65535)
        return VFS_EINVAL;
    hipLaunchKernelGGL(k_residual_partial, dim3(nparts), dim3(256), 0, S(stream), partial, a, b, mask, n, channels);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, S(stream), out, partial, nparts, 2);
    VFS_LAUNCH_STATUS();
}

int vfs_adj_pd_finish(void* stream, float* gub_out, float* gvb_out, const float* gubB, const float* gvbB,
                      const float* um, const float* vm, const uint8_t* mask, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(gub_out);
    VFS_CHECK_PTR(gvb_out);
    VFS_CHECK_PTR(gubB);
    VFS_CHECK_PTR(gvbB);
    VFS_CHECK_PTR(um);
    VFS_CHECK_PTR(vm);
    VFS_CHECK_PTR(mask);
    hipLaunchKernelGGL(k_adj_pd_finish, grid2d(H, W), block2d(), 0, S(stream), gub_out,
                       reinterpret_cast<float2*>(gvb_out), gubB, reinterpret_cast<const float2*>(gvbB), um, vm, mask,
                       H, W);
    VFS_LAUNCH_STATUS();
}

}  // extern "C"
