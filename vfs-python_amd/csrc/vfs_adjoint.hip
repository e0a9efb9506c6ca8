// Hand-written adjoint (reverse mode) of the VFS solve, gfx950.
//
// Each forward kernel of the solve has a reverse kernel here.  The primal-dual
// iteration is reversed step by step (primal, threshold, dual), re-reading the
// forward state of the iteration from a tape that the host recomputes from
// checkpoints.  Contributions to a neighbour's adjoint are written as per-pixel
// messages and collected by the receiving pixel ("pull").  The transposes of the
// fixed stencils and bilinear taps (derivative taps, tensor, Gaussian, pyramid,
// upsampling) are gathers too: each receiving pixel sums its senders' shares in a
// fixed order.  No kernel here uses atomics, so the gradients are the same bits on
// every run.  Gradients of the scalar parameters (lambda, alpha0, alpha1, beta,
// gamma) accumulate per pixel in a (H, W, 5) plane, in the order of the reversed
// steps, and are reduced once at the end (vfs_reduce_sum).
//
// Non-smooth points follow one convention throughout: projections are active
// when the norm exceeds 1; max(exp, 1e-8) differentiates exp when exp >= 1e-8;
// min(d, limit) passes the gradient to d when d < limit; the threshold uses
// the branch its forward took; the median routes to the selected element.
#include <mutex>

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

// Contribution of one bilinear tap (weight v at texel coordinate (x, y)) to the
// adjoint of texel r: the tap's corner weights of every corner that is r, in the
// order i00, i10, i01, i11.  A non-finite coordinate or a zero v contributes
// nothing.  The transposes below are gathers built on this: each receiving texel
// recomputes the taps of every sender that can reach it and adds their shares in
// a fixed order, so the sums do not depend on scheduling.
__device__ __forceinline__ float tap_share(int r, int H, int W, float x, float y, float v)
{
    if (v == 0.0f || !(x == x) || !(y == y)) return 0.0f;
    const Tap t = make_tap(x, y, H, W);
    float acc = 0.0f;
    if (t.i00 == r) acc += v * (1.0f - t.ax) * (1.0f - t.ay);
    if (t.i10 == r) acc += v * t.ax * (1.0f - t.ay);
    if (t.i01 == r) acc += v * (1.0f - t.ax) * t.ay;
    if (t.i11 == r) acc += v * t.ax * t.ay;
    return acc;
}

// Reverse of computeDerivatives (Python/vfs.cu:120-142).  The sender at pixel s
// weights its 5-point taps s + k e (k = -2, -1, 1, 2; e the unit tvf at s) of both
// the reference level image and the warped image by giu * (1, -8, 8, -1), giu =
// g(Iu) / 24, and its centre by -g(Iz) (reference) and +g(Iz) (warped).  |k e| <= 2,
// so a tap's corners lie within [s - 2, s + 3] per axis and, after the mirror,
// every sender of texel r lies within r +- 3.  gref += T - g(Iz), gow += T + g(Iz),
// with T the taps' shares at r.
//
// A block owns a 64 x 16 tile of texels and visits the senders of the tile extended by
// 3 on every side.  Each sender forms its taps' corner shares in float as tap_share
// does (v (1 - ax) (1 - ay), ...) and adds those that fall on the tile into int64
// fixed-point sums in LDS: share x 2^S, rounded to the nearest integer, with S chosen
// per block from the largest |giu| M of its senders so that one share is below 2^54
// and a texel's at most 196 shares sum below 2^62.  Integer sums do not depend on the
// order of the additions, so the result is deterministic; T is the sum rounded once
// to float (shares below 2^-S-1, under 2^-57 M, round to zero).  A non-finite share
// makes its texel's T NaN; taps at non-finite coordinates add nothing.
namespace {
constexpr int DR = 3, AD_TX = 64, AD_TY = 16, AD_NT = 256;
constexpr int AD_SX = AD_TX + 2 * DR, AD_SY = AD_TY + 2 * DR, AD_SN = (AD_SX * AD_SY + AD_NT - 1) / AD_NT;

// the k-th tap of sender (x, y) with unit direction (ex, ey), as derivative_pixel forms it
__device__ __forceinline__ void deriv_tap(int k, float x, float y, float ex, float ey, float& tx, float& ty, float& wt)
{
    if (k == -2) {
        tx = fmaf(-2.0f, ex, x);
        ty = fmaf(-2.0f, ey, y);
        wt = 1.0f;
    } else if (k == -1) {
        tx = x - ex;
        ty = y - ey;
        wt = -8.0f;
    } else if (k == 1) {
        tx = x + ex;
        ty = y + ey;
        wt = 8.0f;
    } else {
        tx = fmaf(2.0f, ex, x);
        ty = fmaf(2.0f, ey, y);
        wt = -1.0f;
    }
}

// one corner share: into the tile's fixed-point sum if the corner is a texel of the tile
__device__ __forceinline__ void add_share(unsigned long long* acc, unsigned* bad, int cx, int cy, int X0, int Y0,
                                          float share, double scale)
{
    const int lx = cx - X0, ly = cy - Y0;
    if (lx < 0 || lx >= AD_TX || ly < 0 || ly >= AD_TY) return;
    const int j = ly * AD_TX + lx;
    if (!isfinite(share)) {
        atomicOr(&bad[j], 1u);
        return;
    }
    const long long q = llrint(static_cast<double>(share) * scale);
    atomicAdd(&acc[j], static_cast<unsigned long long>(q));
}
}  // namespace

__global__ void __launch_bounds__(AD_NT) k_adj_derivatives(float* __restrict__ gref, float* __restrict__ gow,
                                                           const float* __restrict__ gderiv,
                                                           const float2* __restrict__ tvf, int H, int W)
{
    __shared__ unsigned long long s_acc[AD_TX * AD_TY];
    __shared__ unsigned s_bad[AD_TX * AD_TY];
    __shared__ float s_wmax[AD_NT / 64];
    const int X0 = blockIdx.x * AD_TX, Y0 = blockIdx.y * AD_TY;
    const int tid = threadIdx.x;
    for (int j = tid; j < AD_TX * AD_TY; j += AD_NT) {
        s_acc[j] = 0ull;
        s_bad[j] = 0u;
    }
    // the senders' giu (g(Iu) / 24) and the largest finite |giu|
    float g[AD_SN];
    float m = 0.0f;
#pragma unroll
    for (int i = 0; i < AD_SN; ++i) {
        const int e = tid + AD_NT * i;
        const int sx = X0 - DR + e % AD_SX, sy = Y0 - DR + e / AD_SX;
        g[i] = 0.0f;
        if (e < AD_SX * AD_SY && sx >= 0 && sx < W && sy >= 0 && sy < H) {
            g[i] = gderiv[2 * (sy * W + sx)] / 24.0f;
            if (isfinite(g[i])) m = fmaxf(m, fabsf(g[i]));
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((tid & 63) == 0) s_wmax[tid >> 6] = m;
    __syncthreads();
    m = s_wmax[0];
#pragma unroll
    for (int w = 1; w < AD_NT / 64; ++w) m = fmaxf(m, s_wmax[w]);
    // |share| <= 8 M < 2^(e + 3) for M < 2^e: S = 54 - (e + 3)
    int ex2 = 0;
    frexpf(m, &ex2);
    const double scale = m > 0.0f ? ldexp(1.0, 51 - ex2) : 1.0;
#pragma unroll
    for (int i = 0; i < AD_SN; ++i) {
        const int e = tid + AD_NT * i;
        if (g[i] == 0.0f) continue;  // also every sender outside the image
        const int sx = X0 - DR + e % AD_SX, sy = Y0 - DR + e / AD_SX;
        const float2 vec = tvf[sy * W + sx];
        const float nrm = hypotf(vec.x, vec.y);
        const float ex = vec.x / nrm, ey = vec.y / nrm;
        for (int k = -2; k <= 2; ++k) {
            if (k == 0) continue;
            float tx, ty, wt;
            deriv_tap(k, static_cast<float>(sx), static_cast<float>(sy), ex, ey, tx, ty, wt);
            if (!(tx == tx) || !(ty == ty)) continue;
            const float v = g[i] * wt;
            const TapXY t = make_tap_xy(tx, ty, H, W);
            add_share(s_acc, s_bad, t.x0, t.y0, X0, Y0, v * (1.0f - t.ax) * (1.0f - t.ay), scale);
            add_share(s_acc, s_bad, t.x1, t.y0, X0, Y0, v * t.ax * (1.0f - t.ay), scale);
            add_share(s_acc, s_bad, t.x0, t.y1, X0, Y0, v * (1.0f - t.ax) * t.ay, scale);
            add_share(s_acc, s_bad, t.x1, t.y1, X0, Y0, v * t.ax * t.ay, scale);
        }
    }
    __syncthreads();
    for (int j = tid; j < AD_TX * AD_TY; j += AD_NT) {
        const int X = X0 + j % AD_TX, Y = Y0 + j / AD_TX;
        if (X >= W || Y >= H) continue;
        const int idx = Y * W + X;
        const float acc = s_bad[j] ? __builtin_nanf("")
                                   : static_cast<float>(static_cast<double>(static_cast<long long>(s_acc[j])) / scale);
        const float giz = gderiv[2 * idx + 1];
        gow[idx] += acc + giz;
        gref[idx] += acc - giz;
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

// Reverse of computeTensorEtaMasked (Python/vfs.cu:228-276) at the masked-in pixel
// (X, Y): the adjoints of its abc and eta give the adjoints (g_gx, g_gy) of its
// forward differences gx = I(X+1,Y+1) - I(X,Y+1), gy = I(X+1,Y+1) - I(X+1,Y) and,
// when gpp is not null, add its shares of beta, gamma, alpha0, alpha1 to gpp.
__device__ __forceinline__ float2 tensor_adj_pixel(int X, int Y, int H, int W, const float* __restrict__ gabc,
                                                   const float* __restrict__ geta, const float* __restrict__ img,
                                                   float beta, float gamma, float alpha0, float alpha1,
                                                   float* __restrict__ gpp)
{
    const int idx = Y * W + X;
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
    if (gpp != nullptr) {
        const float g_a11 = ge0 * (a * a + b * b + 2.0f * c * c + (a + c) * (a + c) + (b + c) * (b + c)) +
                            ge1 * (b * b + c * c) + ge2 * (a * a + c * c);
        gpp[GP_ALPHA1] += g_a11 * 2.0f * alpha1;
        gpp[GP_ALPHA0] += (ge1 + ge2) * 8.0f * alpha0;
    }
    const float g_w = ga * nx * nx + gb * ny * ny + gc * nx * ny;
    float g_nx = ga * 2.0f * w * nx + gb * 2.0f * nx + gc * (w - 1.0f) * ny;
    float g_ny = ga * 2.0f * ny + gb * 2.0f * w * ny + gc * (w - 1.0f) * nx;
    float g_ng = 0.0f;
    if (wact) {
        const float g_P = -g_w * ex * beta;
        if (gpp != nullptr) gpp[GP_BETA] += -g_w * ex * P;
        if (ng > 0.0f) {
            g_ng += g_P * gamma * P / ng;
            if (gpp != nullptr) gpp[GP_GAMMA] += g_P * P * logf(ng);
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
    return make_float2(g_gx, g_gy);
}

// Reverse of computeTensorEtaMasked: every masked-in pixel s sends g_gx + g_gy to
// i11(s) = (clamp(Y+1), clamp(X+1)), -g_gx to i01(s) = (clamp(Y+1), X) and -g_gy to
// i10(s) = (Y, clamp(X+1)), so the senders of texel r are the pixels r - (0 or 1,
// 0 or 1).  Gather: each texel adds the shares of its up to four senders (row-major,
// then i11, i01, i10), recomputing theirs; its own pass also adds its parameter
// shares to gpar.
__global__ void __launch_bounds__(256) k_adj_tensor(float* __restrict__ gimg, float* __restrict__ gpar,
                                                    const float* __restrict__ gabc, const float* __restrict__ geta,
                                                    const float* __restrict__ img, const uint8_t* __restrict__ mask,
                                                    int H, int W, float beta, float gamma, float alpha0, float alpha1)
{
    VFS_ADJ_PIXEL(H, W)
    float acc = 0.0f;
    for (int dy = -1; dy <= 0; ++dy) {
        const int sy = Y + dy;
        if (sy < 0) continue;
        for (int dx = -1; dx <= 0; ++dx) {
            const int sx = X + dx;
            if (sx < 0) continue;
            const int s = sy * W + sx;
            if (mask[s] == 0) continue;
            const bool own = dx == 0 && dy == 0;
            const float2 g = tensor_adj_pixel(sx, sy, H, W, gabc, geta, img, beta, gamma, alpha0, alpha1,
                                              own ? gpar + GP_N * idx : nullptr);
            const int xp = clamp_idx(sx + 1, W), yp = clamp_idx(sy + 1, H);
            if (yp * W + xp == idx) acc += g.x + g.y;
            if (yp * W + sx == idx) acc += -g.x;
            if (sy * W + xp == idx) acc += -g.y;
        }
    }
    gimg[idx] += acc;
}

// Reverse of gaussianMasked (Python/vfs.cu:81-102,204-226): every masked-in pixel s
// sends (1, 2, 1)^T (1, 2, 1) / 16 of its adjoint to its clamped 3 x 3 neighbourhood.
// Gather over the 3 x 3 senders of each texel (row-major), each sender's taps in the
// order centre, corners, edges.
__global__ void __launch_bounds__(256) k_adj_gaussian(float* __restrict__ gimg, const float* __restrict__ gout,
                                                      const uint8_t* __restrict__ mask, int H, int W)
{
    VFS_ADJ_PIXEL(H, W)
    float acc = 0.0f;
    for (int dy = -1; dy <= 1; ++dy) {
        const int sy = Y + dy;
        if (sy < 0 || sy >= H) continue;
        for (int dx = -1; dx <= 1; ++dx) {
            const int sx = X + dx;
            if (sx < 0 || sx >= W) continue;
            const int s = sy * W + sx;
            if (mask[s] == 0) continue;
            const float g = gout[s];
            if (g == 0.0f) continue;
            const int xm = clamp_idx(sx - 1, W), xp = clamp_idx(sx + 1, W);
            const int ym = clamp_idx(sy - 1, H), yp = clamp_idx(sy + 1, H);
            const float g1 = 0.0625f * g, g2 = 0.125f * g;
            if (s == idx) acc += 0.25f * g;
            if (ym * W + xm == idx) acc += g1;
            if (yp * W + xm == idx) acc += g1;
            if (ym * W + xp == idx) acc += g1;
            if (yp * W + xp == idx) acc += g1;
            if (sy * W + xm == idx) acc += g2;
            if (sy * W + xp == idx) acc += g2;
            if (ym * W + sx == idx) acc += g2;
            if (yp * W + sx == idx) acc += g2;
        }
    }
    gimg[idx] += acc;
}

// Senders s in [0, n_send) of a resampling whose sample coordinate is
// t = fmaf(s + off, ratio, -0.5) with off in [off_lo, off_hi] and whose corners can
// fall on texel r.  Corners are floor(t) and floor(t) + 1, within [-1, n] for these
// resamplings, and the mirror takes -1 to 0 and n to n - 1, so |t - r| <= 1:
// s + off in [(r - 0.5) / ratio, (r + 1.5) / ratio].  One more sender on each side
// absorbs the rounding of the bounds.
__device__ __forceinline__ void sender_range(int r, int n_send, float ratio, float off_lo, float off_hi, int& lo,
                                             int& hi)
{
    lo = static_cast<int>(floorf((static_cast<float>(r) - 0.5f) / ratio - off_hi)) - 1;
    hi = static_cast<int>(ceilf((static_cast<float>(r) + 1.5f) / ratio - off_lo)) + 1;
    lo = lo < 0 ? 0 : lo;
    hi = hi > n_send - 1 ? n_send - 1 : hi;
}

// Reverse of downSamplingImage (four bilinear taps, Python/vfs.cu:35-49,171-174):
// the coarse pixel s sends a quarter of its adjoint through each of its taps
// (s + 1/4, s + 1/2), (s + 3/4, s + 1/2), (s + 1/2, s + 1/4), (s + 1/2, s + 3/4) (in
// coarse pixels, scaled by the size ratio).  Gather over the fine texel's senders
// (row-major, taps in that order).
__global__ void __launch_bounds__(256) k_adj_downsample(float* __restrict__ gsrc, int Hs, int Ws,
                                                        const float* __restrict__ gout, int Hd, int Wd, float rx,
                                                        float ry)
{
    VFS_ADJ_PIXEL(Hs, Ws)
    int xlo, xhi, ylo, yhi;
    sender_range(X, Wd, rx, 0.25f, 0.75f, xlo, xhi);
    sender_range(Y, Hd, ry, 0.25f, 0.75f, ylo, yhi);
    float acc = 0.0f;
    for (int sy = ylo; sy <= yhi; ++sy) {
        for (int sx = xlo; sx <= xhi; ++sx) {
            const float g = 0.25f * gout[sy * Wd + sx];
            if (g == 0.0f) continue;
            const float xc = fmaf(sx + 0.5f, rx, -0.5f), yc = fmaf(sy + 0.5f, ry, -0.5f);
            const float xm = fmaf(sx + 0.25f, rx, -0.5f), xp = fmaf(sx + 0.75f, rx, -0.5f);
            const float ym = fmaf(sy + 0.25f, ry, -0.5f), yp = fmaf(sy + 0.75f, ry, -0.5f);
            acc += tap_share(idx, Hs, Ws, xm, yc, g);
            acc += tap_share(idx, Hs, Ws, xp, yc, g);
            acc += tap_share(idx, Hs, Ws, xc, ym, g);
            acc += tap_share(idx, Hs, Ws, xc, yp, g);
        }
    }
    gsrc[idx] += acc;
}

// Reverse of upSamplingVectorField(1) (Python/vfs.cu:69-79,678-705) for `channels`
// interleaved channels: gsrc += scale * S^T gout.  The fine pixel s samples the
// coarse level at (s + 1/2) (Ws / Wd) - 1/2; gather over the coarse texel's senders
// (row-major).
__global__ void __launch_bounds__(256) k_adj_upsample(float* __restrict__ gsrc, int Hs, int Ws,
                                                      const float* __restrict__ gout, int Hd, int Wd, int channels,
                                                      float rx, float ry, float scale)
{
    VFS_ADJ_PIXEL(Hs, Ws)
    int xlo, xhi, ylo, yhi;
    sender_range(X, Wd, rx, 0.5f, 0.5f, xlo, xhi);
    sender_range(Y, Hd, ry, 0.5f, 0.5f, ylo, yhi);
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int sy = ylo; sy <= yhi; ++sy) {
        for (int sx = xlo; sx <= xhi; ++sx) {
            const Tap t = make_tap(fmaf(sx + 0.5f, rx, -0.5f), fmaf(sy + 0.5f, ry, -0.5f), Hs, Ws);
            if (t.i00 != idx && t.i10 != idx && t.i01 != idx && t.i11 != idx) continue;
            const int s = sy * Wd + sx;
            for (int ch = 0; ch < channels; ++ch) {
                const float g = scale * gout[channels * s + ch];
                if (g == 0.0f) continue;
                if (t.i00 == idx) acc[ch] += g * (1.0f - t.ax) * (1.0f - t.ay);
                if (t.i10 == idx) acc[ch] += g * t.ax * (1.0f - t.ay);
                if (t.i01 == idx) acc[ch] += g * (1.0f - t.ax) * t.ay;
                if (t.i11 == idx) acc[ch] += g * t.ax * t.ay;
            }
        }
    }
    for (int ch = 0; ch < channels; ++ch) gsrc[channels * idx + ch] += acc[ch];
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

// Scratch of vfs_residual_norm_noalloc: the per-workgroup partials (sum of squares,
// pixel count) of at most VFS_NOALLOC_MAX_PARTS workgroups, 16 KiB.  Every device the
// library runs on holds its own copy until the process ends.  Host code never passes
// this name to a kernel: it is a host-side stand-in for the symbol, and the device
// address comes from hipGetSymbolAddress (noalloc_scratch below).
static __device__ __align__(256) double scratch[2 * VFS_NOALLOC_MAX_PARTS];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's copy of `scratch`, resolved once per device
// (the first call on that device) and kept.
int noalloc_scratch(double** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return static_cast<int>(e);
    if (dev < 0 || dev >= kMaxDevices) return VFS_EINVAL;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return static_cast<int>(e);
        if (p == nullptr) return VFS_EINVAL;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

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
    const dim3 grid((W + AD_TX - 1) / AD_TX, (H + AD_TY - 1) / AD_TY, 1);
    hipLaunchKernelGGL(k_adj_derivatives, grid, dim3(AD_NT), 0, S(stream), gref, gow, gderiv,
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
    hipLaunchKernelGGL(k_adj_downsample, grid2d(Hs, Ws), block2d(), 0, S(stream), gsrc, Hs, Ws, gout, Hd, Wd,
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
    hipLaunchKernelGGL(k_adj_upsample, grid2d(Hs, Ws), block2d(), 0, S(stream), gsrc, Hs, Ws, gout, Hd, Wd, channels,
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

int vfs_residual_norm_noalloc(void* stream, double* out, int nparts, const float* a, const float* b,
                              const uint8_t* mask, int64_t n, int channels)
{
    if (out == nullptr || a == nullptr || b == nullptr || n < 0 || channels < 1 || nparts < 1) return VFS_EINVAL;
    if (nparts > VFS_NOALLOC_MAX_PARTS) return VFS_ETOOBIG;
    double* part = nullptr;
    const int rc = noalloc_scratch(&part);
    if (rc != 0) return rc;
    // every partial slot the final block reads (0 .. nparts - 1) is written by this
    // launch's first kernel, so the scratch needs no clearing
    hipLaunchKernelGGL(k_residual_partial, dim3(nparts), dim3(256), 0, S(stream), part, a, b, mask, n, channels);
    hipLaunchKernelGGL(k_reduce_final, dim3(1), dim3(256), 0, S(stream), out, part, nparts, 2);
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
