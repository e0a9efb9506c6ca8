// Per-pixel arithmetic of the reverse primal-dual iteration, shared by the
// one-kernel-per-step path (vfs_adjoint.hip) and the fused level sweep
// (vfs_adjoint_solver.hip), so both compute the same values.
//
// The forward iteration t (Python/vfs.py:505-511) is dual (Python/vfs.cu:326-419),
// threshold (:421-505), primal (:507-581).  Its reverse runs primal^T,
// threshold^T, dual^T.  A contribution to a neighbour's adjoint is written as a
// per-pixel message and collected by the receiving pixel, never scattered.
//
// Non-smooth points: projections are active when the squared norm exceeds 1; the
// threshold uses the branch its forward took.  Forward values are recomputed with
// the same arithmetic as vfs_pd.h (reciprocal steps, rsqrt projection scale).
#pragma once

#include "vfs_device.h"
#include "vfs_pd.h"

namespace vfs {

enum { GP_LAMBDA = 0, GP_ALPHA0 = 1, GP_ALPHA1 = 2, GP_BETA = 3, GP_GAMMA = 4, GP_N = 5 };

struct Nbr {
    int iE, iW, iS, iN;
    bool m, mE, mW, mS, mN;
};

__device__ __forceinline__ Nbr nbr(const uint8_t* __restrict__ mask, int X, int Y, int H, int W)
{
    Nbr n;
    n.iE = Y * W + clamp_idx(X + 1, W);
    n.iW = Y * W + clamp_idx(X - 1, W);
    n.iS = clamp_idx(Y + 1, H) * W + X;
    n.iN = clamp_idx(Y - 1, H) * W + X;
    n.m = mask[Y * W + X] != 0;
    n.mE = mask[n.iE] != 0;
    n.mW = mask[n.iW] != 0;
    n.mS = mask[n.iS] != 0;
    n.mN = mask[n.iN] != 0;
    return n;
}

// x / y neighbour whose u_bar, v_bar the dual step of pixel (X, Y) reads
__device__ __forceinline__ int dual_x_target(const uint8_t* __restrict__ mask, int X, int Y, int W)
{
    const int iE = Y * W + clamp_idx(X + 1, W);
    return mask[iE] != 0 ? iE : Y * W + clamp_idx(X - 1, W);
}

__device__ __forceinline__ int dual_y_target(const uint8_t* __restrict__ mask, int X, int Y, int H, int W)
{
    const int iS = clamp_idx(Y + 1, H) * W + X;
    return mask[iS] != 0 ? iS : clamp_idx(Y - 1, H) * W + X;
}

// Dual^T messages of one pixel: um = (own, to x-target, to y-target) for u_bar,
// vm = (own.x, own.y, x-target.x, x-target.y, y-target.x, y-target.y) for v_bar.
struct DualMsg {
    float um[3];
    float vm[6];
};

// Collect the messages addressed to pixel (X, Y), in the order: own, the x
// senders X-1, X, X+1, the y senders Y-1, Y, Y+1.  msg(i) returns the DualMsg of
// pixel index i (from global memory or from a tile in LDS).
template <typename MsgAt>
__device__ __forceinline__ void gather_msgs(const uint8_t* __restrict__ mask, MsgAt msg, int X, int Y, int H, int W,
                                            float& gub, float2& gvb)
{
    const int j = Y * W + X;
    {
        const DualMsg m = msg(X, Y);
        gub += m.um[0];
        gvb.x += m.vm[0];
        gvb.y += m.vm[1];
    }
    for (int dx = -1; dx <= 1; ++dx) {
        const int xs = X + dx;
        if (xs < 0 || xs >= W) continue;
        const int i = Y * W + xs;
        if (mask[i] == 0 || dual_x_target(mask, xs, Y, W) != j) continue;
        const DualMsg m = msg(xs, Y);
        gub += m.um[1];
        gvb.x += m.vm[2];
        gvb.y += m.vm[3];
    }
    for (int dy = -1; dy <= 1; ++dy) {
        const int ys = Y + dy;
        if (ys < 0 || ys >= H) continue;
        const int i = ys * W + X;
        if (mask[i] == 0 || dual_y_target(mask, X, ys, H, W) != j) continue;
        const DualMsg m = msg(X, ys);
        gub += m.um[2];
        gvb.x += m.vm[4];
        gvb.y += m.vm[5];
    }
}

// Reverse of updatePrimalVariables (Python/vfs.cu:507-570) at a masked-in pixel.
// g_ubo / g_vbo: complete adjoints of u_bar_t / v_bar_t.  State: q_t at the pixel
// and its W / N neighbours (+0 where that neighbour is outside the mask, so that
// x - qW is x there and the forward's gQ needs no test), p_t, abc, eta.y, eta.z.
// Sets gubB / gvbB (own part of the adjoints of u_bar_{t-1} / v_bar_{t-1}), qmsg;
// adds to gu (adjoint of u_t), gp, gq, gabc, geta.y/.z, and the alpha shares.
__device__ __forceinline__ void primal_adj_pixel(const Nbr& n, float4 qc, float4 qW, float4 qN, float a, float b,
                                                 float c, float2 pc, float e1, float e2, float tau, float mu,
                                                 float alpha0, float alpha1, float g_ubo, float2 g_vbo,
                                                 float& gubB, float2& gvbB, float& gu, float2& gp, float4& gq,
                                                 float4& qmsg, float (&gabc)[3], float& geta1, float& geta2,
                                                 float& gpa0, float& gpa1, float2& qmsg_c)
{
    float4 gQ = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (n.mE) {
        gQ.x = qc.x;
        gQ.w = qc.w;
    }
    gQ.x -= qW.x;
    gQ.w -= qW.w;
    if (n.mS) {
        gQ.y = qc.y;
        gQ.z = qc.z;
    }
    gQ.y -= qN.y;
    gQ.z -= qN.z;
    const float2 tu = make_float2(fmaf(a, pc.x, c * pc.y), fmaf(c, pc.x, b * pc.y));
    const bool d1 = !(fabsf(e1) < kEps), d2 = !(fabsf(e2) < kEps);
    const float r1 = recip_eta(e1, kEps), r2 = recip_eta(e2, kEps);
    const float t1 = tau * r1;
    const float t2 = tau * r2;
    const float S1 = gQ.x + gQ.z, S2 = gQ.w + gQ.y;
    const float A1 = fmaf(alpha1, tu.x, alpha0 * S1), A2 = fmaf(alpha1, tu.y, alpha0 * S2);
    // ub_out = (1 + mu) u - mu ub ; vb_out = (1 + mu) v - mu vb ; v = t A + vb
    const float g_u = (1.0f + mu) * g_ubo;
    const float g_vx = (1.0f + mu) * g_vbo.x, g_vy = (1.0f + mu) * g_vbo.y;
    gubB = -mu * g_ubo;
    gvbB = make_float2(g_vx - mu * g_vbo.x, g_vy - mu * g_vbo.y);
    gu += g_u;
    const float g_A1 = g_vx * t1, g_A2 = g_vy * t2;
    const float g_t1 = g_vx * A1, g_t2 = g_vy * A2;
    gpa1 += g_A1 * tu.x + g_A2 * tu.y;
    gpa0 += g_A1 * S1 + g_A2 * S2;
    // d t / d eta = -tau / eta^2 = -t / eta
    if (d1) geta1 += -g_t1 * t1 * r1;
    if (d2) geta2 += -g_t2 * t2 * r2;
    const float2 g_tu = make_float2(g_A1 * alpha1, g_A2 * alpha1);
    gp.x += a * g_tu.x + c * g_tu.y;
    gp.y += c * g_tu.x + b * g_tu.y;
    gabc[0] += g_tu.x * pc.x;
    gabc[1] += g_tu.y * pc.y;
    gabc[2] += g_tu.x * pc.y + g_tu.y * pc.x;
    const float g_S1 = g_A1 * alpha0, g_S2 = g_A2 * alpha0;
    if (n.mE) {
        gq.x += g_S1;
        gq.w += g_S2;
    }
    if (n.mS) {
        gq.z += g_S1;
        gq.y += g_S2;
    }
    // messages: to W (x, w components of q_W), to N (y, z components of q_N).  The W / N
    // gating is the receiver's own mask, so qmsg_c = (-g_S1, -g_S2) carries the same
    // information for a receiver that knows it is inside the mask.
    qmsg = make_float4(n.mW ? -g_S1 : 0.0f, n.mW ? -g_S2 : 0.0f, n.mN ? -g_S2 : 0.0f, n.mN ? -g_S1 : 0.0f);
    qmsg_c = make_float2(-g_S1, -g_S2);
}

// T p of a pixel (zero where masked): the tp the forward threshold read.
__device__ __forceinline__ float2 tp_at(const float2* __restrict__ p, const float* __restrict__ abc,
                                        const uint8_t* __restrict__ mask, int i)
{
    if (mask[i] == 0) return make_float2(0.0f, 0.0f);
    const float2 pp = p[i];
    const float a = abc[3 * i], b = abc[3 * i + 1], c = abc[3 * i + 2];
    return make_float2(fmaf(a, pp.x, c * pp.y), fmaf(c, pp.x, b * pp.y));
}

// Reverse of l1Thresholding (Python/vfs.cu:421-495) at a masked-in pixel.
// gout: complete adjoint of u_t.  tc, tpWx, tpNy: T p_t at the pixel, x of its W
// neighbour, y of its N neighbour (+0 where that neighbour is outside the mask, so
// that tc.x - tpWx is tc.x there).  State: u_{t-1}, u_bar_{t-1}, deriv, eta.x.
// Returns the adjoint of u_{t-1}; sets tmsg = (own tp.x, own tp.y, to W (x), to N (y));
// adds to gubB, geta.x, gderiv and the lambda share.
__device__ __forceinline__ float l1_adj_pixel(const Nbr& n, float gout, float2 tc, float tpWx, float tpNy, float2 d,
                                              float e0, float uc, float ubc, float tau, float lambda, float4& tmsg,
                                              float& gubB, float& geta0, float2& gderiv, float& gplam, float& g_g_out)
{
    float cxc, cxw, cyc, cyn;
    if (n.mE && n.mW) {
        cxc = 1.0f;
        cxw = -1.0f;
    } else if (!n.mE) {
        cxc = 0.0f;
        cxw = -1.0f;
    } else {
        cxc = 1.0f;
        cxw = 0.0f;
    }
    if (n.mS && n.mN) {
        cyc = 1.0f;
        cyn = -1.0f;
    } else if (!n.mS) {
        cyc = 0.0f;
        cyn = -1.0f;
    } else {
        cyc = 1.0f;
        cyn = 0.0f;
    }
    const float gx = n.mE ? tc.x - tpWx : -tpWx;
    const float gy = n.mS ? tc.y - tpNy : -tpNy;
    const float dIu = d.x, dIz = d.y;
    const bool de = !(fabsf(e0) < kEps);
    const float r0 = recip_eta(e0, kEps);
    const float te = tau * r0;
    const float ut = ubc + te * (gx + gy);
    const float rho = dIu * (ut - uc) + dIz;
    const float bound = lambda * te * dIu * dIu;
    const float d0 = ut - uc;
    // The branch the forward took, with each branch's expressions as written for it,
    // formed everywhere and selected (no divergent branch): inside the band with
    // |dIu| >= eps, du = d0 - rho r (r = 1 / dIu); inside with |dIu| < eps, du = d0;
    // below the band du = d0 + lambda te dIu, above it du = d0 - lambda te dIu; none of
    // them for a NaN rho (du = 0).
    const bool in_band = (rho <= bound) && (rho >= -bound);
    const bool lo = !in_band && (rho < -bound);
    const bool hi = !in_band && !lo && (rho > bound);
    const bool flat = fabsf(dIu) < kEps;
    const float r = recip_grad(dIu, kEps);
    const float g_rho = -gout * r;
    const float gi_d0 = gout + g_rho * dIu;
    const float gi_dIu = gout * rho * r * r + g_rho * d0;
    const float glo_lam = gout * te * dIu, glo_te = gout * lambda * dIu, glo_dIu = gout * lambda * te;
    const float ghi_lam = -gout * te * dIu, ghi_te = -gout * lambda * dIu, ghi_dIu = -gout * lambda * te;
    const bool band_grad = in_band && !flat;
    const float g_d0 = band_grad ? gi_d0 : ((in_band || lo || hi) ? gout : 0.0f);
    const float g_dIu = band_grad ? gi_dIu : (lo ? glo_dIu : (hi ? ghi_dIu : 0.0f));
    const float g_dIz = band_grad ? g_rho : 0.0f;
    float g_te = lo ? glo_te : (hi ? ghi_te : 0.0f);
    const float g_lam = lo ? glo_lam : (hi ? ghi_lam : 0.0f);
    // out = uc + du, d0 = ut - uc, ut = ubc + te (gx + gy)
    const float g_ut = g_d0;
    gubB += g_ut;
    g_te += g_ut * (gx + gy);
    const float g_g = g_ut * te;
    if (de) geta0 += -g_te * te * r0;
    gderiv.x += g_dIu;
    gderiv.y += g_dIz;
    gplam += g_lam;
    // cxc = [E inside], cyc = [S inside]; for a receiver inside the mask cxw = cyn = -1,
    // so g_g alone carries the messages
    tmsg = make_float4(cxc * g_g, cyc * g_g, cxw * g_g, cyn * g_g);
    g_g_out = g_g;
    return gout - g_d0;
}

// Adjoints of q_t and tp_t at pixel idx completed with the messages of the
// primal^T and threshold^T steps (senders: the pixel itself at the image's W / N
// edge, its E and S neighbours).
__device__ __forceinline__ void collect_qt(const float4* __restrict__ qmsg, const float4* __restrict__ tmsg,
                                           const Nbr& n, int X, int Y, int H, int W, int idx, float4& gqn,
                                           float2& gtp)
{
    const float4 m0 = qmsg[idx];
    if (n.iW == idx) {
        gqn.x += m0.x;
        gqn.w += m0.y;
    }
    if (X + 1 < W) {
        const float4 me = qmsg[idx + 1];
        gqn.x += me.x;
        gqn.w += me.y;
    }
    if (n.iN == idx) {
        gqn.y += m0.z;
        gqn.z += m0.w;
    }
    if (Y + 1 < H) {
        const float4 ms = qmsg[idx + W];
        gqn.y += ms.z;
        gqn.z += ms.w;
    }
    const float4 t0 = tmsg[idx];
    gtp = make_float2(t0.x, t0.y);
    if (n.iW == idx) gtp.x += t0.z;
    if (X + 1 < W) gtp.x += tmsg[idx + 1].z;
    if (n.iN == idx) gtp.y += t0.w;
    if (Y + 1 < H) gtp.y += tmsg[idx + W].w;
}

// collect_qt for the compact messages qm = (-g_S1, -g_S2) and tm = g_g of the senders;
// valid for a receiver inside the mask (a masked receiver discards the result).
__device__ __forceinline__ void collect_qt_compact(const float2* __restrict__ qm, const float* __restrict__ tm,
                                                   const Nbr& n, int X, int Y, int H, int W, int idx, float4& gqn,
                                                   float2& gtp)
{
    const float2 q0 = qm[idx];
    if (X == 0) {
        gqn.x += q0.x;
        gqn.w += q0.y;
    }
    if (X + 1 < W) {
        const float2 me = qm[idx + 1];
        gqn.x += me.x;
        gqn.w += me.y;
    }
    if (Y == 0) {
        gqn.y += q0.y;
        gqn.z += q0.x;
    }
    if (Y + 1 < H) {
        const float2 ms = qm[idx + W];
        gqn.y += ms.y;
        gqn.z += ms.x;
    }
    const float g0 = tm[idx];
    gtp = make_float2(n.mE ? g0 : 0.0f, n.mS ? g0 : 0.0f);
    if (X == 0) gtp.x += -g0;
    if (X + 1 < W) gtp.x += -tm[idx + 1];
    if (Y == 0) gtp.y += -g0;
    if (Y + 1 < H) gtp.y += -tm[idx + W];
}

// Step sizes of the dual ascent of one iteration and their derivatives in alpha1 and
// alpha0: sp = dual_scale(alpha1, sigma, eta_p) as the forward formed it, dsp = d sp / d
// alpha1 = sigma / eta_p, and likewise sq, dsq.  Formed once per launch on the host.
struct DualSteps {
    float sp, sq, dsp, dsq;
};

__host__ __device__ __forceinline__ DualSteps dual_steps(float sigma, float alpha0, float alpha1, float eta_p,
                                                         float eta_q)
{
    return DualSteps{dual_scale(alpha1, sigma, eta_p), dual_scale(alpha0, sigma, eta_q), sigma / eta_p,
                     sigma / eta_q};
}

// Reverse of updateDualVariables (Python/vfs.cu:326-409) at a masked-in pixel.
// gpn, gqn, gtp: complete adjoints of p_t, q_t, tp_t.  State: p_{t-1}, q_{t-1} at the
// pixel; u_bar_{t-1}, v_bar_{t-1} at the pixel and its x / y targets ix, iy (E / S
// when present, else W / N; sx, sy = +1 / -1).
// Sets gp, gq := adjoints of p_{t-1}, q_{t-1} and the messages; adds to gabc and
// the alpha shares.
__device__ __forceinline__ void dual_adj_pixel(bool fx, bool fy, float uc, float2 vc, float uX, float2 vX, float uY,
                                               float2 vY, float a, float b, float c, float2 pc, float4 qc,
                                               const DualSteps& st, float2 gpn, float4 gqn, float2 gtp,
                                               float2& gp_out, float4& gq_out, DualMsg& msg, float (&gabc)[3],
                                               float& gpa0, float& gpa1)
{
    const float sx = fx ? 1.0f : -1.0f, sy = fy ? 1.0f : -1.0f;
    // the forward differences of dual_pixel, negated for a backward neighbour (b - a is
    // exactly -(a - b)): the same bits without a divergent branch
    const float dux = uX - uc, dvxx = vX.x - vc.x, dvxy = vX.y - vc.y;
    const float duy = uY - uc, dvyy = vY.y - vc.y, dvyx = vY.x - vc.x;
    float2 gU;
    float4 gV;
    gU.x = fmaf(sx, dux, -vc.x);  // (+-dux) - vc.x: the product is exact, one rounding
    gV.x = fx ? dvxx : -dvxx;
    gV.w = fx ? dvxy : -dvxy;
    gU.y = fmaf(sy, duy, -vc.y);
    gV.y = fy ? dvyy : -dvyy;
    gV.z = fy ? dvyx : -dvyx;
    const float sp = st.sp, sq = st.sq;
    const float2 tu = make_float2(fmaf(a, gU.x, c * gU.y), fmaf(c, gU.x, b * gU.y));
    const float2 up = make_float2(fmaf(sp, tu.x, pc.x), fmaf(sp, tu.y, pc.y));
    // the forward projections (dual_pixel): scale r = 1 / |x| where |x|^2 > 1
    const float s2 = sq_norm2(up.x, up.y);
    const bool projp = s2 > 1.0f;
    const float rp = unit_ball_scale(s2);
    const float2 pn = make_float2(up.x * rp, up.y * rp);
    const float4 uq = make_float4(fmaf(sq, gV.x, qc.x), fmaf(sq, gV.y, qc.y), fmaf(sq, gV.z, qc.z),
                                  fmaf(sq, gV.w, qc.w));
    const float s4 = sq_norm4(uq.x, uq.y, uq.z, uq.w);
    const bool projq = s4 > 1.0f;
    const float rq = unit_ball_scale(s4);
    const float4 qn = make_float4(uq.x * rq, uq.y * rq, uq.z * rq, uq.w * rq);
    // tp = T pn
    float ga = gtp.x * pn.x, gb = gtp.y * pn.y, gc = gtp.x * pn.y + gtp.y * pn.x;
    gpn.x += a * gtp.x + c * gtp.y;
    gpn.y += c * gtp.x + b * gtp.y;
    // projection of p
    float2 gup = gpn;
    if (projp) {
        const float dp = pn.x * gpn.x + pn.y * gpn.y;
        gup = make_float2((gpn.x - pn.x * dp) * rp, (gpn.y - pn.y * dp) * rp);
    }
    float4 guq = gqn;
    if (projq) {
        const float dq = qn.x * gqn.x + qn.y * gqn.y + qn.z * gqn.z + qn.w * gqn.w;
        guq = make_float4((gqn.x - qn.x * dq) * rq, (gqn.y - qn.y * dq) * rq, (gqn.z - qn.z * dq) * rq,
                          (gqn.w - qn.w * dq) * rq);
    }
    gp_out = gup;  // d up / d pc = I
    gq_out = guq;
    const float g_sp = gup.x * tu.x + gup.y * tu.y;
    const float g_sq = guq.x * gV.x + guq.y * gV.y + guq.z * gV.z + guq.w * gV.w;
    gpa1 += g_sp * st.dsp;
    gpa0 += g_sq * st.dsq;
    const float2 gtu = make_float2(sp * gup.x, sp * gup.y);
    ga += gtu.x * gU.x;
    gb += gtu.y * gU.y;
    gc += gtu.x * gU.y + gtu.y * gU.x;
    gabc[0] += ga;
    gabc[1] += gb;
    gabc[2] += gc;
    const float2 ggU = make_float2(a * gtu.x + c * gtu.y, c * gtu.x + b * gtu.y);
    const float4 ggV = make_float4(sq * guq.x, sq * guq.y, sq * guq.z, sq * guq.w);
    // gU.x = sx (uX - uc) - vc.x ; gU.y = sy (uY - uc) - vc.y
    // gV.x = sx (vX.x - vc.x) ; gV.w = sx (vX.y - vc.y) ; gV.y = sy (vY.y - vc.y) ; gV.z = sy (vY.x - vc.x)
    msg.um[0] = -sx * ggU.x - sy * ggU.y;
    msg.um[1] = sx * ggU.x;
    msg.um[2] = sy * ggU.y;
    msg.vm[0] = -ggU.x - sx * ggV.x - sy * ggV.z;
    msg.vm[1] = -ggU.y - sx * ggV.w - sy * ggV.y;
    msg.vm[2] = sx * ggV.x;
    msg.vm[3] = sx * ggV.w;
    msg.vm[4] = sy * ggV.z;
    msg.vm[5] = sy * ggV.y;
}

}  // namespace vfs
