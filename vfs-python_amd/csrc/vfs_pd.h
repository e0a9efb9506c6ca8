// Per-pixel arithmetic of one TGV-L1 primal-dual iteration, shared by the
// per-kernel path (vfs_kernels.hip), the temporally blocked path
// (vfs_pd_block.hip) and the adjoint's recomputation (vfs_adj_pixel.h), so that
// all evaluate the same operation sequence.
//
// Dual step: Python/vfs.cu:326-409.  Threshold: Python/vfs.cu:421-495.
// Primal step: Python/vfs.cu:507-570.
//
// Arithmetic form.  The reference divides by the per-pixel metric eta and by the
// image derivative Iu at every iteration, and projects with hypotf / norm4df and a
// division per component.  Here the divisions by per-pixel constants are products
// with their reciprocals (recip_eta, recip_grad; the hardware reciprocal), formed
// once where the constants are loaded, and the projection multiplies by rsqrt of the
// squared norm (unit_ball_scale).  Each differs from the quotient by at most a couple of ulp;
// the iteration is held to the float64 oracle with the same tolerances.
#pragma once

#include <hip/hip_runtime.h>

namespace vfs {

// 1 / eta for the metric steps tau / eta (Python/vfs.cu:430-431,518-521): the step
// is tau * recip_eta(eta), and tau itself where |eta| < eps.  v_rcp_f32 (1 ulp):
// eta >= eps keeps both the input and the result normal.
__device__ __forceinline__ float recip_eta(float e, float eps)
{
    return (fabsf(e) < eps) ? 1.0f : __builtin_amdgcn_rcpf(e);
}

// 1 / Iu for the threshold's inner branch (Python/vfs.cu:476-480), 0 where
// |Iu| < eps so that the branch reduces to the unthresholded step.
__device__ __forceinline__ float recip_grad(float d, float eps)
{
    return (fabsf(d) < eps) ? 0.0f : __builtin_amdgcn_rcpf(d);
}

// Scale of the projection onto the unit ball given the squared norm s:
// 1 / sqrt(s) where s > 1, else 1 (x * 1 is x exactly).  Formed as rsq(max(s, 1)):
// v_rsq_f32 returns exactly 1 at 1, and fmaxf returns 1 for a NaN s, so this is
// the select s > 1 ? rsq(s) : 1 bit for bit, in one instruction fewer.
__device__ __forceinline__ float unit_ball_scale(float s)
{
    return __builtin_amdgcn_rsqf(fmaxf(s, 1.0f));
}

__device__ __forceinline__ float sq_norm2(float x, float y)
{
    return fmaf(x, x, y * y);
}

__device__ __forceinline__ float sq_norm4(float x, float y, float z, float w)
{
    return fmaf(x, x, fmaf(y, y, fmaf(z, z, w * w)));
}

// The two step sizes of the dual ascent, alpha1 sigma / eta_p and alpha0 sigma /
// eta_q (Python/vfs.cu:387,397), as the reference forms them in fp32.
__host__ __device__ __forceinline__ float dual_scale(float alpha, float sigma, float eta)
{
    return alpha * sigma / eta;
}

// Dual ascent with pointwise projection.  The caller selects the neighbour used
// for each axis: (uX, vX) is the east neighbour when fwd_x (forward difference)
// and the west neighbour otherwise; likewise (uY, vY) south / north.
// scale_p / scale_q are dual_scale(alpha1, sigma, eta_p) / dual_scale(alpha0, sigma, eta_q).
// The projection scales by 1 / max(norm, 1) (unit_ball_scale).
__device__ __forceinline__ void dual_pixel(float uc, float2 vc, float uX, float2 vX, bool fwd_x, float uY, float2 vY,
                                           bool fwd_y, float a, float b, float c, float2 pc, float4 qc, float scale_p,
                                           float scale_q, float2& pn, float2& tpn, float4& qn)
{
    // Forward differences, negated for a backward neighbour: b - a is exactly -(a - b).
    // The sign s = +-1 goes into an fma, s * d - v, whose product is exact, so it rounds
    // once, as the difference (+-d) - v does; and into the step, (s * scale_q) d + q,
    // which is scale_q (s d) + q exactly.  Same bits as the differences written out,
    // without a divergent branch and without a select per component.
    const float dux = uX - uc, dvxx = vX.x - vc.x, dvxy = vX.y - vc.y;
    const float duy = uY - uc, dvyy = vY.y - vc.y, dvyx = vY.x - vc.x;
    const float sx = fwd_x ? 1.0f : -1.0f, sy = fwd_y ? 1.0f : -1.0f;
    float2 gU;
    gU.x = fmaf(sx, dux, -vc.x);
    gU.y = fmaf(sy, duy, -vc.y);
    const float sqx = sx * scale_q, sqy = sy * scale_q;  // exact: sx, sy are +-1
    const float2 tu = make_float2(fmaf(a, gU.x, c * gU.y), fmaf(c, gU.x, b * gU.y));
    const float2 up = make_float2(fmaf(scale_p, tu.x, pc.x), fmaf(scale_p, tu.y, pc.y));
    const float rp = unit_ball_scale(sq_norm2(up.x, up.y));
    pn = make_float2(up.x * rp, up.y * rp);
    tpn = make_float2(fmaf(a, pn.x, c * pn.y), fmaf(c, pn.x, b * pn.y));
    const float4 uq = make_float4(fmaf(sqx, dvxx, qc.x), fmaf(sqy, dvyy, qc.y), fmaf(sqy, dvyx, qc.z),
                                  fmaf(sqx, dvxy, qc.w));
    const float rq = unit_ball_scale(sq_norm4(uq.x, uq.y, uq.z, uq.w));
    qn = make_float4(uq.x * rq, uq.y * rq, uq.z * rq, uq.w * rq);
}

// Linearised-data-term threshold: the prox of lambda |Iu (x - u) + Iz| in the
// metric eta.x / tau, after the divergence step of T p.  Returns the new u.
// r0 = recip_eta(eta.x), rIu = recip_grad(Iu).
// ZeroOut: the caller guarantees tpWx (tpNy) is +0 wherever mW (mN) is false, as it
// is where every state outside the mask is zero (the blocked kernel); then
// tc.x - tpWx is tc.x exactly and the mW / mN tests drop out.
template <bool ZeroOut = false>
__device__ __forceinline__ float l1_pixel(float2 tc, float tpWx, float tpNy, bool mE, bool mW, bool mS, bool mN,
                                          float dIu, float rIu, float dIz, float r0, float uc, float ubc, float tau,
                                          float lambda)
{
    float gx, gy;
    if (ZeroOut) {
        gx = mE ? tc.x - tpWx : -tpWx;
        gy = mS ? tc.y - tpNy : -tpNy;
    } else {
        if (mE && mW)
            gx = tc.x - tpWx;
        else if (!mE)
            gx = -tpWx;
        else
            gx = tc.x;
        if (mS && mN)
            gy = tc.y - tpNy;
        else if (!mS)
            gy = -tpNy;
        else
            gy = tc.y;
    }
    const float tauEtaU = tau * r0;
    const float uTauGtp = ubc + tauEtaU * (gx + gy);
    const float rho = dIu * (uTauGtp - uc) + dIz;
    const float bound = lambda * tauEtaU * dIu * dIu;
    // the three branches of the prox, each the same expression as written per branch,
    // selected without a divergent branch (none of them for a NaN rho: du = 0)
    const float du_in = (uTauGtp - uc) - rho * rIu;
    const float du_lo = (uTauGtp - uc) + lambda * tauEtaU * dIu;
    const float du_hi = (uTauGtp - uc) - lambda * tauEtaU * dIu;
    const float du = ((rho <= bound) && (rho >= -bound)) ? du_in
                     : (rho < -bound)                    ? du_lo
                     : (rho > bound)                     ? du_hi
                                                         : 0.0f;
    return uc + du;
}

// Primal descent on v and over-relaxation of u and v.  qW / qN are read only
// when the corresponding neighbour is inside the mask.  r1, r2 = recip_eta(eta.y),
// recip_eta(eta.z).
// ZeroOut: the caller guarantees qW (qN) is +0 wherever mW (mN) is false (every
// state outside the mask is zero, as in the blocked kernel); x - (+0) is x exactly,
// so qW and qN are subtracted without the tests, with the same bits.
template <bool ZeroOut = false>
__device__ __forceinline__ void primal_pixel(float4 qc, float4 qW, float4 qN, bool mE, bool mW, bool mS, bool mN,
                                             float a, float b, float c, float2 pc, float r1, float r2, float2 vb,
                                             float uc, float ub, float tau, float mu, float alpha0, float alpha1,
                                             float& ub_out, float2& vb_out)
{
    float4 gQ = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (mE) {
        gQ.x = qc.x;
        gQ.w = qc.w;
    }
    if (ZeroOut || mW) {
        gQ.x -= qW.x;
        gQ.w -= qW.w;
    }
    if (mS) {
        gQ.y = qc.y;
        gQ.z = qc.z;
    }
    if (ZeroOut || mN) {
        gQ.y -= qN.y;
        gQ.z -= qN.z;
    }
    const float2 tu = make_float2(fmaf(a, pc.x, c * pc.y), fmaf(c, pc.x, b * pc.y));
    const float tauEta1 = tau * r1;
    const float tauEta2 = tau * r2;
    const float vx = fmaf(tauEta1, fmaf(alpha1, tu.x, alpha0 * (gQ.x + gQ.z)), vb.x);
    const float vy = fmaf(tauEta2, fmaf(alpha1, tu.y, alpha0 * (gQ.w + gQ.y)), vb.y);
    ub_out = fmaf(mu, uc - ub, uc);
    vb_out = make_float2(fmaf(mu, vx - vb.x, vx), fmaf(mu, vy - vb.y, vy));
}

}  // namespace vfs
