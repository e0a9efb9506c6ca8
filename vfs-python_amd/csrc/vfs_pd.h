// Per-pixel arithmetic of one TGV-L1 primal-dual iteration, shared by the
// per-kernel path (vfs_kernels.hip) and the temporally blocked path
// (vfs_pd_block.hip) so that both evaluate the same operation sequence.
//
// Dual step: Python/vfs.cu:326-409.  Threshold: Python/vfs.cu:421-495.
// Primal step: Python/vfs.cu:507-570.
#pragma once

#include <hip/hip_runtime.h>

namespace vfs {

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
// The projection divides by max(norm, 1); division by 1 is exact, so it is
// skipped when the norm does not exceed 1.
__device__ __forceinline__ void dual_pixel(float uc, float2 vc, float uX, float2 vX, bool fwd_x, float uY, float2 vY,
                                           bool fwd_y, float a, float b, float c, float2 pc, float4 qc, float scale_p,
                                           float scale_q, float2& pn, float2& tpn, float4& qn)
{
    float2 gU;
    float4 gV;
    if (fwd_x) {
        gU.x = uX - uc - vc.x;
        gV.x = vX.x - vc.x;
        gV.w = vX.y - vc.y;
    } else {
        gU.x = uc - uX - vc.x;
        gV.x = vc.x - vX.x;
        gV.w = vc.y - vX.y;
    }
    if (fwd_y) {
        gU.y = uY - uc - vc.y;
        gV.y = vY.y - vc.y;
        gV.z = vY.x - vc.x;
    } else {
        gU.y = uc - uY - vc.y;
        gV.y = vc.y - vY.y;
        gV.z = vc.x - vY.x;
    }
    const float2 tu = make_float2(fmaf(a, gU.x, c * gU.y), fmaf(c, gU.x, b * gU.y));
    const float2 up = make_float2(fmaf(scale_p, tu.x, pc.x), fmaf(scale_p, tu.y, pc.y));
    float nrm = fmaxf(hypotf(up.x, up.y), 1.0f);
    pn = up;
    if (nrm > 1.0f) {
        pn.x = up.x / nrm;
        pn.y = up.y / nrm;
    }
    tpn = make_float2(fmaf(a, pn.x, c * pn.y), fmaf(c, pn.x, b * pn.y));
    const float4 uq = make_float4(fmaf(scale_q, gV.x, qc.x), fmaf(scale_q, gV.y, qc.y), fmaf(scale_q, gV.z, qc.z),
                                  fmaf(scale_q, gV.w, qc.w));
    nrm = fmaxf(norm4df(uq.x, uq.y, uq.z, uq.w), 1.0f);
    qn = uq;
    if (nrm > 1.0f) {
        qn.x = uq.x / nrm;
        qn.y = uq.y / nrm;
        qn.z = uq.z / nrm;
        qn.w = uq.w / nrm;
    }
}

// Linearised-data-term threshold: the prox of lambda |Iu (x - u) + Iz| in the
// metric eta.x / tau, after the divergence step of T p.  Returns the new u.
__device__ __forceinline__ float l1_pixel(float2 tc, float tpWx, float tpNy, bool mE, bool mW, bool mS, bool mN,
                                          float dIu, float dIz, float e0, float uc, float ubc, float tau,
                                          float lambda, float eps)
{
    float gx, gy;
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
    const float tauEtaU = (fabsf(e0) < eps) ? tau : tau / e0;
    const float uTauGtp = ubc + tauEtaU * (gx + gy);
    const float rho = dIu * (uTauGtp - uc) + dIz;
    const float bound = lambda * tauEtaU * dIu * dIu;
    float du = 0.0f;
    if ((rho <= bound) && (rho >= -bound))
        du = (fabsf(dIu) < eps) ? (uTauGtp - uc) : ((uTauGtp - uc) - rho / dIu);
    else if (rho < -bound)
        du = (uTauGtp - uc) + lambda * tauEtaU * dIu;
    else if (rho > bound)
        du = (uTauGtp - uc) - lambda * tauEtaU * dIu;
    return uc + du;
}

// Primal descent on v and over-relaxation of u and v.  qW / qN are read only
// when the corresponding neighbour is inside the mask.
__device__ __forceinline__ void primal_pixel(float4 qc, float4 qW, float4 qN, bool mE, bool mW, bool mS, bool mN,
                                             float a, float b, float c, float2 pc, float e1, float e2, float2 vb,
                                             float uc, float ub, float tau, float mu, float alpha0, float alpha1,
                                             float eps, float& ub_out, float2& vb_out)
{
    float4 gQ = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (mE) {
        gQ.x = qc.x;
        gQ.w = qc.w;
    }
    if (mW) {
        gQ.x -= qW.x;
        gQ.w -= qW.w;
    }
    if (mS) {
        gQ.y = qc.y;
        gQ.z = qc.z;
    }
    if (mN) {
        gQ.y -= qN.y;
        gQ.z -= qN.z;
    }
    const float2 tu = make_float2(fmaf(a, pc.x, c * pc.y), fmaf(c, pc.x, b * pc.y));
    const float tauEta1 = (fabsf(e1) < eps) ? tau : (tau / e1);
    const float tauEta2 = (fabsf(e2) < eps) ? tau : (tau / e2);
    const float vx = fmaf(tauEta1, fmaf(alpha1, tu.x, alpha0 * (gQ.x + gQ.z)), vb.x);
    const float vy = fmaf(tauEta2, fmaf(alpha1, tu.y, alpha0 * (gQ.w + gQ.y)), vb.y);
    ub_out = fmaf(mu, uc - ub, uc);
    vb_out = make_float2(fmaf(mu, vx - vb.x, vx), fmaf(mu, vy - vb.y, vy));
}

}  // namespace vfs
