// Device-side helpers shared by the VFS kernels (gfx950).
//
// Sampling semantics (stand-in for the reference's CUDA texture objects,
// Python/texture.py:27-55 and the descriptors at Python/vfs.py:96-98,169-171):
//   * texel coordinates with pixel centres on integers;
//   * bilinear interpolation with exact fp32 weights;
//   * half-sample symmetric mirror addressing with period 2N.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vfs_hip.h"

#define VFS_BLOCK_X 64
#define VFS_BLOCK_Y 4

namespace vfs {

// Half-sample symmetric reflection (cudaAddressModeMirror): -1 -> 0, n -> n-1.
__device__ __forceinline__ int mirror(int i, int n)
{
    if (static_cast<unsigned>(i) < static_cast<unsigned>(n)) return i;
    const int p = 2 * n;
    int m = i % p;
    if (m < 0) m += p;
    return m >= n ? p - 1 - m : m;
}

// Neighbour index for +-1 offsets (mirror of -1 is 0, of n is n-1).
__device__ __forceinline__ int clamp_idx(int i, int n)
{
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

struct Tap {
    int i00, i10, i01, i11;  // element offsets (pixel index, not scaled by channels)
    float ax, ay;
};

// The same tap set as columns x0, x1 and rows y0, y1 of the image.
struct TapXY {
    int x0, x1, y0, y1;
    float ax, ay;
};

// Bilinear tap set at texel coordinate (x, y).  Non-finite coordinates keep
// their NaN/inf in the weights so that the sample is non-finite, while the
// indices stay in range.
__device__ __forceinline__ TapXY make_tap_xy(float x, float y, int H, int W)
{
    const float xc = fminf(fmaxf(x, -1.0e7f), 1.0e7f);
    const float yc = fminf(fmaxf(y, -1.0e7f), 1.0e7f);
    const float fx = floorf(xc);
    const float fy = floorf(yc);
    const int ix = static_cast<int>(fx);
    const int iy = static_cast<int>(fy);
    TapXY t;
    t.x0 = mirror(ix, W);
    t.x1 = mirror(ix + 1, W);
    t.y0 = mirror(iy, H);
    t.y1 = mirror(iy + 1, H);
    t.ax = x - fx;
    t.ay = y - fy;
    return t;
}

__device__ __forceinline__ Tap make_tap(float x, float y, int H, int W)
{
    const TapXY u = make_tap_xy(x, y, H, W);
    Tap t;
    t.i00 = u.y0 * W + u.x0;
    t.i10 = u.y0 * W + u.x1;
    t.i01 = u.y1 * W + u.x0;
    t.i11 = u.y1 * W + u.x1;
    t.ax = u.ax;
    t.ay = u.ay;
    return t;
}

__device__ __forceinline__ float lerp2(float t00, float t10, float t01, float t11, float ax, float ay)
{
    const float r0 = fmaf(ax, t10 - t00, t00);
    const float r1 = fmaf(ax, t11 - t01, t01);
    return fmaf(ay, r1 - r0, r0);
}

__device__ __forceinline__ float sample1(const float* __restrict__ img, int H, int W, float x, float y)
{
    const Tap t = make_tap(x, y, H, W);
    return lerp2(img[t.i00], img[t.i10], img[t.i01], img[t.i11], t.ax, t.ay);
}

__device__ __forceinline__ float2 sample2(const float2* __restrict__ img, int H, int W, float x, float y)
{
    const Tap t = make_tap(x, y, H, W);
    const float2 a = img[t.i00], b = img[t.i10], c = img[t.i01], d = img[t.i11];
    return make_float2(lerp2(a.x, b.x, c.x, d.x, t.ax, t.ay), lerp2(a.y, b.y, c.y, d.y, t.ax, t.ay));
}

// computeDerivatives at one pixel, Python/vfs.cu:120-142: Iu = 5-tap directional
// derivative of ref and other along the unit translation direction, Iz = other - ref.
__device__ __forceinline__ float2 derivative_pixel(int X, int Y, int idx, const float* __restrict__ ref,
                                                   const float* __restrict__ other, const float2* __restrict__ tvf,
                                                   int H, int W)
{
    const float2 vec = tvf[idx];
    const float nrm = hypotf(vec.x, vec.y);
    const float ex = vec.x / nrm, ey = vec.y / nrm;
    const float x = static_cast<float>(X), y = static_cast<float>(Y);
    const float xm2 = fmaf(-2.0f, ex, x), ym2 = fmaf(-2.0f, ey, y);
    const float xm1 = x - ex, ym1 = y - ey;
    const float xp1 = x + ex, yp1 = y + ey;
    const float xp2 = fmaf(2.0f, ex, x), yp2 = fmaf(2.0f, ey, y);
    float v0 = sample1(ref, H, W, xm2, ym2);
    v0 -= sample1(ref, H, W, xm1, ym1) * 8;
    v0 += sample1(ref, H, W, xp1, yp1) * 8;
    v0 -= sample1(ref, H, W, xp2, yp2);
    float v1 = sample1(other, H, W, xm2, ym2);
    v1 -= sample1(other, H, W, xm1, ym1) * 8;
    v1 += sample1(other, H, W, xp1, yp1) * 8;
    v1 -= sample1(other, H, W, xp2, yp2);
    const float v2 = ref[idx];
    const float v3 = other[idx];
    return make_float2((v0 + v1) / 24, v3 - v2);
}

// Kernel-side copy of the solver constants.
struct KParams {
    float beta, gamma, alpha0, alpha1, eta_p, eta_q, lambda, limit_range;
    float f_scale, min_grad, min_tensor, a11, a004;
};

inline KParams to_kparams(const vfs_params* p)
{
    KParams k;
    k.beta = p->beta;
    k.gamma = p->gamma;
    k.alpha0 = p->alpha0;
    k.alpha1 = p->alpha1;
    k.eta_p = p->eta_p;
    k.eta_q = p->eta_q;
    k.lambda = p->lambda;
    k.limit_range = p->limit_range;
    k.f_scale = p->f_scale;
    k.min_grad = p->min_grad;
    k.min_tensor = p->min_tensor;
    k.a11 = p->a11;
    k.a004 = p->a004;
    return k;
}

constexpr float kEps = 1e-7f;  // Python/vfs.cu:430,518

inline dim3 grid2d(int H, int W)
{
    return dim3((W + VFS_BLOCK_X - 1) / VFS_BLOCK_X, (H + VFS_BLOCK_Y - 1) / VFS_BLOCK_Y, 1);
}

inline dim3 block2d() { return dim3(VFS_BLOCK_X, VFS_BLOCK_Y, 1); }

// Work item of workgroup b among n.  Workgroups are dispatched to the eight XCDs in
// turn (b to XCD b mod 8), each XCD with its own L2; items are handed out so that
// each XCD works through one contiguous run of items, so tiles that share halo
// rows and cache lines (row-major neighbours) meet in the same L2.
__device__ __forceinline__ int xcd_item(int b, int n)
{
    const int x = b & 7, k = b >> 3;
    int start = 0;
    for (int y = 0; y < x; ++y) start += (n - y + 7) >> 3;
    return start + k;
}

}  // namespace vfs

// Status helpers for the C ABI.
#define VFS_CHECK_DIMS(H, W) \
    do { if ((H) <= 0 || (W) <= 0 || (int64_t)(H) * (int64_t)(W) > (int64_t)1 << 30) return VFS_EINVAL; } while (0)
#define VFS_CHECK_PTR(p) \
    do { if ((p) == nullptr) return VFS_EINVAL; } while (0)
#define VFS_LAUNCH_STATUS() \
    do { hipError_t e_ = hipGetLastError(); return e_ == hipSuccess ? 0 : static_cast<int>(e_); } while (0)
