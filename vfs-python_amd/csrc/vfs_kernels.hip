// Per-kernel gfx950 implementation of the VFS solver kernels (the plain,
// reference-shaped decomposition).  One thread per output pixel; 64x4 blocks so
// that each wavefront covers 64 consecutive pixels of one row (coalesced
// row access).  Each kernel cites the reference kernel whose semantics it
// implements.
#include "vfs_device.h"
#include "vfs_pd.h"
#include "vfs_median.h"

using namespace vfs;

#define VFS_PIXEL(H, W)                                       \
    const int X = blockIdx.x * blockDim.x + threadIdx.x;      \
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;      \
    if (X >= (W) || Y >= (H)) return;                         \
    const int idx = Y * (W) + X;

// ---------------------------------------------------------------------------
// pyramids
// ---------------------------------------------------------------------------

// castAndScale, Python/vfs.cu:145-156.
__global__ void __launch_bounds__(256) k_cast_and_scale(float* __restrict__ out, const uint8_t* __restrict__ in, int H, int W)
{
    VFS_PIXEL(H, W)
    out[idx] = static_cast<float>(in[idx]) / 256.0f;
}

// downSamplingImage / downSamplingVectorField with the Bilinear sampler,
// Python/vfs.cu:35-49,171-184: mean of four linear taps at +-1/4 output pixel.
__global__ void __launch_bounds__(256) k_downsample1(float* __restrict__ out, int Hd, int Wd,
                                                     const float* __restrict__ src, int Hs, int Ws, float rx, float ry)
{
    VFS_PIXEL(Hd, Wd)
    const float xc = fmaf(X + 0.5f, rx, -0.5f), yc = fmaf(Y + 0.5f, ry, -0.5f);
    const float xm = fmaf(X + 0.25f, rx, -0.5f), xp = fmaf(X + 0.75f, rx, -0.5f);
    const float ym = fmaf(Y + 0.25f, ry, -0.5f), yp = fmaf(Y + 0.75f, ry, -0.5f);
    out[idx] = 0.25f * (sample1(src, Hs, Ws, xm, yc) + sample1(src, Hs, Ws, xp, yc) +
                        sample1(src, Hs, Ws, xc, ym) + sample1(src, Hs, Ws, xc, yp));
}

__global__ void __launch_bounds__(256) k_downsample2(float2* __restrict__ out, int Hd, int Wd,
                                                     const float2* __restrict__ src, int Hs, int Ws, float rx, float ry)
{
    VFS_PIXEL(Hd, Wd)
    const float xc = fmaf(X + 0.5f, rx, -0.5f), yc = fmaf(Y + 0.5f, ry, -0.5f);
    const float xm = fmaf(X + 0.25f, rx, -0.5f), xp = fmaf(X + 0.75f, rx, -0.5f);
    const float ym = fmaf(Y + 0.25f, ry, -0.5f), yp = fmaf(Y + 0.75f, ry, -0.5f);
    const float2 a = sample2(src, Hs, Ws, xm, yc), b = sample2(src, Hs, Ws, xp, yc);
    const float2 c = sample2(src, Hs, Ws, xc, ym), d = sample2(src, Hs, Ws, xc, yp);
    out[idx] = make_float2(0.25f * (a.x + b.x + c.x + d.x), 0.25f * (a.y + b.y + c.y + d.y));
}

// downSamplingMask, Python/vfs.cu:54-59,176-179: point sample at floor((X+0.5) Ws / Wd),
// evaluated exactly in integers.
__global__ void __launch_bounds__(256) k_downsample_mask(uint8_t* __restrict__ out, int Hd, int Wd,
                                                         const uint8_t* __restrict__ src, int Hs, int Ws)
{
    VFS_PIXEL(Hd, Wd)
    const int ix = static_cast<int>(((2 * static_cast<int64_t>(X) + 1) * Ws) / (2 * static_cast<int64_t>(Wd)));
    const int iy = static_cast<int>(((2 * static_cast<int64_t>(Y) + 1) * Hs) / (2 * static_cast<int64_t>(Hd)));
    out[idx] = src[iy * Ws + ix];
}

// ---------------------------------------------------------------------------
// per-level / per-warp
// ---------------------------------------------------------------------------

// warpingImage, Python/vfs.cu:61-66,186-202.
__global__ void __launch_bounds__(256) k_warp_image(float* __restrict__ out, const float2* __restrict__ warp, int H, int W,
                                                    const float* __restrict__ src, int Hs, int Ws, float rx, float ry, int same)
{
    VFS_PIXEL(H, W)
    const float2 w = warp[idx];
    float xs, ys;
    if (same) {
        xs = X + w.x;
        ys = Y + w.y;
    } else {
        xs = fmaf(X + w.x + 0.5f, rx, -0.5f);
        ys = fmaf(Y + w.y + 0.5f, ry, -0.5f);
    }
    out[idx] = sample1(src, Hs, Ws, xs, ys);
}

// findWarpingVector, Python/vfs.cu:283-302.
__global__ void __launch_bounds__(256) k_find_warping_vector(float2* __restrict__ out, const float2* __restrict__ warp, int H, int W,
                                                             const float2* __restrict__ tvf, int Hs, int Ws, float rx, float ry, int same)
{
    VFS_PIXEL(H, W)
    const float2 w = warp[idx];
    float xs, ys;
    if (same) {
        xs = X + w.x;
        ys = Y + w.y;
    } else {
        xs = fmaf(X + w.x + 0.5f, rx, -0.5f);
        ys = fmaf(Y + w.y + 0.5f, ry, -0.5f);
    }
    out[idx] = sample2(tvf, Hs, Ws, xs, ys);
}

// gaussianMasked, Python/vfs.cu:81-102,204-226.  The reference adds the
// (.0625 * val1) term in double (a double literal), so the sum is formed in
// double and rounded once.
__global__ void __launch_bounds__(256) k_gaussian_masked(float* __restrict__ out, const float* __restrict__ img,
                                                         const uint8_t* __restrict__ mask, int H, int W)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        out[idx] = 0.0f;
        return;
    }
    const int xm = clamp_idx(X - 1, W), xp = clamp_idx(X + 1, W);
    const int ym = clamp_idx(Y - 1, H), yp = clamp_idx(Y + 1, H);
    const float v0 = img[idx];
    float v1 = img[ym * W + xm];
    v1 += img[yp * W + xm];
    v1 += img[ym * W + xp];
    v1 += img[yp * W + xp];
    float v2 = img[Y * W + xm];
    v2 += img[Y * W + xp];
    v2 += img[ym * W + X];
    v2 += img[yp * W + X];
    out[idx] = static_cast<float>(static_cast<double>(0.25f * v0) + 0.0625 * static_cast<double>(v1) +
                                  static_cast<double>(0.125f * v2));
}

// computeEta, Python/vfs.cu:228-237.
__device__ __forceinline__ float3 compute_eta(float a, float b, float c, float a11, float a004)
{
    return make_float3(fmaf(a11, (fmaf(a, a, b * b) + 2 * c * c), fmaf(a11, fmaf(a + c, a + c, fmaf((b + c), (b + c), 0.0f)), 0.0f)),
                       fmaf(a11, fmaf(b, b, c * c), a004),
                       fmaf(a11, fmaf(a, a, c * c), a004));
}

// computeTensorEtaMasked, Python/vfs.cu:104-118,239-281.
__global__ void __launch_bounds__(256) k_tensor_eta(float* __restrict__ abc, float* __restrict__ eta,
                                                    const float* __restrict__ img, const uint8_t* __restrict__ mask,
                                                    int H, int W, KParams k)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        abc[3 * idx + 0] = abc[3 * idx + 1] = abc[3 * idx + 2] = 0.0f;
        eta[3 * idx + 0] = eta[3 * idx + 1] = eta[3 * idx + 2] = 0.0f;
        return;
    }
    const int xp = clamp_idx(X + 1, W), yp = clamp_idx(Y + 1, H);
    float gx = img[yp * W + xp];
    gx -= img[yp * W + X];
    float gy = img[yp * W + xp];
    gy -= img[Y * W + xp];
    const float ng = hypotf(gx, gy);
    float nx = gx / ng, ny = gy / ng;
    if (ng < k.min_grad) {
        nx = 1.0f;
        ny = 0.0f;
    }
    const float tx = ny, ty = -nx;
    const float w = fmaxf(expf(-k.beta * powf(ng, k.gamma)), k.min_tensor);
    const float a = w * nx * nx + tx * tx;
    const float b = w * ny * ny + ty * ty;
    const float c = w * nx * ny + tx * ty;
    abc[3 * idx + 0] = a;
    abc[3 * idx + 1] = b;
    abc[3 * idx + 2] = c;
    const float3 e = compute_eta(a, b, c, k.a11, k.a004);
    eta[3 * idx + 0] = e.x;
    eta[3 * idx + 1] = e.y;
    eta[3 * idx + 2] = e.z;
}

// computeDerivatives, Python/vfs.cu:120-142,304-324.
__global__ void __launch_bounds__(256) k_derivatives(float2* __restrict__ out, const float* __restrict__ ref,
                                                     const float* __restrict__ other, const float2* __restrict__ tvf,
                                                     int H, int W)
{
    VFS_PIXEL(H, W)
    out[idx] = derivative_pixel(X, Y, idx, ref, other, tvf, H, W);
}

// ---------------------------------------------------------------------------
// primal-dual iteration
// ---------------------------------------------------------------------------

// updateDualVariables, Python/vfs.cu:326-419.
__global__ void __launch_bounds__(256) k_update_dual(float2* __restrict__ p, float2* __restrict__ tp, float4* __restrict__ q,
                                                     float sigma, const float* __restrict__ u, const float2* __restrict__ v,
                                                     const float* __restrict__ abc, const uint8_t* __restrict__ mask,
                                                     int H, int W, float alpha0, float alpha1, float eta_p, float eta_q)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        p[idx] = make_float2(0.0f, 0.0f);
        tp[idx] = make_float2(0.0f, 0.0f);
        q[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const int iE = Y * W + clamp_idx(X + 1, W), iW = Y * W + clamp_idx(X - 1, W);
    const int iS = clamp_idx(Y + 1, H) * W + X, iN = clamp_idx(Y - 1, H) * W + X;
    const bool fx = mask[iE] != 0, fy = mask[iS] != 0;  // E / S present: forward differences
    const int ix = fx ? iE : iW, iy = fy ? iS : iN;
    float2 pn, tpn;
    float4 qn;
    dual_pixel(u[idx], v[idx], u[ix], v[ix], fx, u[iy], v[iy], fy, abc[3 * idx + 0], abc[3 * idx + 1],
               abc[3 * idx + 2], p[idx], q[idx], dual_scale(alpha1, sigma, eta_p), dual_scale(alpha0, sigma, eta_q),
               pn, tpn, qn);
    p[idx] = pn;
    tp[idx] = tpn;
    q[idx] = qn;
}

// l1Thresholding, Python/vfs.cu:421-505.
__global__ void __launch_bounds__(256) k_l1_thresholding(float* out, float tau, const float2* __restrict__ tp,
                                                         const float* u, const float* __restrict__ u_bar,
                                                         const float2* __restrict__ deriv, const float* __restrict__ eta,
                                                         const uint8_t* __restrict__ mask, int H, int W, float lambda)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        out[idx] = 0.0f;
        return;
    }
    const int iW = Y * W + clamp_idx(X - 1, W), iN = clamp_idx(Y - 1, H) * W + X;
    const bool mE = mask[Y * W + clamp_idx(X + 1, W)] != 0, mW = mask[iW] != 0;
    const bool mS = mask[clamp_idx(Y + 1, H) * W + X] != 0, mN = mask[iN] != 0;
    const float2 d = deriv[idx];
    out[idx] = l1_pixel(tp[idx], tp[iW].x, tp[iN].y, mE, mW, mS, mN, d.x, recip_grad(d.x, kEps), d.y,
                        recip_eta(eta[3 * idx], kEps), u[idx], u_bar[idx], tau, lambda);
}

// updatePrimalVariables, Python/vfs.cu:507-581.
__global__ void __launch_bounds__(256) k_update_primal(float* u_out, float2* v_out, float tau, float mu,
                                                       const float2* __restrict__ p, const float4* __restrict__ q,
                                                       const float* __restrict__ u, const float* u_bar, const float2* v_bar,
                                                       const float* __restrict__ abc, const float* __restrict__ eta,
                                                       const uint8_t* __restrict__ mask, int H, int W, float alpha0, float alpha1)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        u_out[idx] = 0.0f;
        v_out[idx] = make_float2(0.0f, 0.0f);
        return;
    }
    const int iW = Y * W + clamp_idx(X - 1, W), iN = clamp_idx(Y - 1, H) * W + X;
    const bool mE = mask[Y * W + clamp_idx(X + 1, W)] != 0, mW = mask[iW] != 0;
    const bool mS = mask[clamp_idx(Y + 1, H) * W + X] != 0, mN = mask[iN] != 0;
    const float4 zero = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 qW = mW ? q[iW] : zero;
    const float4 qN = mN ? q[iN] : zero;
    float ub;
    float2 vb;
    primal_pixel(q[idx], qW, qN, mE, mW, mS, mN, abc[3 * idx + 0], abc[3 * idx + 1], abc[3 * idx + 2], p[idx],
                 recip_eta(eta[3 * idx + 1], kEps), recip_eta(eta[3 * idx + 2], kEps), v_bar[idx], u[idx], u_bar[idx],
                 tau, mu, alpha0, alpha1, ub, vb);
    u_out[idx] = ub;
    v_out[idx] = vb;
}

// medianFilter, Python/vfs.cu:583-639: exact (K*K/2)-th order statistic of the
// clamp-to-edge window by partial selection (the first K*K/2+1 positions).
template <int K>
__global__ void __launch_bounds__(256) k_median(float* __restrict__ out, const float* __restrict__ in,
                                                const uint8_t* __restrict__ mask, int H, int W)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        out[idx] = in[idx];
        return;
    }
    constexpr int N = K * K;
    float v[N];
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int yy = clamp_idx(Y + j - K / 2, H);
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int xx = clamp_idx(X + i - K / 2, W);
            v[j * K + i] = in[yy * W + xx];
        }
    }
    out[idx] = median_select<K>(v);
}

// computeOpticalFlowWithClamp, Python/vfs.cu:641-676 (in place on flow and u0).
__global__ void __launch_bounds__(256) k_flow_clamp(float2* flow, float* u0, const float* __restrict__ uf,
                                                    const float2* __restrict__ tv, const uint8_t* __restrict__ mask,
                                                    int H, int W, float limit)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) return;
    const float delta = fminf(uf[idx] - u0[idx], limit);
    u0[idx] += delta;
    const float2 t = tv[idx];
    float2 f = flow[idx];
    f.x = fmaf(delta, t.x, f.x);
    f.y = fmaf(delta, t.y, f.y);
    flow[idx] = f;
}

// upSamplingVectorField1 / upSamplingVectorField with NearestNeighborScaled,
// Python/vfs.cu:69-79,678-705.
__global__ void __launch_bounds__(256) k_upsample1(float* __restrict__ out, int Hd, int Wd, const float* __restrict__ src,
                                                   int Hs, int Ws, float rx, float ry, float scale)
{
    VFS_PIXEL(Hd, Wd)
    out[idx] = scale * sample1(src, Hs, Ws, fmaf(X + 0.5f, rx, -0.5f), fmaf(Y + 0.5f, ry, -0.5f));
}

__global__ void __launch_bounds__(256) k_upsample2(float2* __restrict__ out, int Hd, int Wd, const float2* __restrict__ src,
                                                   int Hs, int Ws, float rx, float ry, float scale)
{
    VFS_PIXEL(Hd, Wd)
    const float2 s = sample2(src, Hs, Ws, fmaf(X + 0.5f, rx, -0.5f), fmaf(Y + 0.5f, ry, -0.5f));
    out[idx] = make_float2(scale * s.x, scale * s.y);
}

// triangulateRays, Python/vfs.cu:707-745.
__global__ void __launch_bounds__(256) k_triangulate(float* __restrict__ out, const float* __restrict__ rr,
                                                     const float* __restrict__ ro, const float* __restrict__ t,
                                                     const uint8_t* __restrict__ mask, int H, int W)
{
    VFS_PIXEL(H, W)
    if (mask[idx] == 0) {
        out[3 * idx + 0] = 0.0f;
        out[3 * idx + 1] = 0.0f;
        out[3 * idx + 2] = -1.0f;
        return;
    }
    const float xr = rr[3 * idx + 0] / rr[3 * idx + 2];
    const float yr = rr[3 * idx + 1] / rr[3 * idx + 2];
    const float xo = ro[3 * idx + 0] / ro[3 * idx + 2];
    const float z = fmaf(-t[2], xo, t[0]) / fmaf(-1.0f, xr, xo);
    out[3 * idx + 0] = z * xr;
    out[3 * idx + 1] = z * yr;
    out[3 * idx + 2] = z;
}

// ---------------------------------------------------------------------------
// fills (cp.zeros / cp.copy / masked assignment of Python/vfs.py:489-497,531)
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) k_fill(float* __restrict__ out, float value, int64_t n)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[i] = value;
}

__global__ void __launch_bounds__(256) k_copy(float* __restrict__ out, const float* __restrict__ in, int64_t n)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

// The same copy 16 bytes per lane (out and in 16-byte aligned); n4 = n / 4 vectors.
__global__ void __launch_bounds__(256) k_copy4(float4* __restrict__ out, const float4* __restrict__ in, int64_t n4)
{
    for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n4;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        out[i] = in[i];
}

__global__ void __launch_bounds__(256) k_zero_masked(float* __restrict__ flow, const uint8_t* __restrict__ mask,
                                                     int H, int W, int channels)
{
    VFS_PIXEL(H, W)
    if (mask[idx] != 0) return;
    for (int c = 0; c < channels; ++c) flow[channels * idx + c] = 0.0f;
}

// ---------------------------------------------------------------------------
// C ABI launchers
// ---------------------------------------------------------------------------

static inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline unsigned blocks1d(int64_t n)
{
    int64_t b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    if (b < 1) b = 1;
    return static_cast<unsigned>(b);
}

extern "C" {

const char* vfs_version(void) { return "vfs_hip 0.1 gfx950"; }

int vfs_cast_and_scale(void* stream, float* out, const uint8_t* in, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(in);
    hipLaunchKernelGGL(k_cast_and_scale, grid2d(H, W), block2d(), 0, S(stream), out, in, H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_downsample_image(void* stream, float* out, int Hd, int Wd, const float* src, int Hs, int Ws)
{
    VFS_CHECK_DIMS(Hd, Wd);
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(src);
    hipLaunchKernelGGL(k_downsample1, grid2d(Hd, Wd), block2d(), 0, S(stream), out, Hd, Wd, src, Hs, Ws,
                       static_cast<float>(Ws) / static_cast<float>(Wd), static_cast<float>(Hs) / static_cast<float>(Hd));
    VFS_LAUNCH_STATUS();
}

int vfs_downsample_vector_field(void* stream, float* out, int Hd, int Wd, const float* src, int Hs, int Ws)
{
    VFS_CHECK_DIMS(Hd, Wd);
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(src);
    hipLaunchKernelGGL(k_downsample2, grid2d(Hd, Wd), block2d(), 0, S(stream), reinterpret_cast<float2*>(out), Hd, Wd,
                       reinterpret_cast<const float2*>(src), Hs, Ws,
                       static_cast<float>(Ws) / static_cast<float>(Wd), static_cast<float>(Hs) / static_cast<float>(Hd));
    VFS_LAUNCH_STATUS();
}

int vfs_downsample_mask(void* stream, uint8_t* out, int Hd, int Wd, const uint8_t* src, int Hs, int Ws)
{
    VFS_CHECK_DIMS(Hd, Wd);
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(src);
    hipLaunchKernelGGL(k_downsample_mask, grid2d(Hd, Wd), block2d(), 0, S(stream), out, Hd, Wd, src, Hs, Ws);
    VFS_LAUNCH_STATUS();
}

int vfs_warp_image(void* stream, float* out, const float* warp, int H, int W, const float* src, int Hs, int Ws)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(warp);
    VFS_CHECK_PTR(src);
    const int same = (H == Hs && W == Ws);
    hipLaunchKernelGGL(k_warp_image, grid2d(H, W), block2d(), 0, S(stream), out, reinterpret_cast<const float2*>(warp), H, W,
                       src, Hs, Ws, static_cast<float>(Ws) / W, static_cast<float>(Hs) / H, same);
    VFS_LAUNCH_STATUS();
}

int vfs_gaussian_masked(void* stream, float* out, const float* src, const uint8_t* mask, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(src);
    VFS_CHECK_PTR(mask);
    hipLaunchKernelGGL(k_gaussian_masked, grid2d(H, W), block2d(), 0, S(stream), out, src, mask, H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_compute_tensor_eta(void* stream, float* abc, float* eta, const float* img, const uint8_t* mask, int H, int W,
                           const vfs_params* prm)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(abc);
    VFS_CHECK_PTR(eta);
    VFS_CHECK_PTR(img);
    VFS_CHECK_PTR(mask);
    VFS_CHECK_PTR(prm);
    hipLaunchKernelGGL(k_tensor_eta, grid2d(H, W), block2d(), 0, S(stream), abc, eta, img, mask, H, W, to_kparams(prm));
    VFS_LAUNCH_STATUS();
}

int vfs_find_warping_vector(void* stream, float* out, const float* warp, int H, int W, const float* tvf, int Hs, int Ws)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(warp);
    VFS_CHECK_PTR(tvf);
    const int same = (H == Hs && W == Ws);
    hipLaunchKernelGGL(k_find_warping_vector, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(out),
                       reinterpret_cast<const float2*>(warp), H, W, reinterpret_cast<const float2*>(tvf), Hs, Ws,
                       static_cast<float>(Ws) / W, static_cast<float>(Hs) / H, same);
    VFS_LAUNCH_STATUS();
}

int vfs_compute_derivatives(void* stream, float* out, const float* ref, const float* other, const float* tvf, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(ref);
    VFS_CHECK_PTR(other);
    VFS_CHECK_PTR(tvf);
    hipLaunchKernelGGL(k_derivatives, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(out), ref, other,
                       reinterpret_cast<const float2*>(tvf), H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_update_dual(void* stream, float* p, float* tp, float* q, float sigma, const float* u_bar, const float* v_bar,
                    const float* abc, const uint8_t* mask, int H, int W, const vfs_params* prm)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(p);
    VFS_CHECK_PTR(tp);
    VFS_CHECK_PTR(q);
    VFS_CHECK_PTR(u_bar);
    VFS_CHECK_PTR(v_bar);
    VFS_CHECK_PTR(abc);
    VFS_CHECK_PTR(mask);
    VFS_CHECK_PTR(prm);
    hipLaunchKernelGGL(k_update_dual, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(p),
                       reinterpret_cast<float2*>(tp), reinterpret_cast<float4*>(q), sigma, u_bar,
                       reinterpret_cast<const float2*>(v_bar), abc, mask, H, W, prm->alpha0, prm->alpha1, prm->eta_p,
                       prm->eta_q);
    VFS_LAUNCH_STATUS();
}

int vfs_l1_thresholding(void* stream, float* out, float tau, const float* tp, const float* u, const float* u_bar,
                        const float* deriv, const float* eta, const uint8_t* mask, int H, int W, const vfs_params* prm)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(tp);
    VFS_CHECK_PTR(u);
    VFS_CHECK_PTR(u_bar);
    VFS_CHECK_PTR(deriv);
    VFS_CHECK_PTR(eta);
    VFS_CHECK_PTR(mask);
    VFS_CHECK_PTR(prm);
    hipLaunchKernelGGL(k_l1_thresholding, grid2d(H, W), block2d(), 0, S(stream), out, tau,
                       reinterpret_cast<const float2*>(tp), u, u_bar, reinterpret_cast<const float2*>(deriv), eta, mask,
                       H, W, prm->lambda);
    VFS_LAUNCH_STATUS();
}

int vfs_update_primal(void* stream, float* u_bar_out, float* v_bar_out, float tau, float mu, const float* p,
                      const float* q, const float* u, const float* u_bar, const float* v_bar, const float* abc,
                      const float* eta, const uint8_t* mask, int H, int W, const vfs_params* prm)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(u_bar_out);
    VFS_CHECK_PTR(v_bar_out);
    VFS_CHECK_PTR(p);
    VFS_CHECK_PTR(q);
    VFS_CHECK_PTR(u);
    VFS_CHECK_PTR(u_bar);
    VFS_CHECK_PTR(v_bar);
    VFS_CHECK_PTR(abc);
    VFS_CHECK_PTR(eta);
    VFS_CHECK_PTR(mask);
    VFS_CHECK_PTR(prm);
    hipLaunchKernelGGL(k_update_primal, grid2d(H, W), block2d(), 0, S(stream), u_bar_out,
                       reinterpret_cast<float2*>(v_bar_out), tau, mu, reinterpret_cast<const float2*>(p),
                       reinterpret_cast<const float4*>(q), u, u_bar, reinterpret_cast<const float2*>(v_bar), abc, eta,
                       mask, H, W, prm->alpha0, prm->alpha1);
    VFS_LAUNCH_STATUS();
}

int vfs_median_filter(void* stream, float* out, const float* in, const uint8_t* mask, int H, int W, int ksize)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(in);
    VFS_CHECK_PTR(mask);
    if (out == in) return VFS_EINVAL;
    if (ksize == 5)
        hipLaunchKernelGGL(k_median<5>, grid2d(H, W), block2d(), 0, S(stream), out, in, mask, H, W);
    else if (ksize == 3)
        hipLaunchKernelGGL(k_median<3>, grid2d(H, W), block2d(), 0, S(stream), out, in, mask, H, W);
    else
        return VFS_EINVAL;
    VFS_LAUNCH_STATUS();
}

int vfs_optical_flow_with_clamp(void* stream, float* flow, float* u0, const float* uf, const float* tv,
                                const uint8_t* mask, int H, int W, float limit_range)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(flow);
    VFS_CHECK_PTR(u0);
    VFS_CHECK_PTR(uf);
    VFS_CHECK_PTR(tv);
    VFS_CHECK_PTR(mask);
    hipLaunchKernelGGL(k_flow_clamp, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(flow), u0, uf,
                       reinterpret_cast<const float2*>(tv), mask, H, W, limit_range);
    VFS_LAUNCH_STATUS();
}

int vfs_upsample(void* stream, float* out, int Hd, int Wd, const float* src, int Hs, int Ws, int channels, float f_scale)
{
    VFS_CHECK_DIMS(Hd, Wd);
    VFS_CHECK_DIMS(Hs, Ws);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(src);
    const float rx = static_cast<float>(Ws) / static_cast<float>(Wd);
    const float ry = static_cast<float>(Hs) / static_cast<float>(Hd);
    if (channels == 1)
        hipLaunchKernelGGL(k_upsample1, grid2d(Hd, Wd), block2d(), 0, S(stream), out, Hd, Wd, src, Hs, Ws, rx, ry, f_scale);
    else if (channels == 2)
        hipLaunchKernelGGL(k_upsample2, grid2d(Hd, Wd), block2d(), 0, S(stream), reinterpret_cast<float2*>(out), Hd, Wd,
                           reinterpret_cast<const float2*>(src), Hs, Ws, rx, ry, f_scale);
    else
        return VFS_EINVAL;
    VFS_LAUNCH_STATUS();
}

int vfs_triangulate_rays(void* stream, float* out, const float* rays_ref, const float* rays_other, const float* t,
                         const uint8_t* mask, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(rays_ref);
    VFS_CHECK_PTR(rays_other);
    VFS_CHECK_PTR(t);
    VFS_CHECK_PTR(mask);
    hipLaunchKernelGGL(k_triangulate, grid2d(H, W), block2d(), 0, S(stream), out, rays_ref, rays_other, t, mask, H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_fill(void* stream, float* out, float value, int64_t n)
{
    if (n < 0) return VFS_EINVAL;
    if (n == 0) return 0;
    VFS_CHECK_PTR(out);
    hipLaunchKernelGGL(k_fill, dim3(blocks1d(n)), dim3(256), 0, S(stream), out, value, n);
    VFS_LAUNCH_STATUS();
}

int vfs_copy(void* stream, float* out, const float* in, int64_t n)
{
    if (n < 0) return VFS_EINVAL;
    if (n == 0) return 0;
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(in);
    const bool aligned = ((reinterpret_cast<uintptr_t>(out) | reinterpret_cast<uintptr_t>(in)) & 15u) == 0;
    const int64_t n4 = aligned ? n / 4 : 0;
    if (n4 > 0)
        hipLaunchKernelGGL(k_copy4, dim3(blocks1d(n4)), dim3(256), 0, S(stream), reinterpret_cast<float4*>(out),
                           reinterpret_cast<const float4*>(in), n4);
    if (n - 4 * n4 > 0)
        hipLaunchKernelGGL(k_copy, dim3(blocks1d(n - 4 * n4)), dim3(256), 0, S(stream), out + 4 * n4, in + 4 * n4,
                           n - 4 * n4);
    VFS_LAUNCH_STATUS();
}

int vfs_zero_masked(void* stream, float* flow, const uint8_t* mask, int H, int W, int channels)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(flow);
    VFS_CHECK_PTR(mask);
    if (channels < 1 || channels > 4) return VFS_EINVAL;
    hipLaunchKernelGGL(k_zero_masked, grid2d(H, W), block2d(), 0, S(stream), flow, mask, H, W, channels);
    VFS_LAUNCH_STATUS();
}

}  // extern "C"
