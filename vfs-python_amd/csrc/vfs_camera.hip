// Equidistant fisheye camera kernels (r = f * theta), gfx950.
// Model: Python/equidistant_cam.cuh:32-59; kernels: Python/template_cam.cu:25-86.
// The intrinsics (f, u0, v0) are kernel arguments instead of __constant__ memory
// (Python/util_cuda.py:27-30 uploads them there in the reference).
#include "vfs_device.h"

using namespace vfs;

namespace {

// EquidistantCamera::project, Python/equidistant_cam.cuh:32-43.
__device__ __forceinline__ float2 cam_project(float X, float Y, float Z, float f, float u0, float v0)
{
    const float R = hypotf(X, Y);
    const float theta = atan2f(R, Z);
    const float phi = atan2f(Y, X);
    float s, c;
    sincosf(phi, &s, &c);
    return make_float2(fmaf(f * theta, c, u0), fmaf(f * theta, s, v0));
}

// EquidistantCamera::unproject, Python/equidistant_cam.cuh:45-59.
__device__ __forceinline__ float3 cam_unproject(float x, float y, float f, float u0, float v0)
{
    const float dx = x - u0, dy = y - v0;
    const float theta = hypotf(dx, dy) / f;
    const float phi = atan2f(dy, dx);
    float sp, cp, st, ct;
    sincosf(phi, &sp, &cp);
    sincosf(theta, &st, &ct);
    return make_float3(cp * st, sp * st, ct);
}

__global__ void __launch_bounds__(256) k_get_rays(float* __restrict__ out, int H, int W, float f, float u0, float v0)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    const float3 r = cam_unproject(static_cast<float>(X), static_cast<float>(Y), f, u0, v0);
    out[3 * idx + 0] = r.x;
    out[3 * idx + 1] = r.y;
    out[3 * idx + 2] = r.z;
}

__global__ void __launch_bounds__(256) k_uv_from_xyz(float2* __restrict__ uv, const float* __restrict__ xyz, int H, int W,
                                                     float f, float u0, float v0)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    uv[idx] = cam_project(xyz[3 * idx + 0], xyz[3 * idx + 1], xyz[3 * idx + 2], f, u0, v0);
}

__global__ void __launch_bounds__(256) k_xyz_from_uv(float* __restrict__ xyz, const float2* __restrict__ uv, int H, int W,
                                                     float f, float u0, float v0)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    const float2 p = uv[idx];
    const float3 r = cam_unproject(p.x, p.y, f, u0, v0);
    xyz[3 * idx + 0] = r.x;
    xyz[3 * idx + 1] = r.y;
    xyz[3 * idx + 2] = r.z;
}

// getThetas, Python/template_cam.cu:39-54.
__global__ void __launch_bounds__(256) k_thetas(float* __restrict__ out, const float* __restrict__ rays, int H, int W)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    const float x = rays[3 * idx + 0], y = rays[3 * idx + 1], z = rays[3 * idx + 2];
    out[3 * idx + 0] = atan2f(x, z);
    out[3 * idx + 1] = atan2f(y, z);
    out[3 * idx + 2] = atan2f(hypotf(x, y), z);
}

// getXYZFromDistance, Python/template_cam.cu:88-112 (the mask texture there is
// Border + Point with unnormalised coordinates, i.e. an exact read of mask[Y, X]).
__global__ void __launch_bounds__(256) k_xyz_from_distance(float* __restrict__ out, const float* __restrict__ rays,
                                                           const float* __restrict__ dist, const uint8_t* __restrict__ mask,
                                                           int H, int W)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    if (mask[idx] == 0) {
        out[3 * idx + 0] = 0.0f;
        out[3 * idx + 1] = 0.0f;
        out[3 * idx + 2] = -1.0f;
        return;
    }
    const float d = dist[idx];
    out[3 * idx + 0] = d * rays[3 * idx + 0];
    out[3 * idx + 1] = d * rays[3 * idx + 1];
    out[3 * idx + 2] = d * rays[3 * idx + 2];
}

struct Mat3 {
    float m[9];
};

// Python/vector_field.py:34-54 after the normaliser, fused per pixel:
// tvf = unit((project(xyz + s0 t) - project(xyz - s1 t)) / 2), cvf = project_other(R xyz) - uv.
__global__ void __launch_bounds__(256) k_vector_fields(float2* __restrict__ tvf, float2* __restrict__ cvf,
                                                       const float* __restrict__ xyz, const float2* __restrict__ uv,
                                                       int H, int W, float fr, float ur, float vr, float fo, float uo,
                                                       float vo, Mat3 R, float3 st0, float3 st1)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    const float px = xyz[3 * idx + 0], py = xyz[3 * idx + 1], pz = xyz[3 * idx + 2];
    const float2 f = cam_project(px + st0.x, py + st0.y, pz + st0.z, fr, ur, vr);
    const float2 b = cam_project(px - st1.x, py - st1.y, pz - st1.z, fr, ur, vr);
    const float dx = (f.x - b.x) / 2, dy = (f.y - b.y) / 2;
    const float n = hypotf(dx, dy);
    tvf[idx] = make_float2(dx / n, dy / n);
    const float rx = R.m[0] * px + R.m[1] * py + R.m[2] * pz;
    const float ry = R.m[3] * px + R.m[4] * py + R.m[5] * pz;
    const float rz = R.m[6] * px + R.m[7] * py + R.m[8] * pz;
    const float2 c = cam_project(rx, ry, rz, fo, uo, vo);
    const float2 g = uv[idx];
    cvf[idx] = make_float2(c.x - g.x, c.y - g.y);
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace

extern "C" {

int vfs_cam_get_rays(void* stream, float* out, int H, int W, float f, float u0, float v0)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    hipLaunchKernelGGL(k_get_rays, grid2d(H, W), block2d(), 0, S(stream), out, H, W, f, u0, v0);
    VFS_LAUNCH_STATUS();
}

int vfs_cam_get_uv_from_xyz(void* stream, float* uv, const float* xyz, int H, int W, float f, float u0, float v0)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(uv);
    VFS_CHECK_PTR(xyz);
    hipLaunchKernelGGL(k_uv_from_xyz, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(uv), xyz, H, W, f,
                       u0, v0);
    VFS_LAUNCH_STATUS();
}

int vfs_cam_get_xyz_from_uv(void* stream, float* xyz, const float* uv, int H, int W, float f, float u0, float v0)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(xyz);
    VFS_CHECK_PTR(uv);
    hipLaunchKernelGGL(k_xyz_from_uv, grid2d(H, W), block2d(), 0, S(stream), xyz, reinterpret_cast<const float2*>(uv), H,
                       W, f, u0, v0);
    VFS_LAUNCH_STATUS();
}

int vfs_cam_get_thetas(void* stream, float* out, const float* rays, int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(rays);
    hipLaunchKernelGGL(k_thetas, grid2d(H, W), block2d(), 0, S(stream), out, rays, H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_cam_xyz_from_distance(void* stream, float* out, const float* rays, const float* dist, const uint8_t* mask,
                              int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(out);
    VFS_CHECK_PTR(rays);
    VFS_CHECK_PTR(dist);
    VFS_CHECK_PTR(mask);
    hipLaunchKernelGGL(k_xyz_from_distance, grid2d(H, W), block2d(), 0, S(stream), out, rays, dist, mask, H, W);
    VFS_LAUNCH_STATUS();
}

int vfs_vector_fields(void* stream, float* tvf, float* cvf, const float* xyz, const float* uv, int H, int W,
                      const float* cam_ref, const float* cam_other, const float* R, const float* st0, const float* st1)
{
    VFS_CHECK_DIMS(H, W);
    VFS_CHECK_PTR(tvf);
    VFS_CHECK_PTR(cvf);
    VFS_CHECK_PTR(xyz);
    VFS_CHECK_PTR(uv);
    VFS_CHECK_PTR(cam_ref);
    VFS_CHECK_PTR(cam_other);
    VFS_CHECK_PTR(R);
    VFS_CHECK_PTR(st0);
    VFS_CHECK_PTR(st1);
    Mat3 m;
    for (int i = 0; i < 9; ++i) m.m[i] = R[i];
    hipLaunchKernelGGL(k_vector_fields, grid2d(H, W), block2d(), 0, S(stream), reinterpret_cast<float2*>(tvf),
                       reinterpret_cast<float2*>(cvf), xyz, reinterpret_cast<const float2*>(uv), H, W, cam_ref[0],
                       cam_ref[1], cam_ref[2], cam_other[0], cam_other[1], cam_other[2], m,
                       make_float3(st0[0], st0[1], st0[2]), make_float3(st1[0], st1[1], st1[2]));
    VFS_LAUNCH_STATUS();
}

}  // extern "C"
