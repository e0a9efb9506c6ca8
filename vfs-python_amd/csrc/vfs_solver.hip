// Host-side level loop of VFS.solve_stereo (Python/vfs.py:473-521) with the
// reference's kernel decomposition, enqueued without host synchronisation, and
// its HIP-graph form (the loop recorded once by stream capture and replayed).
#include <new>

#include "vfs_device.h"

using namespace vfs;

namespace {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Per-warp reset of Python/vfs.py:489-497: p = q = v_bar = 0 and u0 = copy(u).
__global__ void __launch_bounds__(256) k_warp_reset(float2* __restrict__ p, float4* __restrict__ q,
                                                    float2* __restrict__ v_bar, float* __restrict__ u0,
                                                    const float* __restrict__ u, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = make_float2(0.0f, 0.0f);
    v_bar[i] = make_float2(0.0f, 0.0f);
    q[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    u0[i] = u[i];
}

// computeOpticalFlowWithClamp followed by u = u_before_solver (Python/vfs.py:519-521):
// u0 holds the warp-start u; the clamped result is written to u, and u0 is
// left as is (it is re-copied at the next warp start).
__global__ void __launch_bounds__(256) k_flow_clamp_to_u(float2* __restrict__ flow, float* __restrict__ u,
                                                         const float* __restrict__ u0, const float* __restrict__ uf,
                                                         const float2* __restrict__ tv, const uint8_t* __restrict__ mask,
                                                         int n, float limit)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float base = u0[i];
    if (mask[i] == 0) {
        u[i] = base;
        return;
    }
    const float delta = fminf(uf[i] - base, limit);
    u[i] = base + delta;
    const float2 t = tv[i];
    float2 f = flow[i];
    f.x = fmaf(delta, t.x, f.x);
    f.y = fmaf(delta, t.y, f.y);
    flow[i] = f;
}

int check_level(const vfs_level* lv)
{
    if (lv == nullptr) return VFS_EINVAL;
    VFS_CHECK_DIMS(lv->H, lv->W);
    const void* ptrs[] = {lv->ref, lv->other, lv->tvf, lv->mask, lv->flow, lv->u, lv->u_bar, lv->ref_smoothed,
                          lv->abc, lv->eta, lv->tv, lv->warped, lv->deriv, lv->p, lv->q, lv->tp, lv->v_bar,
                          lv->u0, lv->uf};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    return 0;
}

#define VFS_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != 0) return rc_;  \
    } while (0)

int solve_level_plain(hipStream_t s, const vfs_level* lv, const vfs_params* prm, const float* taus,
                      const float* sigmas, const float* mus, int n_iter, int n_warp)
{
    VFS_TRY(check_level(lv));
    if (prm == nullptr || n_iter < 0 || n_warp < 0) return VFS_EINVAL;
    if (n_iter > 0 && (taus == nullptr || sigmas == nullptr || mus == nullptr)) return VFS_EINVAL;
    if (prm->median_k != 3 && prm->median_k != 5) return VFS_EINVAL;
    const int H = lv->H, W = lv->W, n = H * W;
    void* st = reinterpret_cast<void*>(s);
    VFS_TRY(vfs_gaussian_masked(st, lv->ref_smoothed, lv->ref, lv->mask, H, W));
    VFS_TRY(vfs_compute_tensor_eta(st, lv->abc, lv->eta, lv->ref_smoothed, lv->mask, H, W, prm));
    const unsigned nb = static_cast<unsigned>((n + 255) / 256);
    for (int w = 0; w < n_warp; ++w) {
        VFS_TRY(vfs_find_warping_vector(st, lv->tv, lv->flow, H, W, lv->tvf, H, W));
        VFS_TRY(vfs_warp_image(st, lv->warped, lv->flow, H, W, lv->other, H, W));
        VFS_TRY(vfs_compute_derivatives(st, lv->deriv, lv->ref, lv->warped, lv->tvf, H, W));
        hipLaunchKernelGGL(k_warp_reset, dim3(nb), dim3(256), 0, s, reinterpret_cast<float2*>(lv->p),
                           reinterpret_cast<float4*>(lv->q), reinterpret_cast<float2*>(lv->v_bar), lv->u0, lv->u, n);
        if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
        for (int it = 0; it < n_iter; ++it) {
            VFS_TRY(vfs_update_dual(st, lv->p, lv->tp, lv->q, sigmas[it], lv->u_bar, lv->v_bar, lv->abc, lv->mask, H, W,
                                    prm));
            VFS_TRY(vfs_l1_thresholding(st, lv->u, taus[it], lv->tp, lv->u, lv->u_bar, lv->deriv, lv->eta, lv->mask, H,
                                        W, prm));
            VFS_TRY(vfs_update_primal(st, lv->u_bar, lv->v_bar, taus[it], mus[it], lv->p, lv->q, lv->u, lv->u_bar,
                                      lv->v_bar, lv->abc, lv->eta, lv->mask, H, W, prm));
        }
        VFS_TRY(vfs_median_filter(st, lv->uf, lv->u, lv->mask, H, W, prm->median_k));
        hipLaunchKernelGGL(k_flow_clamp_to_u, dim3(nb), dim3(256), 0, s, reinterpret_cast<float2*>(lv->flow), lv->u,
                           lv->u0, lv->uf, reinterpret_cast<const float2*>(lv->tv), lv->mask, n, prm->limit_range);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return static_cast<int>(e);
    }
    return 0;
}

struct Plan {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

}  // namespace

extern "C" {

int vfs_solve_level_plain(void* stream, const vfs_level* lv, const vfs_params* prm, const float* taus,
                          const float* sigmas, const float* mus, int n_iter, int n_warp)
{
    return solve_level_plain(S(stream), lv, prm, taus, sigmas, mus, n_iter, n_warp);
}

int vfs_plan_create_level_plain(void** plan, const vfs_level* lv, const vfs_params* prm, const float* taus,
                                const float* sigmas, const float* mus, int n_iter, int n_warp)
{
    if (plan == nullptr) return VFS_EINVAL;
    *plan = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return static_cast<int>(e);
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hipStreamDestroy(s);
        return static_cast<int>(e);
    }
    const int rc = solve_level_plain(s, lv, prm, taus, sigmas, mus, n_iter, n_warp);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(s, &g);
    hipStreamDestroy(s);
    if (rc != 0 || e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return rc != 0 ? rc : static_cast<int>(e);
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return static_cast<int>(e);
    }
    Plan* p = new (std::nothrow) Plan();
    if (p == nullptr) {
        hipGraphExecDestroy(x);
        hipGraphDestroy(g);
        return VFS_EINVAL;
    }
    p->graph = g;
    p->exec = x;
    *plan = p;
    return 0;
}

int vfs_capture_begin(void** capture_stream)
{
    if (capture_stream == nullptr) return VFS_EINVAL;
    *capture_stream = nullptr;
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return static_cast<int>(e);
    e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        hipStreamDestroy(s);
        return static_cast<int>(e);
    }
    *capture_stream = s;
    return 0;
}

int vfs_capture_end(void* capture_stream, void** plan)
{
    if (capture_stream == nullptr) return VFS_EINVAL;
    hipStream_t s = S(capture_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s, &g);
    hipStreamDestroy(s);
    if (plan == nullptr) {
        if (g) hipGraphDestroy(g);
        return VFS_EINVAL;
    }
    *plan = nullptr;
    if (e != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return static_cast<int>(e);
    }
    hipGraphExec_t x = nullptr;
    e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(g);
        return static_cast<int>(e);
    }
    Plan* p = new (std::nothrow) Plan();
    if (p == nullptr) {
        hipGraphExecDestroy(x);
        hipGraphDestroy(g);
        return VFS_EINVAL;
    }
    p->graph = g;
    p->exec = x;
    *plan = p;
    return 0;
}

int vfs_plan_launch(void* plan, void* stream)
{
    if (plan == nullptr) return VFS_EINVAL;
    Plan* p = static_cast<Plan*>(plan);
    hipError_t e = hipGraphLaunch(p->exec, S(stream));
    return e == hipSuccess ? 0 : static_cast<int>(e);
}

int vfs_plan_destroy(void* plan)
{
    if (plan == nullptr) return 0;
    Plan* p = static_cast<Plan*>(plan);
    if (p->exec) hipGraphExecDestroy(p->exec);
    if (p->graph) hipGraphDestroy(p->graph);
    delete p;
    return 0;
}

}  // extern "C"
