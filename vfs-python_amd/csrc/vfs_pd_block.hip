// Temporally blocked primal-dual iterations (K iterations per launch) and the
// fused per-warp kernels of the blocked level loop, gfx950.
//
// One iteration (dual -> threshold -> primal, Python/vfs.py:498-515) reads, at a
// pixel, u_bar / v_bar at its east and south neighbours (the dual step's forward
// differences; west / north where those are masked), and tp and q at its west and
// north neighbours (threshold, primal), both taken from the dual step of the same
// iteration.  The new u_bar of a pixel x therefore depends on the old u_bar at
// x - 1 .. x + 1 only (through p(x - 1), which reads u_bar(x - 1) and u_bar(x), and
// p(x), which reads u_bar(x + 1)), and p, q are exact wherever u_bar was: the
// region on which the state is exact shrinks by one pixel on every side per
// iteration.  (A missing neighbour at the region's border is replaced by the pixel
// itself, as at the image border; only pixels within k of the border after k
// iterations see the difference.)  A workgroup that must produce a TX x TY tile
// after K iterations loads the tile extended by K on every side, keeps each owned
// pixel's state and constants in registers, exchanges the neighbour fields through
// LDS (two barriers per iteration), and writes back only the tile.  Neighbouring
// tiles read their halos from the same input buffers, so the state is ping-ponged
// between two buffer sets across launches.
//
// The per-pixel arithmetic is the same inline code as the per-kernel path
// (vfs_pd.h).
#include <new>
#include <vector>

#include "vfs_device.h"
#include "vfs_internal.h"
#include "vfs_pd.h"
#include "vfs_median.h"

using namespace vfs;

namespace {

constexpr int KMAX = 12;

struct PDSched {
    float tau[KMAX];
    float sp[KMAX];  // dual_scale(alpha1, sigma, eta_p) per iteration
    float sq[KMAX];  // dual_scale(alpha0, sigma, eta_q) per iteration
    float mu[KMAX];
};

struct PDConsts {
    float alpha0, alpha1, eta_p, eta_q, lambda;
};

// One pyramid level of one pair for one launch: constant inputs, the state read
// (in_*) and the state written (out_*), all C-contiguous (H, W[, C]) float32.
struct PDLevel {
    int H, W;
    int tiles_x, tiles_y;
    int first;  // 1: p = q = v_bar = 0 on input (the reset at warp start)
    const uint8_t* mask;
    const float* abc;
    const float* eta;
    const float* deriv;
    const float* u_in;
    const float* ub_in;
    const float* vb_in;
    const float* p_in;
    const float* q_in;
    float* u_out;
    float* ub_out;
    float* vb_out;
    float* p_out;
    float* q_out;
    // Optional: the tile's state after each of the first tape_count iterations of
    // the launch, into tape + k * tape_stride (p at offset 0, then q, u, u_bar, v_bar
    // at the given offsets, in floats).  Used to recompute a tape for the adjoint.
    float* tape = nullptr;
    int64_t tape_stride = 0, tape_q = 0, tape_u = 0, tape_ub = 0, tape_vb = 0;
    int tape_count = 0;
};

enum : unsigned {
    F_ACT = 1u,     // inside the image and inside the mask
    F_ME = 2u,      // mask at the east neighbour (mirror at the border)
    F_MW = 4u,
    F_MS = 8u,
    F_MN = 16u,
    F_HE = 32u,     // east neighbour is a different pixel of the loaded region
    F_HW = 64u,
    F_HS = 128u,
    F_HN = 256u,
    F_IMG = 512u,   // inside the image
};

// The tile pixels' state into tape slot k (see PDLevel::tape).
template <int TX, int TY, int K, int NT, int S>
__device__ __forceinline__ void pd_tape_write(const PDLevel& L, int k, int ox, int oy, const unsigned (&fl)[S],
                                              const float (&u)[S], const float (&ub)[S], const float (&vbx)[S],
                                              const float (&vby)[S], const float (&px)[S], const float (&py)[S],
                                              const float (&qx)[S], const float (&qy)[S], const float (&qz)[S],
                                              const float (&qw)[S])
{
    constexpr int EX = TX + 2 * K, EY = TY + 2 * K, E = EX * EY;
    float* base = L.tape + static_cast<int64_t>(k) * L.tape_stride;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int e = threadIdx.x + NT * j;
        if (e >= E) continue;
        const int ex = e % EX, ey = e / EX;
        if (ex < K || ex >= K + TX || ey < K || ey >= K + TY) continue;
        if (!(fl[j] & F_IMG)) continue;
        const int64_t idx = static_cast<int64_t>(oy + ey) * L.W + (ox + ex);
        reinterpret_cast<float2*>(base)[idx] = make_float2(px[j], py[j]);
        reinterpret_cast<float4*>(base + L.tape_q)[idx] = make_float4(qx[j], qy[j], qz[j], qw[j]);
        base[L.tape_u + idx] = u[j];
        base[L.tape_ub + idx] = ub[j];
        reinterpret_cast<float2*>(base + L.tape_vb)[idx] = make_float2(vbx[j], vby[j]);
    }
}

// One pixel's inputs of a tile launch as they come from memory (the forward state,
// abc, eta, deriv) and its image-position flags.
struct PDRaw {
    float ub, vbx, vby, u, px, py, qx, qy, qz, qw, a, b, c, e0, e1, e2, iu, iz;
    unsigned f;  // F_IMG, F_ACT (mask), F_HE .. F_HN
};

// Global loads of the tile (tx, ty) extended by K: every pixel of the region in the
// image, its u_bar / v_bar (also where masked: the dual step of an inside pixel whose
// east (south) neighbour is masked reads the west (north) one), and inside the mask
// u, p, q and the constants.  Zeros elsewhere, and for v_bar, p, q at a warp's first launch.
template <int TX, int TY, int K, int NT>
__device__ __forceinline__ void pd_load_raw(const PDLevel& L, int tx, int ty,
                                            PDRaw (&r)[((TX + 2 * K) * (TY + 2 * K) + NT - 1) / NT])
{
    constexpr int EX = TX + 2 * K, EY = TY + 2 * K, E = EX * EY;
    constexpr int S = (E + NT - 1) / NT;
    const int tid = threadIdx.x;
    const int H = L.H, W = L.W;
    const int ox = tx * TX - K, oy = ty * TY - K;
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int e = tid + NT * j;
        const int ex = e % EX, ey = e / EX;
        const int gx = ox + ex, gy = oy + ey;
        PDRaw& q = r[j];
        unsigned f = 0;
        q.ub = q.vbx = q.vby = q.u = q.px = q.py = q.qx = q.qy = q.qz = q.qw = 0.0f;
        q.a = q.b = q.c = q.e0 = q.e1 = q.e2 = q.iu = q.iz = 0.0f;
        if (e < E && gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const int idx = gy * W + gx;
            f |= F_IMG;
            if (gx < W - 1 && ex < EX - 1) f |= F_HE;
            if (gx > 0 && ex > 0) f |= F_HW;
            if (gy < H - 1 && ey < EY - 1) f |= F_HS;
            if (gy > 0 && ey > 0) f |= F_HN;
            q.ub = L.ub_in[idx];
            if (!L.first) {
                const float2 v = reinterpret_cast<const float2*>(L.vb_in)[idx];
                q.vbx = v.x;
                q.vby = v.y;
            }
            if (L.mask[idx] != 0) {
                f |= F_ACT;
                q.u = L.u_in[idx];
                if (!L.first) {
                    const float2 p = reinterpret_cast<const float2*>(L.p_in)[idx];
                    const float4 qq = reinterpret_cast<const float4*>(L.q_in)[idx];
                    q.px = p.x;
                    q.py = p.y;
                    q.qx = qq.x;
                    q.qy = qq.y;
                    q.qz = qq.z;
                    q.qw = qq.w;
                }
                q.a = L.abc[3 * idx + 0];
                q.b = L.abc[3 * idx + 1];
                q.c = L.abc[3 * idx + 2];
                q.e0 = L.eta[3 * idx + 0];
                q.e1 = L.eta[3 * idx + 1];
                q.e2 = L.eta[3 * idx + 2];
                const float2 d = reinterpret_cast<const float2*>(L.deriv)[idx];
                q.iu = d.x;
                q.iz = d.y;
            }
        }
        q.f = f;
    }
}

// k_count iterations on the tile (tx, ty) from its loaded inputs r, then the write-back
// of the tile interior (and, with L.tape, the tape of its intermediate states).
template <int TX, int TY, int K, int NT>
__device__ __forceinline__ void pd_tile_run(const PDLevel& L, int tx, int ty, int k_count, const PDSched& sc,
                                            const PDConsts& cst, float* __restrict__ sA, float* __restrict__ sB,
                                            const PDRaw (&r)[((TX + 2 * K) * (TY + 2 * K) + NT - 1) / NT])
{
    constexpr int EX = TX + 2 * K, EY = TY + 2 * K, E = EX * EY;
    constexpr int S = (E + NT - 1) / NT;
    const int tid = threadIdx.x;
    const int W = L.W;
    const int ox = tx * TX - K, oy = ty * TY - K;

    float u[S], ub[S], vbx[S], vby[S], px[S], py[S], qx[S], qy[S], qz[S], qw[S];
    // constants: abc, 1/eta (recip_eta), Iu, 1/Iu (recip_grad), Iz
    float ca[S], cb[S], cc[S], r0[S], r1[S], r2[S], iu[S], riu[S], iz[S];
    unsigned fl[S];

#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int e = tid + NT * j;
        const PDRaw& q = r[j];
        const unsigned f = q.f;
        u[j] = q.u;
        ub[j] = q.ub;
        vbx[j] = q.vbx;
        vby[j] = q.vby;
        px[j] = q.px;
        py[j] = q.py;
        qx[j] = q.qx;
        qy[j] = q.qy;
        qz[j] = q.qz;
        qw[j] = q.qw;
        ca[j] = q.a;
        cb[j] = q.b;
        cc[j] = q.c;
        r0[j] = r1[j] = r2[j] = iu[j] = riu[j] = iz[j] = 0.0f;
        if (f & F_ACT) {
            r0[j] = recip_eta(q.e0, kEps);
            r1[j] = recip_eta(q.e1, kEps);
            r2[j] = recip_eta(q.e2, kEps);
            iu[j] = q.iu;
            riu[j] = recip_grad(q.iu, kEps);
            iz[j] = q.iz;
        }
        fl[j] = f;
        if (e < E) {
            sA[e] = ub[j];
            sA[E + e] = vbx[j];
            sA[2 * E + e] = vby[j];
            sB[e] = (f & F_ACT) ? 1.0f : 0.0f;  // the mask, staged for the neighbour flags
        }
    }

    // A region with no pixel inside the mask ends every iteration with all state
    // zero (masked pixels are zeroed by each of the three steps).
    int any_active = 0;
#pragma unroll
    for (int j = 0; j < S; ++j) any_active |= static_cast<int>(fl[j] & F_ACT);
    const bool skip = __syncthreads_or(any_active) == 0 && k_count > 0;
    // Neighbour mask flags from the staged mask.  A neighbour outside the loaded region
    // is the pixel itself: the mirror at the image border, and at the region's own
    // border a pixel whose state is inexact after the first iteration in any case.
    // (The first barrier of the loop orders these reads before the dual step's sB writes.)
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int e = tid + NT * j;
        const unsigned f = fl[j];
        if (e >= E || !(f & F_ACT)) continue;
        unsigned g = f;
        if (sB[(f & F_HE) ? e + 1 : e] != 0.0f) g |= F_ME;
        if (sB[(f & F_HW) ? e - 1 : e] != 0.0f) g |= F_MW;
        if (sB[(f & F_HS) ? e + EX : e] != 0.0f) g |= F_MS;
        if (sB[(f & F_HN) ? e - EX : e] != 0.0f) g |= F_MN;
        fl[j] = g;
    }
    // Outside the mask every state variable is zero after the first iteration.  u, p and
    // q were never loaded there and the steps below write them only inside the mask; u_bar
    // and v_bar, which the first dual step reads through sA as loaded, are zeroed now.
#pragma unroll
    for (int j = 0; j < S; ++j)
        if (skip || !(fl[j] & F_ACT)) ub[j] = vbx[j] = vby[j] = 0.0f;

    float tpx[S], tpy[S];
    for (int k = 0; k < (skip ? 0 : k_count); ++k) {
        const float tau = sc.tau[k], mu = sc.mu[k], sp = sc.sp[k], sq = sc.sq[k];
        __syncthreads();  // sA (u_bar, v_bar) complete
        // dual step
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int e = tid + NT * j;
            if (e >= E) continue;
            const unsigned f = fl[j];
            if (f & F_ACT) {
                const bool fx = (f & F_ME) != 0, fy = (f & F_MS) != 0;
                const int eE = (f & F_HE) ? e + 1 : e, eW = (f & F_HW) ? e - 1 : e;
                const int eS = (f & F_HS) ? e + EX : e, eN = (f & F_HN) ? e - EX : e;
                const int ix = fx ? eE : eW, iy = fy ? eS : eN;
                float2 pn, tpn;
                float4 qn;
                dual_pixel(ub[j], make_float2(vbx[j], vby[j]), sA[ix], make_float2(sA[E + ix], sA[2 * E + ix]), fx,
                           sA[iy], make_float2(sA[E + iy], sA[2 * E + iy]), fy, ca[j], cb[j], cc[j],
                           make_float2(px[j], py[j]), make_float4(qx[j], qy[j], qz[j], qw[j]), sp, sq, pn,
                           tpn, qn);
                px[j] = pn.x;
                py[j] = pn.y;
                qx[j] = qn.x;
                qy[j] = qn.y;
                qz[j] = qn.z;
                qw[j] = qn.w;
            }
            // T p (dual_pixel's tpn; zero outside the mask, where p and abc are zero), kept
            // in registers for the pixel's own threshold step
            tpx[j] = fmaf(ca[j], px[j], cc[j] * py[j]);
            tpy[j] = fmaf(cc[j], px[j], cb[j] * py[j]);
            sB[e] = tpx[j];
            sB[E + e] = tpy[j];
            sB[2 * E + e] = qx[j];
            sB[3 * E + e] = qy[j];
            sB[4 * E + e] = qz[j];
            sB[5 * E + e] = qw[j];
        }
        __syncthreads();  // sB (tp, q) complete
        // threshold, then primal step (reads sB only; writes sA, last read before the barrier above)
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int e = tid + NT * j;
            if (e >= E) continue;
            const unsigned f = fl[j];
            if (f & F_ACT) {
                const int eW = (f & F_HW) ? e - 1 : e, eN = (f & F_HN) ? e - EX : e;
                const bool mE = (f & F_ME) != 0, mW = (f & F_MW) != 0, mS = (f & F_MS) != 0, mN = (f & F_MN) != 0;
                // ZeroOut: a W / N neighbour outside the mask has T p = q = +0 in sB (its p,
                // q and abc are zero here), which the two steps may then subtract untested
                u[j] = l1_pixel<true>(make_float2(tpx[j], tpy[j]), sB[eW], sB[E + eN], mE, mW, mS, mN, iu[j],
                                      riu[j], iz[j], r0[j], u[j], ub[j], tau, cst.lambda);
                const float4 qW = make_float4(sB[2 * E + eW], sB[3 * E + eW], sB[4 * E + eW], sB[5 * E + eW]);
                const float4 qN = make_float4(sB[2 * E + eN], sB[3 * E + eN], sB[4 * E + eN], sB[5 * E + eN]);
                float ubn;
                float2 vbn;
                primal_pixel<true>(make_float4(qx[j], qy[j], qz[j], qw[j]), qW, qN, mE, mW, mS, mN, ca[j], cb[j], cc[j],
                             make_float2(px[j], py[j]), r1[j], r2[j], make_float2(vbx[j], vby[j]), u[j], ub[j], tau,
                             mu, cst.alpha0, cst.alpha1, ubn, vbn);
                ub[j] = ubn;
                vbx[j] = vbn.x;
                vby[j] = vbn.y;
            }
            sA[e] = ub[j];
            sA[E + e] = vbx[j];
            sA[2 * E + e] = vby[j];
        }
        if (L.tape != nullptr && k < L.tape_count)
            pd_tape_write<TX, TY, K, NT>(L, k, ox, oy, fl, u, ub, vbx, vby, px, py, qx, qy, qz, qw);
    }
    if (skip && L.tape != nullptr) {
        // a fully masked region: every iteration's state is zero
        for (int k = 0; k < k_count && k < L.tape_count; ++k)
            pd_tape_write<TX, TY, K, NT>(L, k, ox, oy, fl, u, ub, vbx, vby, px, py, qx, qy, qz, qw);
    }

    // write back the tile interior
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int e = tid + NT * j;
        if (e >= E) continue;
        const int ex = e % EX, ey = e / EX;
        if (ex < K || ex >= K + TX || ey < K || ey >= K + TY) continue;
        if (!(fl[j] & F_IMG)) continue;
        const int idx = (oy + ey) * W + (ox + ex);
        L.u_out[idx] = u[j];
        L.ub_out[idx] = ub[j];
        reinterpret_cast<float2*>(L.vb_out)[idx] = make_float2(vbx[j], vby[j]);
        reinterpret_cast<float2*>(L.p_out)[idx] = make_float2(px[j], py[j]);
        reinterpret_cast<float4*>(L.q_out)[idx] = make_float4(qx[j], qy[j], qz[j], qw[j]);
    }
}


template <int TX, int TY, int K, int NT>
__device__ __forceinline__ void pd_block_tile(const PDLevel& L, int tx, int ty, int k_count, const PDSched& sc,
                                              const PDConsts& cst, float* __restrict__ sA, float* __restrict__ sB)
{
    PDRaw r[((TX + 2 * K) * (TY + 2 * K) + NT - 1) / NT];
    pd_load_raw<TX, TY, K, NT>(L, tx, ty, r);
    pd_tile_run<TX, TY, K, NT>(L, tx, ty, k_count, sc, cst, sA, sB, r);
}

template <int TX, int TY, int K, int NT>
constexpr int lds_floats()
{
    return 9 * (TX + 2 * K) * (TY + 2 * K);
}

// Tiles of one level; blockIdx.x enumerates tiles row by row.
template <int TX, int TY, int K, int NT>
__global__ void __launch_bounds__(NT) k_pd_block(PDLevel L, int k_count, PDSched sc, PDConsts cst)
{
    constexpr int E = (TX + 2 * K) * (TY + 2 * K);
    __shared__ float sA[3 * E];
    __shared__ float sB[6 * E];
    const int t = xcd_item(blockIdx.x, gridDim.x);
    pd_block_tile<TX, TY, K, NT>(L, t % L.tiles_x, t / L.tiles_x, k_count, sc, cst, sA, sB);
}

// Tiles of several levels / pairs in one launch.  tile_start[i] is the first
// tile index of levels[i] (prefix sums, tile_start[n] = total).
template <int TX, int TY, int K, int NT>
__global__ void __launch_bounds__(NT) k_pd_block_batched(const PDLevel* __restrict__ levels,
                                                         const int* __restrict__ tile_start, int n, int k_count,
                                                         PDSched sc, PDConsts cst)
{
    constexpr int E = (TX + 2 * K) * (TY + 2 * K);
    __shared__ float sA[3 * E];
    __shared__ float sB[6 * E];
    const int t = xcd_item(blockIdx.x, gridDim.x);
    int lo = 0, hi = n - 1;  // last i with tile_start[i] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const PDLevel& L = levels[lo];
    const int lt = t - tile_start[lo];
    pd_block_tile<TX, TY, K, NT>(L, lt % L.tiles_x, lt / L.tiles_x, k_count, sc, cst, sA, sB);
}

// ---------------------------------------------------------------------------
// per-warp kernels of the blocked loop
// ---------------------------------------------------------------------------

// findWarpingVector + warpingImage at one pixel (Python/vfs.py:480,483).
__device__ __forceinline__ void warp_prep_pixel(int X, int Y, int idx, float2 f, const float2* __restrict__ tvf,
                                                const float* __restrict__ other, int H, int W, float2* tv, float* ow)
{
    const float xs = X + f.x, ys = Y + f.y;
    tv[idx] = sample2(tvf, H, W, xs, ys);
    ow[idx] = sample1(other, H, W, xs, ys);
}

__global__ void __launch_bounds__(256) k_warp_prep(float2* __restrict__ tv, float* __restrict__ ow,
                                                   const float2* __restrict__ flow, const float2* __restrict__ tvf,
                                                   const float* __restrict__ other, int H, int W)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    warp_prep_pixel(X, Y, idx, flow[idx], tvf, other, H, W, tv, ow);
}

// medianFilter + computeOpticalFlowWithClamp (+ u = u_before_solver), and the
// next warp's findWarpingVector / warpingImage at the updated flow
// (Python/vfs.py:518-521, then :480,483).  u_ws holds the warp-start u and
// receives the clamped u; u_th is the thresholded u of the primal-dual loop.
template <int KM>
__device__ __forceinline__ void median_clamp_prep_pixel(int X, int Y, float* __restrict__ u_ws,
                                                        float2* __restrict__ flow, float2* __restrict__ tv,
                                                        float* __restrict__ ow, const float* __restrict__ u_th,
                                                        const uint8_t* __restrict__ mask,
                                                        const float2* __restrict__ tvf,
                                                        const float* __restrict__ other, int H, int W, float limit,
                                                        int prep)
{
    const int idx = Y * W + X;
    float2 f = flow[idx];
    if (mask[idx] != 0) {
        constexpr int N = KM * KM;
        float v[N];
#pragma unroll
        for (int j = 0; j < KM; ++j) {
            const int yy = clamp_idx(Y + j - KM / 2, H);
#pragma unroll
            for (int i = 0; i < KM; ++i) v[j * KM + i] = u_th[yy * W + clamp_idx(X + i - KM / 2, W)];
        }
        const float med = median_select<KM>(v);
        const float base = u_ws[idx];
        const float delta = fminf(med - base, limit);
        u_ws[idx] = base + delta;
        const float2 t = tv[idx];
        f.x = fmaf(delta, t.x, f.x);
        f.y = fmaf(delta, t.y, f.y);
        flow[idx] = f;
    }
    if (prep) warp_prep_pixel(X, Y, idx, f, tvf, other, H, W, tv, ow);
}

template <int KM>
__global__ void __launch_bounds__(256) k_median_clamp_prep(float* __restrict__ u_ws, float2* __restrict__ flow,
                                                           float2* __restrict__ tv, float* __restrict__ ow,
                                                           const float* __restrict__ u_th,
                                                           const uint8_t* __restrict__ mask,
                                                           const float2* __restrict__ tvf,
                                                           const float* __restrict__ other, int H, int W, float limit,
                                                           int prep)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    median_clamp_prep_pixel<KM>(X, Y, u_ws, flow, tv, ow, u_th, mask, tvf, other, H, W, limit, prep);
}

// ---------------------------------------------------------------------------
// batched per-warp kernels: many pairs (of different sizes) in one launch
// ---------------------------------------------------------------------------

// One pair's level for the per-warp kernels.  Tiles are 64 x 4 pixels (one
// wavefront per row segment); tiles_x = ceil(W / 64).
struct WarpDesc {
    int H, W, tiles_x;
    const float* ref;
    const float* other;
    const float2* tvf;
    const uint8_t* mask;
    float2* flow;
    float* u;
    float2* tv;
    float* warped;
    float2* deriv;
    const float* u_th;
};

__device__ __forceinline__ int find_item(const int* __restrict__ start, int n, int t)
{
    int lo = 0, hi = n - 1;  // last i with start[i] <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(256) k_derivatives_batched(const WarpDesc* __restrict__ d,
                                                             const int* __restrict__ start, int n)
{
    const int t = blockIdx.x;
    const int i = find_item(start, n, t);
    const WarpDesc& D = d[i];
    const int lt = t - start[i];
    const int X = (lt % D.tiles_x) * 64 + (threadIdx.x & 63);
    const int Y = (lt / D.tiles_x) * 4 + (threadIdx.x >> 6);
    if (X >= D.W || Y >= D.H) return;
    const int idx = Y * D.W + X;
    D.deriv[idx] = derivative_pixel(X, Y, idx, D.ref, D.warped, D.tvf, D.H, D.W);
}

template <int KM>
__global__ void __launch_bounds__(256) k_median_clamp_prep_batched(const WarpDesc* __restrict__ d,
                                                                   const int* __restrict__ start, int n, float limit,
                                                                   int prep)
{
    const int t = blockIdx.x;
    const int i = find_item(start, n, t);
    const WarpDesc& D = d[i];
    const int lt = t - start[i];
    const int X = (lt % D.tiles_x) * 64 + (threadIdx.x & 63);
    const int Y = (lt / D.tiles_x) * 4 + (threadIdx.x >> 6);
    if (X >= D.W || Y >= D.H) return;
    median_clamp_prep_pixel<KM>(X, Y, D.u, D.flow, D.tv, D.warped, D.u_th, D.mask, D.tvf, D.other, D.H, D.W, limit,
                                prep);
}

// ---------------------------------------------------------------------------
// launch helpers
// ---------------------------------------------------------------------------

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

using PDFn = void (*)(PDLevel, int, PDSched, PDConsts);
using PDBFn = void (*)(const PDLevel*, const int*, int, int, PDSched, PDConsts);

struct Variant {
    int tx, ty, k, nt;
    PDFn single;
    PDBFn batched;
};

#define VFS_PD_VARIANT(TX, TY, K, NT) \
    {TX, TY, K, NT, k_pd_block<TX, TY, K, NT>, k_pd_block_batched<TX, TY, K, NT>}

// Tile shapes and iteration counts available to the autotuner.  LDS use is
// 36 bytes per pixel of the extended (TX + 2K) x (TY + 2K) region.
const Variant kVariants[] = {
    VFS_PD_VARIANT(16, 16, 1, 256),  VFS_PD_VARIANT(16, 16, 2, 512),  VFS_PD_VARIANT(16, 16, 4, 1024),
    VFS_PD_VARIANT(32, 32, 1, 1024), VFS_PD_VARIANT(32, 32, 2, 1024), VFS_PD_VARIANT(32, 32, 4, 1024),
    VFS_PD_VARIANT(64, 16, 2, 1024), VFS_PD_VARIANT(64, 16, 4, 1024), VFS_PD_VARIANT(32, 16, 4, 1024),
    VFS_PD_VARIANT(32, 16, 8, 1024), VFS_PD_VARIANT(8, 8, 8, 512),    VFS_PD_VARIANT(16, 16, 2, 256),
    VFS_PD_VARIANT(8, 8, 1, 128),    VFS_PD_VARIANT(8, 8, 2, 256),    VFS_PD_VARIANT(16, 8, 2, 256),
    // tiles whose extended region fills the workgroup exactly (no idle lanes),
    // 1024 lanes on 32 x 32: TX + 2K = 32
    VFS_PD_VARIANT(30, 30, 1, 1024), VFS_PD_VARIANT(28, 28, 2, 1024), VFS_PD_VARIANT(26, 26, 3, 1024),
    VFS_PD_VARIANT(24, 24, 4, 1024), VFS_PD_VARIANT(22, 22, 5, 1024), VFS_PD_VARIANT(20, 20, 6, 1024),
    VFS_PD_VARIANT(18, 18, 7, 1024), VFS_PD_VARIANT(16, 16, 8, 1024),
    // deeper, for the latency-bound coarse levels: fewer launches per warp
    VFS_PD_VARIANT(12, 12, 10, 1024), VFS_PD_VARIANT(8, 8, 12, 1024),
    // two pixels per lane (64 x 32), 512 lanes on 32 x 16, 256 lanes on 16 x 16
    VFS_PD_VARIANT(60, 28, 2, 1024), VFS_PD_VARIANT(58, 26, 3, 1024), VFS_PD_VARIANT(56, 24, 4, 1024),
    VFS_PD_VARIANT(28, 12, 2, 512),  VFS_PD_VARIANT(26, 10, 3, 512),  VFS_PD_VARIANT(24, 8, 4, 512),
    VFS_PD_VARIANT(14, 14, 1, 256),  VFS_PD_VARIANT(12, 12, 2, 256),  VFS_PD_VARIANT(10, 10, 3, 256),
    VFS_PD_VARIANT(8, 8, 4, 256),
    // coarse levels (one workgroup per CU there): a 27 x 27 region on 768 lanes puts three
    // wavefronts on each SIMD instead of four, and 7 x 7 tiles still cover 100 x 100 with
    // 225 workgroups
    VFS_PD_VARIANT(7, 7, 10, 768),
};
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// The variant (tx, ty, k) with nt threads, or with the first listed thread count for nt == 0.
const Variant* find_variant(int tx, int ty, int k, int nt)
{
    for (int i = 0; i < kNumVariants; ++i)
        if (kVariants[i].tx == tx && kVariants[i].ty == ty && kVariants[i].k == k && (nt == 0 || kVariants[i].nt == nt))
            return &kVariants[i];
    return nullptr;
}

PDConsts make_consts(const vfs_params* prm)
{
    PDConsts c;
    c.alpha0 = prm->alpha0;
    c.alpha1 = prm->alpha1;
    c.eta_p = prm->eta_p;
    c.eta_q = prm->eta_q;
    c.lambda = prm->lambda;
    return c;
}

// Schedule of one launch: tau, mu and the two dual step sizes (formed on the
// host with the same fp32 operations as the per-kernel path forms them).
PDSched make_sched(const vfs_params* prm, const float* taus, const float* sigmas, const float* mus, int n)
{
    PDSched sc;
    for (int i = 0; i < KMAX; ++i) {
        const bool in = i < n;
        sc.tau[i] = in ? taus[i] : 0.0f;
        sc.mu[i] = in ? mus[i] : 0.0f;
        sc.sp[i] = in ? dual_scale(prm->alpha1, sigmas[i], prm->eta_p) : 0.0f;
        sc.sq[i] = in ? dual_scale(prm->alpha0, sigmas[i], prm->eta_q) : 0.0f;
    }
    return sc;
}

int launch_pd_single(hipStream_t s, const Variant* v, PDLevel L, int k_count, const PDSched& sc, const PDConsts& c)
{
    L.tiles_x = (L.W + v->tx - 1) / v->tx;
    L.tiles_y = (L.H + v->ty - 1) / v->ty;
    hipLaunchKernelGGL(v->single, dim3(L.tiles_x * L.tiles_y), dim3(v->nt), 0, s, L, k_count, sc, c);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : static_cast<int>(e);
}

#define VFS_TRY(expr)              \
    do {                           \
        int rc_ = (expr);          \
        if (rc_ != 0) return rc_;  \
    } while (0)

}  // namespace

extern "C" {

int vfs_pd_variants(int* out, int max_variants)
{
    if (out == nullptr) return kNumVariants;
    const int n = max_variants < kNumVariants ? max_variants : kNumVariants;
    for (int i = 0; i < n; ++i) {
        out[4 * i + 0] = kVariants[i].tx;
        out[4 * i + 1] = kVariants[i].ty;
        out[4 * i + 2] = kVariants[i].k;
        out[4 * i + 3] = kVariants[i].nt;
    }
    return n;
}

int vfs_pd_blocks(void* stream, int H, int W, const uint8_t* mask, const float* abc, const float* eta,
                  const float* deriv, const float* u_in, const float* ub_in, const float* vb_in, const float* p_in,
                  const float* q_in, float* u_out, float* ub_out, float* vb_out, float* p_out, float* q_out,
                  int first, const float* taus, const float* sigmas, const float* mus, int n_iter, int tile_x,
                  int tile_y, int k, int threads, const vfs_params* prm)
{
    VFS_CHECK_DIMS(H, W);
    const void* ptrs[] = {mask, abc, eta, deriv, u_in, ub_in, u_out, ub_out, vb_out, p_out, q_out, taus, sigmas, mus,
                          prm};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    if (!first && (vb_in == nullptr || p_in == nullptr || q_in == nullptr)) return VFS_EINVAL;
    if (n_iter < 1 || n_iter > k) return VFS_EINVAL;
    const Variant* v = find_variant(tile_x, tile_y, k, threads);
    if (v == nullptr) return VFS_EINVAL;
    PDLevel L;
    L.H = H;
    L.W = W;
    L.first = first;
    L.mask = mask;
    L.abc = abc;
    L.eta = eta;
    L.deriv = deriv;
    L.u_in = u_in;
    L.ub_in = ub_in;
    L.vb_in = vb_in;
    L.p_in = p_in;
    L.q_in = q_in;
    L.u_out = u_out;
    L.ub_out = ub_out;
    L.vb_out = vb_out;
    L.p_out = p_out;
    L.q_out = q_out;
    const PDSched sc = make_sched(prm, taus, sigmas, mus, n_iter);
    return launch_pd_single(S(stream), v, L, n_iter, sc, make_consts(prm));
}

}  // extern "C"

namespace vfs {

int pd_blocks_tape(hipStream_t s, const PdTapeLaunch& a, const vfs_params* prm)
{
    VFS_CHECK_DIMS(a.H, a.W);
    if (prm == nullptr || a.n_iter < 1 || a.n_iter > a.k || a.tape_count < 0 || a.tape_count > a.n_iter)
        return VFS_EINVAL;
    if (a.tape_count > 0 && a.tape == nullptr) return VFS_EINVAL;
    const Variant* v = find_variant(a.tile_x, a.tile_y, a.k, a.threads);
    if (v == nullptr) return VFS_EINVAL;
    PDLevel L;
    L.H = a.H;
    L.W = a.W;
    L.first = a.first ? 1 : 0;
    L.mask = a.mask;
    L.abc = a.abc;
    L.eta = a.eta;
    L.deriv = a.deriv;
    L.p_in = a.in[0];
    L.q_in = a.in[1];
    L.u_in = a.in[2];
    L.ub_in = a.in[3];
    L.vb_in = a.in[4];
    L.p_out = a.out[0];
    L.q_out = a.out[1];
    L.u_out = a.out[2];
    L.ub_out = a.out[3];
    L.vb_out = a.out[4];
    L.tape = a.tape;
    L.tape_stride = a.tape_stride;
    L.tape_q = a.tape_off[0];
    L.tape_u = a.tape_off[1];
    L.tape_ub = a.tape_off[2];
    L.tape_vb = a.tape_off[3];
    L.tape_count = a.tape_count;
    const PDSched sc = make_sched(prm, a.taus, a.sigmas, a.mus, a.n_iter);
    return launch_pd_single(s, v, L, a.n_iter, sc, make_consts(prm));
}

int warp_prep_launch(hipStream_t s, float* tv, float* warped, const float* flow, const float* tvf, const float* other,
                     int H, int W)
{
    VFS_CHECK_DIMS(H, W);
    hipLaunchKernelGGL(k_warp_prep, grid2d(H, W), block2d(), 0, s, reinterpret_cast<float2*>(tv), warped,
                       reinterpret_cast<const float2*>(flow), reinterpret_cast<const float2*>(tvf), other, H, W);
    VFS_LAUNCH_STATUS();
}

int median_clamp_prep_launch(hipStream_t s, int median_k, float* u_ws, float* flow, float* tv, float* warped,
                             const float* u_th, const uint8_t* mask, const float* tvf, const float* other, int H,
                             int W, float limit, int prep)
{
    VFS_CHECK_DIMS(H, W);
    if (median_k == 5)
        hipLaunchKernelGGL(k_median_clamp_prep<5>, grid2d(H, W), block2d(), 0, s, u_ws, reinterpret_cast<float2*>(flow),
                           reinterpret_cast<float2*>(tv), warped, u_th, mask, reinterpret_cast<const float2*>(tvf),
                           other, H, W, limit, prep);
    else if (median_k == 3)
        hipLaunchKernelGGL(k_median_clamp_prep<3>, grid2d(H, W), block2d(), 0, s, u_ws, reinterpret_cast<float2*>(flow),
                           reinterpret_cast<float2*>(tv), warped, u_th, mask, reinterpret_cast<const float2*>(tvf),
                           other, H, W, limit, prep);
    else
        return VFS_EINVAL;
    VFS_LAUNCH_STATUS();
}

}  // namespace vfs

extern "C" {

int vfs_solve_level_blocked(void* stream, const vfs_level* lv, const vfs_pd_ws* ws, const vfs_params* prm,
                            const float* taus, const float* sigmas, const float* mus, int n_iter, int n_warp,
                            int tile_x, int tile_y, int k, int threads)
{
    if (lv == nullptr || ws == nullptr || prm == nullptr) return VFS_EINVAL;
    VFS_CHECK_DIMS(lv->H, lv->W);
    const void* ptrs[] = {lv->ref, lv->other, lv->tvf, lv->mask, lv->flow, lv->u, lv->u_bar, lv->ref_smoothed,
                          lv->abc, lv->eta, lv->tv, lv->warped, lv->deriv, lv->p, lv->q, lv->v_bar,
                          ws->u_th0, ws->u_th1, ws->u_bar2, ws->v_bar2, ws->p2, ws->q2};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    if (n_iter < 0 || n_warp < 0 || (n_iter > 0 && (taus == nullptr || sigmas == nullptr || mus == nullptr)))
        return VFS_EINVAL;
    if (prm->median_k != 3 && prm->median_k != 5) return VFS_EINVAL;
    const Variant* v = find_variant(tile_x, tile_y, k, threads);
    if (v == nullptr || k > KMAX) return VFS_EINVAL;
    hipStream_t s = S(stream);
    const int H = lv->H, W = lv->W;
    VFS_TRY(vfs_gaussian_masked(stream, lv->ref_smoothed, lv->ref, lv->mask, H, W));
    VFS_TRY(vfs_compute_tensor_eta(stream, lv->abc, lv->eta, lv->ref_smoothed, lv->mask, H, W, prm));
    if (n_warp == 0) return 0;
    hipLaunchKernelGGL(k_warp_prep, grid2d(H, W), block2d(), 0, s, reinterpret_cast<float2*>(lv->tv), lv->warped,
                       reinterpret_cast<const float2*>(lv->flow), reinterpret_cast<const float2*>(lv->tvf),
                       lv->other, H, W);
    // state set 0 / 1 (ping-pong): thresholded u, u_bar, v_bar, p, q
    float* su[2] = {ws->u_th0, ws->u_th1};
    float* sub[2] = {lv->u_bar, ws->u_bar2};
    float* svb[2] = {lv->v_bar, ws->v_bar2};
    float* sp[2] = {lv->p, ws->p2};
    float* sq[2] = {lv->q, ws->q2};
    int cur = 0;  // set holding u_bar
    const PDConsts cst = make_consts(prm);
    for (int w = 0; w < n_warp; ++w) {
        VFS_TRY(vfs_compute_derivatives(stream, lv->deriv, lv->ref, lv->warped, lv->tvf, H, W));
        int done = 0;
        bool first = true;
        while (done < n_iter) {
            const int kc = (n_iter - done) < k ? (n_iter - done) : k;
            PDLevel L;
            L.H = H;
            L.W = W;
            L.first = first ? 1 : 0;
            L.mask = lv->mask;
            L.abc = lv->abc;
            L.eta = lv->eta;
            L.deriv = lv->deriv;
            L.u_in = first ? lv->u : su[cur];
            L.ub_in = sub[cur];
            L.vb_in = svb[cur];
            L.p_in = sp[cur];
            L.q_in = sq[cur];
            L.u_out = su[1 - cur];
            L.ub_out = sub[1 - cur];
            L.vb_out = svb[1 - cur];
            L.p_out = sp[1 - cur];
            L.q_out = sq[1 - cur];
            const PDSched sc = make_sched(prm, taus + done, sigmas + done, mus + done, kc);
            VFS_TRY(launch_pd_single(s, v, L, kc, sc, cst));
            cur = 1 - cur;
            done += kc;
            first = false;
        }
        // with n_iter == 0 the thresholded u is the warp-start u
        const float* uth = n_iter > 0 ? su[cur] : lv->u;
        const int prep = (w + 1 < n_warp) ? 1 : 0;
        if (prm->median_k == 5)
            hipLaunchKernelGGL(k_median_clamp_prep<5>, grid2d(H, W), block2d(), 0, s, lv->u,
                               reinterpret_cast<float2*>(lv->flow), reinterpret_cast<float2*>(lv->tv), lv->warped,
                               uth, lv->mask, reinterpret_cast<const float2*>(lv->tvf), lv->other, H, W,
                               prm->limit_range, prep);
        else
            hipLaunchKernelGGL(k_median_clamp_prep<3>, grid2d(H, W), block2d(), 0, s, lv->u,
                               reinterpret_cast<float2*>(lv->flow), reinterpret_cast<float2*>(lv->tv), lv->warped,
                               uth, lv->mask, reinterpret_cast<const float2*>(lv->tvf), lv->other, H, W,
                               prm->limit_range, prep);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return static_cast<int>(e);
    }
    if (cur == 1) VFS_TRY(vfs_copy(stream, lv->u_bar, ws->u_bar2, static_cast<int64_t>(H) * W));
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches of pairs: one level of every pair per call
// ---------------------------------------------------------------------------

namespace {

constexpr size_t kAlign = 256;

inline size_t align_up(size_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

struct BatchLayout {
    size_t pd, pd_start, wd, wd_start, total;
};

BatchLayout batch_layout(int n)
{
    BatchLayout l;
    l.pd = 0;
    l.pd_start = align_up(l.pd + 4 * static_cast<size_t>(n) * sizeof(PDLevel));
    l.wd = align_up(l.pd_start + (static_cast<size_t>(n) + 1) * sizeof(int));
    l.wd_start = align_up(l.wd + 2 * static_cast<size_t>(n) * sizeof(WarpDesc));
    l.total = align_up(l.wd_start + (static_cast<size_t>(n) + 1) * sizeof(int));
    return l;
}

int check_pair(const vfs_level* lv, const vfs_pd_ws* ws)
{
    if (lv == nullptr || ws == nullptr) return VFS_EINVAL;
    VFS_CHECK_DIMS(lv->H, lv->W);
    const void* ptrs[] = {lv->ref, lv->other, lv->tvf, lv->mask, lv->flow, lv->u, lv->u_bar, lv->ref_smoothed,
                          lv->abc, lv->eta, lv->tv, lv->warped, lv->deriv, lv->p, lv->q, lv->v_bar,
                          ws->u_th0, ws->u_th1, ws->u_bar2, ws->v_bar2, ws->p2, ws->q2};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    return 0;
}

}  // namespace

extern "C" {

size_t vfs_batch_workspace_bytes(int n_pairs)
{
    return n_pairs > 0 ? batch_layout(n_pairs).total : 0;
}

int vfs_batch_prepare_level(void* d_workspace, size_t workspace_bytes, const vfs_level* lv, const vfs_pd_ws* ws,
                            int n_pairs, int tile_x, int tile_y, int k, int threads)
{
    if (d_workspace == nullptr || lv == nullptr || ws == nullptr || n_pairs <= 0) return VFS_EINVAL;
    const BatchLayout lay = batch_layout(n_pairs);
    if (workspace_bytes < lay.total) return VFS_EINVAL;
    if (find_variant(tile_x, tile_y, k, threads) == nullptr) return VFS_EINVAL;
    std::vector<unsigned char> host(lay.total, 0);
    PDLevel* pd = reinterpret_cast<PDLevel*>(host.data() + lay.pd);
    int* pd_start = reinterpret_cast<int*>(host.data() + lay.pd_start);
    WarpDesc* wd = reinterpret_cast<WarpDesc*>(host.data() + lay.wd);
    int* wd_start = reinterpret_cast<int*>(host.data() + lay.wd_start);
    int pt = 0, wt = 0;
    for (int i = 0; i < n_pairs; ++i) {
        VFS_TRY(check_pair(&lv[i], &ws[i]));
        const vfs_level& l = lv[i];
        const vfs_pd_ws& w = ws[i];
        float* su[2] = {w.u_th0, w.u_th1};
        float* sub[2] = {l.u_bar, w.u_bar2};
        float* svb[2] = {l.v_bar, w.v_bar2};
        float* sp[2] = {l.p, w.p2};
        float* sq[2] = {l.q, w.q2};
        // tables: 0 / 1 = first launch of a warp with u_bar in set 0 / 1; 2 / 3 = later launches reading set 0 / 1
        for (int t = 0; t < 4; ++t) {
            const int cur = t & 1, first = t < 2;
            PDLevel& L = pd[t * n_pairs + i];
            L.H = l.H;
            L.W = l.W;
            L.tiles_x = (l.W + tile_x - 1) / tile_x;
            L.tiles_y = (l.H + tile_y - 1) / tile_y;
            L.first = first;
            L.mask = l.mask;
            L.abc = l.abc;
            L.eta = l.eta;
            L.deriv = l.deriv;
            L.u_in = first ? l.u : su[cur];
            L.ub_in = sub[cur];
            L.vb_in = svb[cur];
            L.p_in = sp[cur];
            L.q_in = sq[cur];
            L.u_out = su[1 - cur];
            L.ub_out = sub[1 - cur];
            L.vb_out = svb[1 - cur];
            L.p_out = sp[1 - cur];
            L.q_out = sq[1 - cur];
        }
        pd_start[i] = pt;
        pt += pd[i].tiles_x * pd[i].tiles_y;
        for (int t = 0; t < 2; ++t) {
            WarpDesc& D = wd[t * n_pairs + i];
            D.H = l.H;
            D.W = l.W;
            D.tiles_x = (l.W + 63) / 64;
            D.ref = l.ref;
            D.other = l.other;
            D.tvf = reinterpret_cast<const float2*>(l.tvf);
            D.mask = l.mask;
            D.flow = reinterpret_cast<float2*>(l.flow);
            D.u = l.u;
            D.tv = reinterpret_cast<float2*>(l.tv);
            D.warped = l.warped;
            D.deriv = reinterpret_cast<float2*>(l.deriv);
            D.u_th = su[t];
        }
        wd_start[i] = wt;
        wt += ((l.W + 63) / 64) * ((l.H + 3) / 4);
    }
    pd_start[n_pairs] = pt;
    wd_start[n_pairs] = wt;
    hipError_t e = hipMemcpy(d_workspace, host.data(), lay.total, hipMemcpyHostToDevice);
    return e == hipSuccess ? 0 : static_cast<int>(e);
}

int vfs_pd_blocks_batch(void* stream, const void* d_workspace, const vfs_level* lv, const vfs_pd_ws* ws, int n_pairs,
                        const vfs_params* prm, const float* taus, const float* sigmas, const float* mus, int n_iter,
                        int tile_x, int tile_y, int k, int threads)
{
    if (d_workspace == nullptr || lv == nullptr || ws == nullptr || prm == nullptr || n_pairs <= 0) return VFS_EINVAL;
    if (n_iter < 1 || taus == nullptr || sigmas == nullptr || mus == nullptr) return VFS_EINVAL;
    const Variant* v = find_variant(tile_x, tile_y, k, threads);
    if (v == nullptr) return VFS_EINVAL;
    hipStream_t s = S(stream);
    const BatchLayout lay = batch_layout(n_pairs);
    const PDLevel* pd = reinterpret_cast<const PDLevel*>(static_cast<const unsigned char*>(d_workspace) + lay.pd);
    const int* pd_start = reinterpret_cast<const int*>(static_cast<const unsigned char*>(d_workspace) + lay.pd_start);
    int pd_tiles = 0;
    for (int i = 0; i < n_pairs; ++i) {
        VFS_TRY(check_pair(&lv[i], &ws[i]));
        pd_tiles += ((lv[i].W + tile_x - 1) / tile_x) * ((lv[i].H + tile_y - 1) / tile_y);
    }
    const PDConsts cst = make_consts(prm);
    int cur = 0, done = 0;
    bool first = true;
    while (done < n_iter) {
        const int kc = (n_iter - done) < k ? (n_iter - done) : k;
        const PDSched sc = make_sched(prm, taus + done, sigmas + done, mus + done, kc);
        hipLaunchKernelGGL(v->batched, dim3(pd_tiles), dim3(v->nt), 0, s, pd + ((first ? 0 : 2) + cur) * n_pairs,
                           pd_start, n_pairs, kc, sc, cst);
        cur = 1 - cur;
        done += kc;
        first = false;
    }
    VFS_LAUNCH_STATUS();
}

int vfs_solve_level_blocked_batch(void* stream, const void* d_workspace, const vfs_level* lv, const vfs_pd_ws* ws,
                                  int n_pairs, const vfs_params* prm, const float* taus, const float* sigmas,
                                  const float* mus, int n_iter, int n_warp, int tile_x, int tile_y, int k,
                                  int threads)
{
    if (d_workspace == nullptr || lv == nullptr || ws == nullptr || prm == nullptr || n_pairs <= 0) return VFS_EINVAL;
    if (n_iter < 0 || n_warp < 0 || (n_iter > 0 && (taus == nullptr || sigmas == nullptr || mus == nullptr)))
        return VFS_EINVAL;
    if (prm->median_k != 3 && prm->median_k != 5) return VFS_EINVAL;
    if (n_warp > 0 && n_iter == 0) return VFS_EINVAL;  // the batched median reads this warp's thresholded u
    const Variant* v = find_variant(tile_x, tile_y, k, threads);
    if (v == nullptr) return VFS_EINVAL;
    hipStream_t s = S(stream);
    const BatchLayout lay = batch_layout(n_pairs);
    const unsigned char* base = static_cast<const unsigned char*>(d_workspace);
    const PDLevel* pd = reinterpret_cast<const PDLevel*>(base + lay.pd);
    const int* pd_start = reinterpret_cast<const int*>(base + lay.pd_start);
    const WarpDesc* wd = reinterpret_cast<const WarpDesc*>(base + lay.wd);
    const int* wd_start = reinterpret_cast<const int*>(base + lay.wd_start);
    int pd_tiles = 0, px_tiles = 0;
    for (int i = 0; i < n_pairs; ++i) {
        VFS_TRY(check_pair(&lv[i], &ws[i]));
        const vfs_level& l = lv[i];
        pd_tiles += ((l.W + tile_x - 1) / tile_x) * ((l.H + tile_y - 1) / tile_y);
        px_tiles += ((l.W + 63) / 64) * ((l.H + 3) / 4);
        VFS_TRY(vfs_gaussian_masked(stream, l.ref_smoothed, l.ref, l.mask, l.H, l.W));
        VFS_TRY(vfs_compute_tensor_eta(stream, l.abc, l.eta, l.ref_smoothed, l.mask, l.H, l.W, prm));
        if (n_warp > 0)
            hipLaunchKernelGGL(k_warp_prep, grid2d(l.H, l.W), block2d(), 0, s, reinterpret_cast<float2*>(l.tv),
                               l.warped, reinterpret_cast<const float2*>(l.flow),
                               reinterpret_cast<const float2*>(l.tvf), l.other, l.H, l.W);
    }
    const PDConsts cst = make_consts(prm);
    int cur = 0;
    for (int w = 0; w < n_warp; ++w) {
        hipLaunchKernelGGL(k_derivatives_batched, dim3(px_tiles), dim3(256), 0, s, wd, wd_start, n_pairs);
        int done = 0;
        bool first = true;
        while (done < n_iter) {
            const int kc = (n_iter - done) < k ? (n_iter - done) : k;
            const int table = (first ? 0 : 2) + cur;
            const PDSched sc = make_sched(prm, taus + done, sigmas + done, mus + done, kc);
            hipLaunchKernelGGL(v->batched, dim3(pd_tiles), dim3(v->nt), 0, s, pd + table * n_pairs, pd_start,
                               n_pairs, kc, sc, cst);
            cur = 1 - cur;
            done += kc;
            first = false;
        }
        const int prep = (w + 1 < n_warp) ? 1 : 0;
        if (prm->median_k == 5)
            hipLaunchKernelGGL(k_median_clamp_prep_batched<5>, dim3(px_tiles), dim3(256), 0, s, wd + cur * n_pairs,
                               wd_start, n_pairs, prm->limit_range, prep);
        else
            hipLaunchKernelGGL(k_median_clamp_prep_batched<3>, dim3(px_tiles), dim3(256), 0, s, wd + cur * n_pairs,
                               wd_start, n_pairs, prm->limit_range, prep);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return static_cast<int>(e);
    }
    if (cur == 1)
        for (int i = 0; i < n_pairs; ++i)
            VFS_TRY(vfs_copy(stream, lv[i].u_bar, ws[i].u_bar2, static_cast<int64_t>(lv[i].H) * lv[i].W));
    return 0;
}

}  // extern "C"
