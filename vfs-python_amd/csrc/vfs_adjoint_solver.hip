// Native level driver of the adjoint: the forward warp loop of one level that also
// records what the reverse sweep needs, and the whole reverse sweep of one level.
//
// Save area of a level (caller-allocated, vfs_adj_save_bytes): for every warp the
// warp-start u, u_bar and flow, then ceil(n_iter / ckpt) checkpoints of the
// primal-dual state (p, q, u, u_bar, v_bar) taken before iterations 0, ckpt,
// 2 ckpt, ...  Workspace (vfs_adj_workspace_bytes): a tape of ckpt states that
// one checkpoint segment is recomputed into, and the adjoint buffers.  Every
// array starts on a 256-byte boundary, so the float2 / float4 views the
// kernels take stay aligned for any H x W.
//
// Everything is enqueued on one stream without host synchronisation and
// without memsets (clearing is done by a fill kernel), so both functions can
// be recorded into a HIP graph by stream capture.
#include "vfs_adj_pixel.h"
#include "vfs_internal.h"

using namespace vfs;

namespace {

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

constexpr int64_t kAlign = 64;  // floats (256 bytes)
inline int64_t al(int64_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

struct State {
    float *p, *q, *u, *ub, *vb;
};

inline int64_t state_floats(int64_t n) { return al(2 * n) + al(4 * n) + 2 * al(n) + al(2 * n); }

inline State state_at(float* base, int64_t n)
{
    State s;
    s.p = base;
    s.q = s.p + al(2 * n);
    s.u = s.q + al(4 * n);
    s.ub = s.u + al(n);
    s.vb = s.ub + al(n);
    return s;
}

inline int n_ckpt(int n_iter, int ck) { return n_iter <= 0 ? 0 : (n_iter + ck - 1) / ck; }

// iterations of the last checkpoint segment (the states a tape of it holds)
inline int last_seg_len(int n_iter, int ck)
{
    const int nck = n_ckpt(n_iter, ck);
    return nck > 0 ? n_iter - (nck - 1) * ck : 0;
}

// the last-segment tape of warp w, or nullptr (vfs_adj_opts.tape_save / tape_warps)
inline float* warp_tape_at(const vfs_adj_opts* o, int64_t n, int n_iter, int ck, int n_warp, int w)
{
    if (o == nullptr || o->tape_warps <= 0 || w < n_warp - o->tape_warps) return nullptr;
    return o->tape_save + static_cast<int64_t>(w - (n_warp - o->tape_warps)) * last_seg_len(n_iter, ck) *
                              state_floats(n);
}

// per warp: u, u_bar, flow, the warp's derivatives (Iu, Iz), then the checkpoints
inline int64_t warp_rec_floats(int64_t n) { return 2 * al(n) + 2 * al(2 * n); }
inline int64_t warp_save_floats(int64_t n, int n_iter, int ck)
{
    return warp_rec_floats(n) + static_cast<int64_t>(n_ckpt(n_iter, ck)) * state_floats(n);
}

struct WarpSave {
    float *u, *ub, *flow, *deriv;
    float* ckpt0;
};

inline WarpSave warp_save_at(float* save, int64_t n, int n_iter, int ck, int w)
{
    WarpSave r;
    r.u = save + static_cast<int64_t>(w) * warp_save_floats(n, n_iter, ck);
    r.ub = r.u + al(n);
    r.flow = r.ub + al(n);
    r.deriv = r.flow + al(2 * n);
    r.ckpt0 = r.deriv + al(2 * n);
    return r;
}

// adjoint workspace (after the tape)
struct AdjWs {
    float* tp;  // forward recompute's T p
    float *gp, *gq, *um, *vm, *qmsg, *tmsg, *gabc, *geta, *gderiv, *gvbB, *gubB, *guf, *gtv, *gow, *guth, *gvb_fin,
        *gIs;
    int* pos;
    // second copies: a fused launch reads one and writes the other (its ring reads
    // a neighbouring tile's values, which that tile's launch also rewrites)
    float *gp2, *gq2, *qmsg2, *tmsg2;
    // second copies of the state the blocked sweep also reads across tiles
    float *gu2, *gub2, *gvb2;
    float* tape;
};

// fused sweep tile (32 x 8, one row per thread)
constexpr int FT_X = 32, FT_Y = 8;

inline int64_t adj_ws_floats(int64_t n)
{
    // tp 2, gp 2, gq 4, um 3, vm 6, qmsg 4, tmsg 4, gabc 3, geta 3, gderiv 2, gvbB 2, gubB 1, guf 1, gtv 2,
    // gow 1, guth 1, gvb_fin 2, gIs 1, pos 1, the second copies gp 2, gq 4, qmsg 4, tmsg 4, gu 1, gubB 1, gvbB 2
    const int ch[] = {2, 2, 4, 3, 6, 4, 4, 3, 3, 2, 2, 1, 1, 2, 1, 1, 2, 1, 1, 2, 4, 4, 4, 1, 1, 2};
    int64_t t = 0;
    for (int c : ch) t += al(c * n);
    return t;
}

inline AdjWs adj_ws_at(float* base, int64_t n)
{
    AdjWs a;
    float* c = base;
    auto take = [&](int ch) {
        float* r = c;
        c += al(ch * n);
        return r;
    };
    a.tp = take(2);
    a.gp = take(2);
    a.gq = take(4);
    a.um = take(3);
    a.vm = take(6);
    a.qmsg = take(4);
    a.tmsg = take(4);
    a.gabc = take(3);
    a.geta = take(3);
    a.gderiv = take(2);
    a.gvbB = take(2);
    a.gubB = take(1);
    a.guf = take(1);
    a.gtv = take(2);
    a.gow = take(1);
    a.guth = take(1);
    a.gvb_fin = take(2);
    a.gIs = take(1);
    a.pos = reinterpret_cast<int*>(take(1));
    a.gp2 = take(2);
    a.gq2 = take(4);
    a.qmsg2 = take(4);
    a.tmsg2 = take(4);
    a.gu2 = take(1);
    a.gub2 = take(1);
    a.gvb2 = take(2);
    a.tape = c;
    return a;
}

// updateDualVariables (Python/vfs.cu:326-419) reading (p, q) from one state and
// writing them to another; the per-pixel arithmetic is the shared dual_pixel, so
// a recomputed iteration reproduces the in-place kernel's values.
__global__ void __launch_bounds__(256) k_dual_oop(float2* __restrict__ p_out, float2* __restrict__ tp,
                                                  float4* __restrict__ q_out, const float2* __restrict__ p_in,
                                                  const float4* __restrict__ q_in, float sigma,
                                                  const float* __restrict__ u, const float2* __restrict__ v,
                                                  const float* __restrict__ abc, const uint8_t* __restrict__ mask,
                                                  int H, int W, float alpha0, float alpha1, float eta_p, float eta_q)
{
    const int X = blockIdx.x * blockDim.x + threadIdx.x;
    const int Y = blockIdx.y * blockDim.y + threadIdx.y;
    if (X >= W || Y >= H) return;
    const int idx = Y * W + X;
    if (mask[idx] == 0) {
        p_out[idx] = make_float2(0.0f, 0.0f);
        tp[idx] = make_float2(0.0f, 0.0f);
        q_out[idx] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const int iE = Y * W + clamp_idx(X + 1, W), iW = Y * W + clamp_idx(X - 1, W);
    const int iS = clamp_idx(Y + 1, H) * W + X, iN = clamp_idx(Y - 1, H) * W + X;
    const bool fx = mask[iE] != 0, fy = mask[iS] != 0;
    const int ix = fx ? iE : iW, iy = fy ? iS : iN;
    float2 pn, tpn;
    float4 qn;
    dual_pixel(u[idx], v[idx], u[ix], v[ix], fx, u[iy], v[iy], fy, abc[3 * idx + 0], abc[3 * idx + 1],
               abc[3 * idx + 2], p_in[idx], q_in[idx], dual_scale(alpha1, sigma, eta_p),
               dual_scale(alpha0, sigma, eta_q), pn, tpn, qn);
    p_out[idx] = pn;
    tp[idx] = tpn;
    q_out[idx] = qn;
}

// Per-warp reset + warp-start record: p = q = v_bar = 0, u0 = u, and the saved
// copies of u, u_bar and flow.
__global__ void __launch_bounds__(256) k_warp_reset_save(float2* __restrict__ p, float4* __restrict__ q,
                                                         float2* __restrict__ v_bar, float* __restrict__ u0,
                                                         const float* __restrict__ u, const float* __restrict__ u_bar,
                                                         const float2* __restrict__ flow, float* __restrict__ su,
                                                         float* __restrict__ sub, float2* __restrict__ sflow, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    p[i] = make_float2(0.0f, 0.0f);
    v_bar[i] = make_float2(0.0f, 0.0f);
    q[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float uu = u[i];
    u0[i] = uu;
    su[i] = uu;
    sub[i] = u_bar[i];
    sflow[i] = flow[i];
}

// Checkpoint of the primal-dual state; with zero_pqv (the warp start of the blocked
// path, whose kernel reads no p, q, v_bar there) p, q, v_bar are stored as zeros.
__global__ void __launch_bounds__(256) k_save_state(float2* __restrict__ sp, float4* __restrict__ sq,
                                                    float* __restrict__ su, float* __restrict__ sub,
                                                    float2* __restrict__ svb, const float2* __restrict__ p,
                                                    const float4* __restrict__ q, const float* __restrict__ u,
                                                    const float* __restrict__ ub, const float2* __restrict__ vb, int n,
                                                    int zero_pqv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    sp[i] = zero_pqv ? make_float2(0.0f, 0.0f) : p[i];
    sq[i] = zero_pqv ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : q[i];
    su[i] = u[i];
    sub[i] = ub[i];
    svb[i] = zero_pqv ? make_float2(0.0f, 0.0f) : vb[i];
}

// Warp-start record of the blocked path and its first checkpoint: u, u_bar and flow, and
// the state before iteration 0 (p = q = v_bar = 0, u, u_bar).
__global__ void __launch_bounds__(256) k_warp_start_save(float* __restrict__ su, float* __restrict__ sub,
                                                         float2* __restrict__ sflow, float2* __restrict__ cp,
                                                         float4* __restrict__ cq, float* __restrict__ cu,
                                                         float* __restrict__ cub, float2* __restrict__ cvb,
                                                         const float* __restrict__ u, const float* __restrict__ u_bar,
                                                         const float2* __restrict__ flow, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float ui = u[i], ubi = u_bar[i];
    su[i] = ui;
    sub[i] = ubi;
    sflow[i] = flow[i];
    cp[i] = make_float2(0.0f, 0.0f);
    cq[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    cu[i] = ui;
    cub[i] = ubi;
    cvb[i] = make_float2(0.0f, 0.0f);
}

// computeOpticalFlowWithClamp with u0 = u in place (the blocked path keeps the
// warp-start u in u; the thresholded iterate lives elsewhere).
__global__ void __launch_bounds__(256) k_flow_clamp_inplace(float2* __restrict__ flow, float* __restrict__ u,
                                                            const float* __restrict__ uf,
                                                            const float2* __restrict__ tv,
                                                            const uint8_t* __restrict__ mask, int n, float limit)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || mask[i] == 0) return;
    const float base = u[i];
    const float delta = fminf(uf[i] - base, limit);
    u[i] = base + delta;
    const float2 t = tv[i];
    float2 f = flow[i];
    f.x = fmaf(delta, t.x, f.x);
    f.y = fmaf(delta, t.y, f.y);
    flow[i] = f;
}

// Same as the solver's clamp: flow += delta tv, u = u0 + delta (Python/vfs.cu:641-676,
// Python/vfs.py:519-521).
__global__ void __launch_bounds__(256) k_flow_clamp_u(float2* __restrict__ flow, float* __restrict__ u,
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

__global__ void __launch_bounds__(256) k_add(float* __restrict__ out, const float* __restrict__ in, int64_t n)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i < n) out[i] += in[i];
}

// A warp's reverse sweep starts from gubB := gub (the adjoint of u_bar at the warp's
// end) and zero v_bar, p, q adjoints, derivative and warped-image accumulators.
__global__ void __launch_bounds__(256) k_warp_reset(float* __restrict__ gubB, const float* __restrict__ gub,
                                                    float2* __restrict__ gvbB, float2* __restrict__ gp,
                                                    float4* __restrict__ gq, float2* __restrict__ gderiv,
                                                    float* __restrict__ gow, int64_t n)
{
    const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
    if (i >= n) return;
    const float2 z2 = make_float2(0.0f, 0.0f);
    gubB[i] = gub[i];
    gvbB[i] = z2;
    gp[i] = z2;
    gq[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    gderiv[i] = z2;
    gow[i] = 0.0f;
}

// One step of the fused reverse sweep over a tile of 32 x 8R pixels (256 threads,
// R rows each).  With S_i the state before iteration i (S_0 the warp start),
// launch F(i) runs
//   DO_D:   dual^T of iteration i (reads S_i: p, q, u_bar, v_bar) for the tile and a
//           one-pixel ring, its u_bar / v_bar messages kept in LDS;
//   DO_PL:  primal^T and threshold^T of iteration i - 1 (reads S_i: p, q and
//           S_{i-1}: u, u_bar), collecting the messages from LDS;
//   FINISH: instead of DO_PL, the adjoint of u_bar at the loop start (gub_out).
// The head of a sweep is DO_PL alone (iteration n - 1, no messages yet), the tail
// DO_D + FINISH (iteration 0).  Per-pixel arithmetic is that of the one-kernel-
// per-step path; the primal^T / threshold^T messages travel in compact form
// (qmsg = (-g_S1, -g_S2), tmsg = g_g).  Each pixel adds its lambda / alpha shares
// to its own slots of gpar in the order of the reversed steps, as the one-kernel-
// per-step path does, so the two sweeps give the same bits.
struct FusedArgs {
    float* gu;
    float* gubB;
    float2* gvbB;
    const float2* gp;  // read (own pixel and, for the ring, the neighbouring tile's)
    const float4* gq;
    const float2* qmsg;
    const float* tmsg;
    float2* gp_out;  // written (own pixel)
    float4* gq_out;
    float2* qmsg_out;
    float* tmsg_out;
    float* gabc;
    float* geta;
    float2* gderiv;
    float* gpar;  // (H, W, 5) per-pixel parameter shares
    float* gub_out;
    const float2* p1;  // S_i
    const float4* q1;
    const float* ub1;
    const float2* vb1;
    const float* u2;  // S_{i-1}
    const float* ub2;
    const float* abc;
    const float* eta;
    const float2* deriv;
    const uint8_t* mask;
    int H, W;
    int nbx, nby;         // tiles
    DualSteps st;         // iteration i (dual_steps of its sigma)
    float tau_pl, mu_pl;  // iteration i - 1 (the head: n - 1)
    float alpha0, alpha1, eta_p, eta_q, lambda;
};

// dual^T at pixel (X, Y) (in the image).  gab: adjoint of abc on entry (the owner's
// accumulator, or zeros for a ring pixel), updated.
__device__ __forceinline__ void fused_dual(const FusedArgs& a, int X, int Y, float2& gpo, float4& gqo, DualMsg& msg,
                                           float (&gab)[3], float& gpa0, float& gpa1)
{
    const int H = a.H, W = a.W, idx = Y * W + X;
    const Nbr n = nbr(a.mask, X, Y, H, W);
    if (!n.m) {
        gpo = make_float2(0.0f, 0.0f);
        gqo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 3; ++k) msg.um[k] = 0.0f;
        for (int k = 0; k < 6; ++k) msg.vm[k] = 0.0f;
        return;
    }
    float4 gqn = a.gq[idx];
    float2 gtp;
    collect_qt_compact(a.qmsg, a.tmsg, n, X, Y, H, W, idx, gqn, gtp);
    const bool fx = n.mE, fy = n.mS;
    const int ix = fx ? n.iE : n.iW, iy = fy ? n.iS : n.iN;
    dual_adj_pixel(fx, fy, a.ub1[idx], a.vb1[idx], a.ub1[ix], a.vb1[ix], a.ub1[iy], a.vb1[iy], a.abc[3 * idx],
                   a.abc[3 * idx + 1], a.abc[3 * idx + 2], a.p1[idx], a.q1[idx], a.st, a.gp[idx], gqn, gtp, gpo, gqo,
                   msg, gab, gpa0, gpa1);
}

template <int LX>
struct LdsMsgs {
    const float* m;  // rows x LX x 9
    int X0, Y0;
    __device__ __forceinline__ DualMsg operator()(int x, int y) const
    {
        const float* r = m + 9 * ((y - Y0 + 1) * LX + (x - X0 + 1));
        DualMsg d;
        for (int k = 0; k < 3; ++k) d.um[k] = r[k];
        for (int k = 0; k < 6; ++k) d.vm[k] = r[3 + k];
        return d;
    }
};

__device__ __forceinline__ void put_msg(float* m, int l, const DualMsg& d)
{
    float* r = m + 9 * l;
    for (int k = 0; k < 3; ++k) r[k] = d.um[k];
    for (int k = 0; k < 6; ++k) r[3 + k] = d.vm[k];
}

// Tile of a block, in the XCD-aware order of xcd_item.
__device__ __forceinline__ void tile_of(int b, int nbx, int nby, int& bx, int& by)
{
    const int t = xcd_item(b, nbx * nby);
    bx = t % nbx;
    by = t / nbx;
}

template <int R, bool DO_D, bool DO_PL, bool FINISH>
__global__ void __launch_bounds__(256) k_adj_fused(const FusedArgs a)
{
    constexpr int TX = FT_X, TY = 8 * R, LX = TX + 2, LY = TY + 2, RING = 2 * LX + 2 * TY;
    __shared__ float s_msg[DO_D ? LY * LX * 9 : 1];
    const int tid = threadIdx.x, tx = tid & (TX - 1), ty = tid / TX;
    int bx, by;
    tile_of(blockIdx.x, a.nbx, a.nby, bx, by);
    const int X0 = bx * TX, Y0 = by * TY;
    const int H = a.H, W = a.W;
    const int X = X0 + tx;
    float gpl[R], gpa0[R], gpa1[R];
    float2 gpo[R];
    float4 gqo[R];
    float gab[R][3];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int Y = Y0 + ty + 8 * r;
        gpo[r] = make_float2(0.0f, 0.0f);
        gqo[r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        for (int k = 0; k < 3; ++k) gab[r][k] = 0.0f;
        gpl[r] = gpa0[r] = gpa1[r] = 0.0f;
        if (X < W && Y < H) {
            const int idx = Y * W + X;
            for (int k = 0; k < 3; ++k) gab[r][k] = a.gabc[3 * idx + k];
            const float* gp = a.gpar + GP_N * idx;
            gpl[r] = gp[GP_LAMBDA];
            gpa0[r] = gp[GP_ALPHA0];
            gpa1[r] = gp[GP_ALPHA1];
        }
    }
    if constexpr (DO_D) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int Y = Y0 + ty + 8 * r;
            if (X < W && Y < H) {
                DualMsg m;
                fused_dual(a, X, Y, gpo[r], gqo[r], m, gab[r], gpa0[r], gpa1[r]);
                put_msg(s_msg, (ty + 8 * r + 1) * LX + tx + 1, m);
            }
        }
        for (int t = tid; t < RING; t += 256) {
            int lx, ly;
            if (t < LX) {
                lx = t;
                ly = 0;
            } else if (t < 2 * LX) {
                lx = t - LX;
                ly = LY - 1;
            } else if (t < 2 * LX + TY) {
                lx = 0;
                ly = 1 + t - 2 * LX;
            } else {
                lx = LX - 1;
                ly = 1 + t - 2 * LX - TY;
            }
            const int gx = X0 - 1 + lx, gy = Y0 - 1 + ly;
            if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
                float2 p_;
                float4 q_;
                float g_[3] = {0.0f, 0.0f, 0.0f};
                float d0 = 0.0f, d1 = 0.0f;
                DualMsg m;
                fused_dual(a, gx, gy, p_, q_, m, g_, d0, d1);
                put_msg(s_msg, ly * LX + lx, m);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int Y = Y0 + ty + 8 * r;
        if (X >= W || Y >= H) continue;
        const int idx = Y * W + X;
        if constexpr (DO_PL) {
            float g_ubo = a.gubB[idx];
            float2 g_vbo = a.gvbB[idx];
            if constexpr (DO_D) gather_msgs(a.mask, LdsMsgs<LX>{s_msg, X0, Y0}, X, Y, H, W, g_ubo, g_vbo);
            const Nbr n = nbr(a.mask, X, Y, H, W);
            if (!n.m) {
                a.gubB[idx] = 0.0f;
                a.gvbB[idx] = make_float2(0.0f, 0.0f);
                a.qmsg_out[idx] = make_float2(0.0f, 0.0f);
                a.gu[idx] = 0.0f;
                a.tmsg_out[idx] = 0.0f;
                a.gp_out[idx] = DO_D ? gpo[r] : a.gp[idx];
                a.gq_out[idx] = DO_D ? gqo[r] : a.gq[idx];
            } else {
                const float4 zero4 = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                float2 gp_ = DO_D ? gpo[r] : a.gp[idx];
                float4 gq_ = DO_D ? gqo[r] : a.gq[idx];
                float gu_ = a.gu[idx];
                float e0 = a.geta[3 * idx], e1 = a.geta[3 * idx + 1], e2 = a.geta[3 * idx + 2];
                float ubB, g_g;
                float2 vbB, qmc;
                float4 qm, tm;
                const float ea = a.abc[3 * idx], eb = a.abc[3 * idx + 1], ec = a.abc[3 * idx + 2];
                primal_adj_pixel(n, a.q1[idx], n.mW ? a.q1[n.iW] : zero4, n.mN ? a.q1[n.iN] : zero4, ea, eb, ec,
                                 a.p1[idx], a.eta[3 * idx + 1], a.eta[3 * idx + 2], a.tau_pl, a.mu_pl, a.alpha0,
                                 a.alpha1, g_ubo, g_vbo, ubB, vbB, gu_, gp_, gq_, qm, gab[r], e1, e2, gpa0[r], gpa1[r],
                                 qmc);
                float2 gd = a.gderiv[idx];
                gu_ = l1_adj_pixel(n, gu_, tp_at(a.p1, a.abc, a.mask, idx), tp_at(a.p1, a.abc, a.mask, n.iW).x,
                                   tp_at(a.p1, a.abc, a.mask, n.iN).y, a.deriv[idx], a.eta[3 * idx], a.u2[idx],
                                   a.ub2[idx], a.tau_pl, a.lambda, tm, ubB, e0, gd, gpl[r], g_g);
                a.gu[idx] = gu_;
                a.gubB[idx] = ubB;
                a.gvbB[idx] = vbB;
                a.gp_out[idx] = gp_;
                a.gq_out[idx] = gq_;
                a.qmsg_out[idx] = qmc;
                a.tmsg_out[idx] = g_g;
                a.geta[3 * idx] = e0;
                a.geta[3 * idx + 1] = e1;
                a.geta[3 * idx + 2] = e2;
                a.gderiv[idx] = gd;
                for (int k = 0; k < 3; ++k) a.gabc[3 * idx + k] = gab[r][k];
                float* gp = a.gpar + GP_N * idx;
                gp[GP_LAMBDA] = gpl[r];
                gp[GP_ALPHA0] = gpa0[r];
                gp[GP_ALPHA1] = gpa1[r];
            }
        } else {
            // DO_D alone: the dual^T outputs of the owned pixel
            a.gp_out[idx] = gpo[r];
            a.gq_out[idx] = gqo[r];
            for (int k = 0; k < 3; ++k) a.gabc[3 * idx + k] = gab[r][k];
            if (a.mask[idx] != 0) {
                float* gp = a.gpar + GP_N * idx;
                gp[GP_ALPHA0] = gpa0[r];
                gp[GP_ALPHA1] = gpa1[r];
            }
            if constexpr (FINISH) {
                float g = a.gubB[idx];
                float2 gv = a.gvbB[idx];
                gather_msgs(a.mask, LdsMsgs<LX>{s_msg, X0, Y0}, X, Y, H, W, g, gv);
                a.gub_out[idx] = g;
            }
        }
    }
}

// Rows per thread of the fused sweep at a level of n pixels.  Measured on MI355X at
// 800 x 800 / 400 x 400 / 200 x 200: one row (the most resident waves) is the
// fastest at every level; the longer tiles trade occupancy for less ring work
// and lose.
inline int fused_rows(int64_t) { return 1; }

inline unsigned nb1(int64_t n) { return static_cast<unsigned>((n + 255) / 256); }

// kind 0: head (PL), 1: D + PL, 2: tail (D + finish)
template <int R>
int launch_fused(int kind, int nblocks, hipStream_t s, const FusedArgs& fa)
{
    if (kind == 0)
        hipLaunchKernelGGL((k_adj_fused<R, false, true, false>), dim3(nblocks), dim3(256), 0, s, fa);
    else if (kind == 1)
        hipLaunchKernelGGL((k_adj_fused<R, true, true, false>), dim3(nblocks), dim3(256), 0, s, fa);
    else
        hipLaunchKernelGGL((k_adj_fused<R, true, false, true>), dim3(nblocks), dim3(256), 0, s, fa);
    return hipGetLastError() == hipSuccess ? 0 : VFS_EINVAL;
}

int launch_fused_rows(int rows, int kind, int nblocks, hipStream_t s, const FusedArgs& fa)
{
    if (rows == 4) return launch_fused<4>(kind, nblocks, s, fa);
    if (rows == 2) return launch_fused<2>(kind, nblocks, s, fa);
    return launch_fused<1>(kind, nblocks, s, fa);
}

// ---------------------------------------------------------------------------
// Temporally blocked reverse sweep: up to KR consecutive units F(i) = dual^T(i) +
// primal^T / threshold^T(i - 1) per launch, on a tile extended by KR pixels on every
// side.
//
// Per unit the region on which the adjoint state is exact shrinks by one pixel on
// every side.  dual^T at x reads the messages of x and x + 1 (E; S likewise), and
// primal^T at x gathers the dual^T messages of x - 1 (whose dual step read x as its
// east neighbour) and of x + 1 only when x + 1 read x as its west neighbour, i.e.
// when x + 2 is outside the mask; the messages of such a pixel are zero, so that
// sender is exact wherever x + 1 is.  A neighbour beyond the loaded region
// contributes nothing (or is the pixel itself, for the forward state read): the
// pixels within one of the region's border are inexact after the first unit, and
// the inexact ring grows by one per unit, so after KR units the tile itself is
// exact.  The adjoint state (gu, g u_bar,
// g v_bar, gp, gq) and the accumulators of abc, eta and the derivatives stay in
// registers; the messages and, per unit, the neighbours' forward state go through LDS
// (each pixel loads its own state from the tape).  Per-pixel arithmetic is the
// fused kernel's (vfs_adj_pixel.h), in the same order, and each tile pixel keeps its
// (lambda, alpha0, alpha1) shares in registers that start from and return to its
// gpar slots, so the tile's values equal the fused sweep's bit for bit.
// ---------------------------------------------------------------------------

constexpr int KRMAX = 6;

struct BlockRevArgs {
    // adjoint state read (one set) and written (the other): every pixel of the level
    const float* gu_in;
    const float* gub_in;
    const float2* gvb_in;
    const float2* gp_in;
    const float4* gq_in;
    const float2* qm_in;
    const float* tm_in;
    float* gu_out;
    float* gub_out;
    float2* gvb_out;
    float2* gp_out;
    float4* gq_out;
    float2* qm_out;
    float* tm_out;
    // accumulators, read and written at the tile's own pixels inside the mask
    float* gabc;
    float* geta;
    float2* gderiv;
    float* gpar;  // (H, W, 5) per-pixel parameter shares
    const float* abc;
    const float* eta;
    const float2* deriv;
    const uint8_t* mask;
    // unit u reverses iteration i = i0 - u: forward states S_i and S_{i-1}
    const float2* p1[KRMAX];
    const float4* q1[KRMAX];
    const float* ub1[KRMAX];
    const float2* vb1[KRMAX];
    const float* u2[KRMAX];
    const float* ub2[KRMAX];
    DualSteps st[KRMAX];  // dual_steps of iteration i's sigma
    float tau_pl[KRMAX], mu_pl[KRMAX];
    int n_units;
    int H, W, tiles_x, tiles_y;
    float alpha0, alpha1, eta_p, eta_q, lambda;
};

enum : unsigned {
    RB_IMG = 1u,    // inside the image
    RB_M = 2u,      // inside the mask
    RB_ME = 4u,     // mask at the east neighbour (the pixel itself at the image border)
    RB_MW = 8u,
    RB_MS = 16u,
    RB_MN = 32u,
    RB_TXP = 64u,   // dual x-target: east neighbour (RB_TXM: west; neither: the pixel itself)
    RB_TXM = 128u,
    RB_TYP = 256u,  // dual y-target: south neighbour (RB_TYM: north)
    RB_TYM = 512u,
    RB_TILE = 1024u,
};

// A pixel's forward state read by unit u from the tape: S_i (u_bar, v_bar, and inside
// the mask p, q) and, inside the mask, u and u_bar of S_{i-1}.  Zero where not read.
struct UnitTape {
    float ub1, vb1x, vb1y, p1x, p1y, q1x, q1y, q1z, q1w, u2, ub2;
};

__device__ __forceinline__ UnitTape load_unit_tape(const BlockRevArgs& a, int u, unsigned f, int idx)
{
    UnitTape t;
    t.ub1 = t.vb1x = t.vb1y = t.p1x = t.p1y = t.q1x = t.q1y = t.q1z = t.q1w = t.u2 = t.ub2 = 0.0f;
    if (f & RB_IMG) {
        t.ub1 = a.ub1[u][idx];
        const float2 vb = a.vb1[u][idx];
        t.vb1x = vb.x;
        t.vb1y = vb.y;
        if (f & RB_M) {
            const float2 p = a.p1[u][idx];
            const float4 q = a.q1[u][idx];
            t.p1x = p.x;
            t.p1y = p.y;
            t.q1x = q.x;
            t.q1y = q.y;
            t.q1z = q.z;
            t.q1w = q.w;
            t.u2 = a.u2[u][idx];
            t.ub2 = a.ub2[u][idx];
        }
    }
    return t;
}

// LEAN: one buffer of the units' forward state in LDS instead of two (a third barrier
// per unit orders its rewrite after the previous unit's reads) and each unit's tape
// loaded at the unit's start instead of during the previous unit: 22 instead of 31
// floats of LDS per extended pixel and fewer registers, so that three 512-thread
// workgroups fit on a CU instead of two.
template <int TX, int TY, int KR, int NT, bool LEAN = false>
__global__ void __launch_bounds__(NT) k_adj_block(const BlockRevArgs a)
{
    constexpr int EX = TX + 2 * KR, EY = TY + 2 * KR, E = EX * EY, S = (E + NT - 1) / NT;
    __shared__ float sQM[2 * E];     // primal^T messages (-g_S1, -g_S2), compact
    __shared__ float sTM[E];         // threshold^T messages g_g
    __shared__ float sDM[9 * E];     // dual^T messages (DualMsg um[3], vm[6])
    // the unit's forward state S_i as its neighbours read it: T p (2), u_bar (1), v_bar (2),
    // q (4); two buffers, alternating between units
    __shared__ float sFS[LEAN ? 1 : 2][9 * E];
    __shared__ unsigned sFL[E];
    const int tid = threadIdx.x;
    const int H = a.H, W = a.W;
    const int t = xcd_item(blockIdx.x, gridDim.x);
    const int btx = t % a.tiles_x, bty = t / a.tiles_x;
    const int ox = btx * TX - KR, oy = bty * TY - KR;

    float gu[S], gub[S], gvx[S], gvy[S], gpx[S], gpy[S], gqx[S], gqy[S], gqz[S], gqw[S];
    float ab[S][3], et[S][3], dd[S][2];   // accumulators (tile pixels)
    float ca[S], cb[S], cc[S], e0[S], e1[S], e2[S], iu[S], iz[S];
    float gpl[S], gpa0[S], gpa1[S];       // parameter shares (tile pixels)
    int gi[S];
    unsigned fl[S];

#pragma unroll
    for (int j = 0; j < S; ++j) {
        const int e = tid + NT * j;
        const int ex = e % EX, ey = e / EX, gx = ox + ex, gy = oy + ey;
        unsigned f = 0;
        gi[j] = 0;
        gu[j] = gub[j] = gvx[j] = gvy[j] = gpx[j] = gpy[j] = gqx[j] = gqy[j] = gqz[j] = gqw[j] = 0.0f;
        ca[j] = cb[j] = cc[j] = e0[j] = e1[j] = e2[j] = iu[j] = iz[j] = 0.0f;
        for (int c = 0; c < 3; ++c) ab[j][c] = et[j][c] = 0.0f;
        dd[j][0] = dd[j][1] = 0.0f;
        gpl[j] = gpa0[j] = gpa1[j] = 0.0f;
        float qm0 = 0.0f, qm1 = 0.0f, tm0 = 0.0f;
        if (e < E && gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const int idx = gy * W + gx;
            gi[j] = idx;
            f |= RB_IMG;
            if (ex >= KR && ex < KR + TX && ey >= KR && ey < KR + TY) f |= RB_TILE;
            // the pixel's and its four neighbours' mask bytes in one round trip to memory (the
            // empty asm takes them, so that the neighbours' loads are not put behind the test of
            // the pixel's own)
            const unsigned m0 = a.mask[idx], mEb = a.mask[gy * W + clamp_idx(gx + 1, W)];
            const unsigned mSb = a.mask[clamp_idx(gy + 1, H) * W + gx], mWb = a.mask[gy * W + clamp_idx(gx - 1, W)];
            const unsigned mNb = a.mask[clamp_idx(gy - 1, H) * W + gx];
            asm volatile("" ::"v"(m0), "v"(mEb), "v"(mSb), "v"(mWb), "v"(mNb));
            if (m0 != 0) {
                f |= RB_M;
                const bool mE = mEb != 0;
                const bool mS = mSb != 0;
                if (mE) f |= RB_ME;
                if (mWb != 0) f |= RB_MW;
                if (mS) f |= RB_MS;
                if (mNb != 0) f |= RB_MN;
                // the neighbour whose u_bar, v_bar this pixel's dual step read (dual_x_target)
                if (mE ? gx + 1 < W : gx > 0) f |= mE ? RB_TXP : RB_TXM;
                if (mS ? gy + 1 < H : gy > 0) f |= mS ? RB_TYP : RB_TYM;
                ca[j] = a.abc[3 * idx];
                cb[j] = a.abc[3 * idx + 1];
                cc[j] = a.abc[3 * idx + 2];
                e0[j] = a.eta[3 * idx];
                e1[j] = a.eta[3 * idx + 1];
                e2[j] = a.eta[3 * idx + 2];
                const float2 d = a.deriv[idx];
                iu[j] = d.x;
                iz[j] = d.y;
                if (f & RB_TILE) {
                    for (int c = 0; c < 3; ++c) {
                        ab[j][c] = a.gabc[3 * idx + c];
                        et[j][c] = a.geta[3 * idx + c];
                    }
                    const float2 g = a.gderiv[idx];
                    dd[j][0] = g.x;
                    dd[j][1] = g.y;
                    const float* gp = a.gpar + GP_N * idx;
                    gpl[j] = gp[GP_LAMBDA];
                    gpa0[j] = gp[GP_ALPHA0];
                    gpa1[j] = gp[GP_ALPHA1];
                }
            }
            gu[j] = a.gu_in[idx];
            gub[j] = a.gub_in[idx];
            const float2 gv = a.gvb_in[idx];
            gvx[j] = gv.x;
            gvy[j] = gv.y;
            const float2 gp = a.gp_in[idx];
            gpx[j] = gp.x;
            gpy[j] = gp.y;
            const float4 gq = a.gq_in[idx];
            gqx[j] = gq.x;
            gqy[j] = gq.y;
            gqz[j] = gq.z;
            gqw[j] = gq.w;
            const float2 qm = a.qm_in[idx];
            qm0 = qm.x;
            qm1 = qm.y;
            tm0 = a.tm_in[idx];
        }
        fl[j] = f;
        if (e < E) {
            sQM[e] = qm0;
            sQM[E + e] = qm1;
            sTM[e] = tm0;
            sFL[e] = f;
        }
    }

    // Outside the mask every unit leaves the adjoint state zero (dual^T and primal^T
    // produce nothing there), so those lanes' registers are zeroed once here and not
    // touched again; only their messages in LDS, which neighbours read as loaded in the
    // first unit, are cleared by primal^T below.
#pragma unroll
    for (int j = 0; j < S; ++j)
        if (!(fl[j] & RB_M)) gu[j] = gub[j] = gvx[j] = gvy[j] = gpx[j] = gpy[j] = gqx[j] = gqy[j] = gqz[j] = gqw[j] = 0.0f;

    // The tape of unit u + 1 is loaded while unit u runs (the loads stay in flight
    // across the unit's barriers), so its latency is not on the unit's path.
    UnitTape nxt[S];
#pragma unroll
    for (int j = 0; j < S; ++j) nxt[j] = load_unit_tape(a, 0, fl[j], gi[j]);
    for (int u = 0; u < a.n_units; ++u) {
        float* fsb = sFS[LEAN ? 0 : (u & 1)];
        if (LEAN && u > 0) {
#pragma unroll
            for (int j = 0; j < S; ++j) nxt[j] = load_unit_tape(a, u, fl[j], gi[j]);
            __syncthreads();  // the previous unit's reads of the forward state are done
        }
        // own forward state of this unit
        float p1x[S], p1y[S], q1x[S], q1y[S], q1z[S], q1w[S], ub1[S], vb1x[S], vb1y[S], u2[S], ub2[S];
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const UnitTape t = nxt[j];
            p1x[j] = t.p1x;
            p1y[j] = t.p1y;
            q1x[j] = t.q1x;
            q1y[j] = t.q1y;
            q1z[j] = t.q1z;
            q1w[j] = t.q1w;
            ub1[j] = t.ub1;
            vb1x[j] = t.vb1x;
            vb1y[j] = t.vb1y;
            u2[j] = t.u2;
            ub2[j] = t.ub2;
        }
        if (!LEAN && u + 1 < a.n_units) {
#pragma unroll
            for (int j = 0; j < S; ++j) nxt[j] = load_unit_tape(a, u + 1, fl[j], gi[j]);
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int e = tid + NT * j;
            float tx_ = 0.0f, ty_ = 0.0f;
            if (fl[j] & RB_M) {
                tx_ = fmaf(ca[j], p1x[j], cc[j] * p1y[j]);  // tp_at
                ty_ = fmaf(cc[j], p1x[j], cb[j] * p1y[j]);
            }
            if (e < E) {
                fsb[e] = tx_;
                fsb[E + e] = ty_;
                fsb[2 * E + e] = ub1[j];
                fsb[3 * E + e] = vb1x[j];
                fsb[4 * E + e] = vb1y[j];
                fsb[5 * E + e] = q1x[j];
                fsb[6 * E + e] = q1y[j];
                fsb[7 * E + e] = q1z[j];
                fsb[8 * E + e] = q1w[j];
            }
        }
        __syncthreads();  // messages of the previous unit, T p of this one, flags
        // dual^T of iteration i
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int e = tid + NT * j;
            if (e >= E) continue;
            const unsigned f = fl[j];
            // outside the mask: gp, gq stay zero, and no pixel reads this lane's dual
            // messages (primal^T gathers only from senders inside the mask)
            if (!(f & RB_M)) continue;
            DualMsg m;
            float2 gpo = make_float2(0.0f, 0.0f);
            float4 gqo = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (f & RB_M) {
                const int ex = e % EX, ey = e / EX, gx = ox + ex, gy = oy + ey;
                // collect_qt_compact with the messages in LDS (a sender beyond the loaded
                // region contributes nothing: such a pixel is inexact after this unit)
                float4 gqn = make_float4(gqx[j], gqy[j], gqz[j], gqw[j]);
                const float q0x = sQM[e], q0y = sQM[E + e], g0 = sTM[e];
                float2 gtp = make_float2((f & RB_ME) ? g0 : 0.0f, (f & RB_MS) ? g0 : 0.0f);
                if (gx == 0) {
                    gqn.x += q0x;
                    gqn.w += q0y;
                    gtp.x += -g0;
                }
                if (gx + 1 < W && ex + 1 < EX) {
                    gqn.x += sQM[e + 1];
                    gqn.w += sQM[E + e + 1];
                    gtp.x += -sTM[e + 1];
                }
                if (gy == 0) {
                    gqn.y += q0y;
                    gqn.z += q0x;
                    gtp.y += -g0;
                }
                if (gy + 1 < H && ey + 1 < EY) {
                    gqn.y += sQM[E + e + EX];
                    gqn.z += sQM[e + EX];
                    gtp.y += -sTM[e + EX];
                }
                const bool fx = (f & RB_ME) != 0, fy = (f & RB_MS) != 0;
                // the neighbours read (the pixel itself at the image border, clamped; and at the
                // region's own border, where the result is inexact after this unit in any case)
                const int eE = (gx + 1 < W && ex + 1 < EX) ? e + 1 : e, eW = (gx > 0 && ex > 0) ? e - 1 : e;
                const int eS = (gy + 1 < H && ey + 1 < EY) ? e + EX : e, eN = (gy > 0 && ey > 0) ? e - EX : e;
                const int ix = fx ? eE : eW, iy = fy ? eS : eN;
                // the per-pixel accumulators run at every pixel of the region (only the tile's
                // are stored)
                dual_adj_pixel(fx, fy, ub1[j], make_float2(vb1x[j], vb1y[j]), fsb[2 * E + ix],
                               make_float2(fsb[3 * E + ix], fsb[4 * E + ix]), fsb[2 * E + iy],
                               make_float2(fsb[3 * E + iy], fsb[4 * E + iy]), ca[j], cb[j], cc[j],
                               make_float2(p1x[j], p1y[j]),
                               make_float4(q1x[j], q1y[j], q1z[j], q1w[j]), a.st[u], make_float2(gpx[j], gpy[j]), gqn,
                               gtp, gpo, gqo, m, ab[j], gpa0[j], gpa1[j]);
            }
            gpx[j] = gpo.x;
            gpy[j] = gpo.y;
            gqx[j] = gqo.x;
            gqy[j] = gqo.y;
            gqz[j] = gqo.z;
            gqw[j] = gqo.w;
            for (int c = 0; c < 3; ++c) sDM[c * E + e] = m.um[c];
            for (int c = 0; c < 6; ++c) sDM[(3 + c) * E + e] = m.vm[c];
        }
        __syncthreads();  // dual^T messages
        // primal^T and threshold^T of iteration i - 1
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int e = tid + NT * j;
            if (e >= E) continue;
            const unsigned f = fl[j];
            if (!(f & RB_M)) {
                sQM[e] = sQM[E + e] = 0.0f;
                sTM[e] = 0.0f;
                continue;
            }
            const int ex = e % EX, ey = e / EX, gx = ox + ex, gy = oy + ey;
            // gather_msgs from LDS: own, then the x senders X-1, X, X+1, the y senders Y-1, Y, Y+1
            float g_ubo = gub[j];
            float2 g_vbo = make_float2(gvx[j], gvy[j]);
            g_ubo += sDM[e];
            g_vbo.x += sDM[3 * E + e];
            g_vbo.y += sDM[4 * E + e];
            for (int dx = -1; dx <= 1; ++dx) {
                const int xs = gx + dx, es = e + dx;
                if (xs < 0 || xs >= W || ex + dx < 0 || ex + dx >= EX) continue;
                const unsigned fs = sFL[es];
                const unsigned want = dx < 0 ? RB_TXP : (dx > 0 ? RB_TXM : 0u);
                if (!(fs & RB_M) || (fs & (RB_TXP | RB_TXM)) != want) continue;
                g_ubo += sDM[E + es];
                g_vbo.x += sDM[5 * E + es];
                g_vbo.y += sDM[6 * E + es];
            }
            for (int dy = -1; dy <= 1; ++dy) {
                const int ys = gy + dy, es = e + dy * EX;
                if (ys < 0 || ys >= H || ey + dy < 0 || ey + dy >= EY) continue;
                const unsigned fs = sFL[es];
                const unsigned want = dy < 0 ? RB_TYP : (dy > 0 ? RB_TYM : 0u);
                if (!(fs & RB_M) || (fs & (RB_TYP | RB_TYM)) != want) continue;
                g_ubo += sDM[2 * E + es];
                g_vbo.x += sDM[7 * E + es];
                g_vbo.y += sDM[8 * E + es];
            }
            Nbr n;
            n.iE = n.iW = n.iS = n.iN = gi[j];
            n.m = true;
            n.mE = (f & RB_ME) != 0;
            n.mW = (f & RB_MW) != 0;
            n.mS = (f & RB_MS) != 0;
            n.mN = (f & RB_MN) != 0;
            // W / N neighbours (the pixel itself at the image border, or beyond the region)
            const int eW = (gx > 0 && ex > 0) ? e - 1 : e, eN = (gy > 0 && ey > 0) ? e - EX : e;
            float2 gd = make_float2(dd[j][0], dd[j][1]);
            float ubB, g_g;
            float2 vbB, qmc;
            float4 qm, tm;
            float2 gp_ = make_float2(gpx[j], gpy[j]);
            float4 gq_ = make_float4(gqx[j], gqy[j], gqz[j], gqw[j]);
            float gu_ = gu[j];
            // read whether or not the neighbour is inside the mask: primal_adj_pixel uses
            // qW (qN) only when it is
            const float4 qW = make_float4(fsb[5 * E + eW], fsb[6 * E + eW], fsb[7 * E + eW], fsb[8 * E + eW]);
            const float4 qN = make_float4(fsb[5 * E + eN], fsb[6 * E + eN], fsb[7 * E + eN], fsb[8 * E + eN]);
            primal_adj_pixel(n, make_float4(q1x[j], q1y[j], q1z[j], q1w[j]), qW, qN, ca[j], cb[j], cc[j],
                             make_float2(p1x[j], p1y[j]), e1[j],
                             e2[j], a.tau_pl[u], a.mu_pl[u], a.alpha0, a.alpha1, g_ubo, g_vbo, ubB, vbB, gu_, gp_, gq_,
                             qm, ab[j], et[j][1], et[j][2], gpa0[j], gpa1[j], qmc);
            gu_ = l1_adj_pixel(n, gu_, make_float2(fsb[e], fsb[E + e]), fsb[eW], fsb[E + eN], make_float2(iu[j], iz[j]),
                               e0[j], u2[j], ub2[j], a.tau_pl[u], a.lambda, tm, ubB, et[j][0], gd,
                               gpl[j], g_g);
            dd[j][0] = gd.x;
            dd[j][1] = gd.y;
            gu[j] = gu_;
            gub[j] = ubB;
            gvx[j] = vbB.x;
            gvy[j] = vbB.y;
            gpx[j] = gp_.x;
            gpy[j] = gp_.y;
            gqx[j] = gq_.x;
            gqy[j] = gq_.y;
            gqz[j] = gq_.z;
            gqw[j] = gq_.w;
            sQM[e] = qmc.x;
            sQM[E + e] = qmc.y;
            sTM[e] = g_g;
        }
    }

    // the tile's state and accumulators
#pragma unroll
    for (int j = 0; j < S; ++j) {
        const unsigned f = fl[j];
        if (!(f & RB_TILE)) continue;
        const int e = tid + NT * j;
        const int idx = gi[j];
        a.gu_out[idx] = gu[j];
        a.gub_out[idx] = gub[j];
        a.gvb_out[idx] = make_float2(gvx[j], gvy[j]);
        a.gp_out[idx] = make_float2(gpx[j], gpy[j]);
        a.gq_out[idx] = make_float4(gqx[j], gqy[j], gqz[j], gqw[j]);
        a.qm_out[idx] = make_float2(sQM[e], sQM[E + e]);
        a.tm_out[idx] = sTM[e];
        if (f & RB_M) {
            for (int c = 0; c < 3; ++c) {
                a.gabc[3 * idx + c] = ab[j][c];
                a.geta[3 * idx + c] = et[j][c];
            }
            a.gderiv[idx] = make_float2(dd[j][0], dd[j][1]);
            float* gp = a.gpar + GP_N * idx;
            gp[GP_LAMBDA] = gpl[j];
            gp[GP_ALPHA0] = gpa0[j];
            gp[GP_ALPHA1] = gpa1[j];
        }
    }
}

// Reverse variants (sweep modes): 20 + KR (1-6 units) with 1024 threads on a 32 x 32
// extended tile, 30 + KR (1-3 units) with 512 threads on 32 x 16, 40 + KR the LEAN
// form of the latter.
struct RevVariant {
    int mode, tx, ty, kr, nt;
};
constexpr RevVariant kRevVariants[] = {
    {21, 30, 30, 1, 1024}, {22, 28, 28, 2, 1024}, {23, 26, 26, 3, 1024}, {24, 24, 24, 4, 1024},
    {25, 22, 22, 5, 1024}, {26, 20, 20, 6, 1024}, {31, 30, 14, 1, 512},  {32, 28, 12, 2, 512},
    {33, 26, 10, 3, 512},  {41, 30, 14, 1, 512},  {42, 28, 12, 2, 512},  {43, 26, 10, 3, 512},
};

const RevVariant* find_rev_variant(int mode)
{
    for (const RevVariant& v : kRevVariants)
        if (v.mode == mode) return &v;
    return nullptr;
}

template <int TX, int TY, int KR, int NT, bool LEAN = false>
int launch_adj_block_t(hipStream_t s, int tiles, const BlockRevArgs& ba)
{
    hipLaunchKernelGGL((k_adj_block<TX, TY, KR, NT, LEAN>), dim3(tiles), dim3(NT), 0, s, ba);
    return hipGetLastError() == hipSuccess ? 0 : VFS_EINVAL;
}

int launch_adj_block(const RevVariant& v, hipStream_t s, int tiles, const BlockRevArgs& ba)
{
    switch (v.mode) {
        case 21: return launch_adj_block_t<30, 30, 1, 1024>(s, tiles, ba);
        case 22: return launch_adj_block_t<28, 28, 2, 1024>(s, tiles, ba);
        case 23: return launch_adj_block_t<26, 26, 3, 1024>(s, tiles, ba);
        case 24: return launch_adj_block_t<24, 24, 4, 1024>(s, tiles, ba);
        case 25: return launch_adj_block_t<22, 22, 5, 1024>(s, tiles, ba);
        case 26: return launch_adj_block_t<20, 20, 6, 1024>(s, tiles, ba);
        case 31: return launch_adj_block_t<30, 14, 1, 512>(s, tiles, ba);
        case 32: return launch_adj_block_t<28, 12, 2, 512>(s, tiles, ba);
        case 33: return launch_adj_block_t<26, 10, 3, 512>(s, tiles, ba);
        case 41: return launch_adj_block_t<30, 14, 1, 512, true>(s, tiles, ba);
        case 42: return launch_adj_block_t<28, 12, 2, 512, true>(s, tiles, ba);
        case 43: return launch_adj_block_t<26, 10, 3, 512, true>(s, tiles, ba);
        default: return VFS_EINVAL;
    }
}

#define VFS_TRY(expr)             \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != 0) return rc_; \
    } while (0)

int check_args(const vfs_level* lv, const vfs_params* prm, const float* taus, const float* sigmas, const float* mus,
               int n_iter, int n_warp, int ck)
{
    if (lv == nullptr || prm == nullptr || n_iter < 0 || n_warp < 0 || ck < 1) return VFS_EINVAL;
    VFS_CHECK_DIMS(lv->H, lv->W);
    if (n_iter > 0 && (taus == nullptr || sigmas == nullptr || mus == nullptr)) return VFS_EINVAL;
    if (prm->median_k != 3 && prm->median_k != 5) return VFS_EINVAL;
    const void* ptrs[] = {lv->ref, lv->other, lv->tvf, lv->mask, lv->flow, lv->u, lv->u_bar, lv->ref_smoothed,
                          lv->abc, lv->eta, lv->tv, lv->warped, lv->deriv, lv->p, lv->q, lv->tp, lv->v_bar,
                          lv->u0, lv->uf};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    return 0;
}

// State S_{i+1} from S_i (iteration i of the schedule), written out of place by the
// per-kernel steps.
int pd_step(void* st, const vfs_level* lv, const vfs_params* prm, const State& a, const State& b, float* tp, float tau,
            float sigma, float mu)
{
    const int H = lv->H, W = lv->W;
    hipLaunchKernelGGL(k_dual_oop, grid2d(H, W), block2d(), 0, S(st), reinterpret_cast<float2*>(b.p),
                       reinterpret_cast<float2*>(tp), reinterpret_cast<float4*>(b.q),
                       reinterpret_cast<const float2*>(a.p), reinterpret_cast<const float4*>(a.q), sigma, a.ub,
                       reinterpret_cast<const float2*>(a.vb), lv->abc, lv->mask, H, W, prm->alpha0, prm->alpha1,
                       prm->eta_p, prm->eta_q);
    if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
    VFS_TRY(vfs_l1_thresholding(st, b.u, tau, tp, a.u, a.ub, lv->deriv, lv->eta, lv->mask, H, W, prm));
    VFS_TRY(vfs_update_primal(st, b.ub, b.vb, tau, mu, b.p, b.q, b.u, a.ub, a.vb, lv->abc, lv->eta, lv->mask, H, W,
                              prm));
    return 0;
}

}  // namespace

extern "C" {

size_t vfs_adj_save_bytes(int H, int W, int n_warp, int n_iter, int ckpt)
{
    if (H <= 0 || W <= 0 || n_warp < 0 || n_iter < 0 || ckpt < 1) return 0;
    const int64_t n = static_cast<int64_t>(H) * W;
    return static_cast<size_t>(n_warp) * static_cast<size_t>(warp_save_floats(n, n_iter, ckpt)) * sizeof(float);
}

size_t vfs_adj_tape_bytes(int H, int W, int n_iter, int ckpt, int tape_warps)
{
    if (H <= 0 || W <= 0 || n_iter < 0 || ckpt < 1 || tape_warps < 0) return 0;
    const int64_t n = static_cast<int64_t>(H) * W;
    return static_cast<size_t>(tape_warps) * static_cast<size_t>(last_seg_len(n_iter, ckpt)) *
           static_cast<size_t>(state_floats(n)) * sizeof(float);
}

size_t vfs_adj_workspace_bytes(int H, int W, int ckpt)
{
    if (H <= 0 || W <= 0 || ckpt < 1) return 0;
    const int64_t n = static_cast<int64_t>(H) * W;
    return static_cast<size_t>(adj_ws_floats(n) +
                               static_cast<int64_t>(ckpt) * state_floats(n)) *
           sizeof(float);
}

int vfs_adj_forward_level(void* stream, const vfs_level* lv, const vfs_params* prm, const float* taus,
                          const float* sigmas, const float* mus, int n_iter, int n_warp, int ckpt, float* save,
                          float* workspace, const vfs_adj_opts* tiles)
{
    VFS_TRY(check_args(lv, prm, taus, sigmas, mus, n_iter, n_warp, ckpt));
    if (save == nullptr && n_warp > 0) return VFS_EINVAL;
    const int fk = tiles ? tiles->fwd_k : 0;
    if (fk > 0 && workspace == nullptr) return VFS_EINVAL;
    const int tw = tiles ? tiles->tape_warps : 0;
    // last-segment tapes are written by the blocked forward only
    if (tw < 0 || tw > n_warp || (tw > 0 && (tiles->tape_save == nullptr || fk <= 0))) return VFS_EINVAL;
    const int H = lv->H, W = lv->W;
    const int64_t n = static_cast<int64_t>(H) * W;
    hipStream_t s = S(stream);
    VFS_TRY(vfs_gaussian_masked(stream, lv->ref_smoothed, lv->ref, lv->mask, H, W));
    VFS_TRY(vfs_compute_tensor_eta(stream, lv->abc, lv->eta, lv->ref_smoothed, lv->mask, H, W, prm));
    if (fk > 0) {
        // blocked path: k-iteration launches between two state sets, with a launch
        // boundary at every checkpoint; set 1 is the first tape slot of the workspace
        const AdjWs A = adj_ws_at(workspace, n);
        const State t1 = state_at(A.tape, n);
        float* su[2] = {lv->u0, t1.u};
        float* sub[2] = {lv->u_bar, t1.ub};
        float* svb[2] = {lv->v_bar, t1.vb};
        float* sp[2] = {lv->p, t1.p};
        float* sq[2] = {lv->q, t1.q};
        int cur = 0;  // set holding u_bar
        // the first warp's warping vector and warped image; each warp's end forms the next's
        VFS_TRY(warp_prep_launch(s, lv->tv, lv->warped, lv->flow, lv->tvf, lv->other, H, W));
        for (int w = 0; w < n_warp; ++w) {
            const WarpSave ws = warp_save_at(save, n, n_iter, ckpt, w);
            // the warp's derivatives go to its save record, which the reverse reads again
            VFS_TRY(vfs_compute_derivatives(stream, ws.deriv, lv->ref, lv->warped, lv->tvf, H, W));
            const State c0 = state_at(ws.ckpt0, n);
            hipLaunchKernelGGL(k_warp_start_save, dim3(nb1(n)), dim3(256), 0, s, ws.u, ws.ub,
                               reinterpret_cast<float2*>(ws.flow), reinterpret_cast<float2*>(c0.p),
                               reinterpret_cast<float4*>(c0.q), c0.u, c0.ub, reinterpret_cast<float2*>(c0.vb), lv->u,
                               sub[cur], reinterpret_cast<const float2*>(lv->flow), static_cast<int>(n));
            if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
            int done = 0;
            bool first = true;
            while (done < n_iter) {
                if (done > 0 && done % ckpt == 0) {
                    const State c = state_at(ws.ckpt0 + static_cast<int64_t>(done / ckpt) * state_floats(n), n);
                    hipLaunchKernelGGL(k_save_state, dim3(nb1(n)), dim3(256), 0, s, reinterpret_cast<float2*>(c.p),
                                       reinterpret_cast<float4*>(c.q), c.u, c.ub, reinterpret_cast<float2*>(c.vb),
                                       reinterpret_cast<const float2*>(sp[cur]),
                                       reinterpret_cast<const float4*>(sq[cur]), first ? lv->u : su[cur], sub[cur],
                                       reinterpret_cast<const float2*>(svb[cur]), static_cast<int>(n), first ? 1 : 0);
                    if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
                }
                int kc = n_iter - done < fk ? n_iter - done : fk;
                const int to_ck = ckpt - done % ckpt;
                if (kc > to_ck) kc = to_ck;
                const int a_last = (n_ckpt(n_iter, ckpt) - 1) * ckpt;
                float* wt = warp_tape_at(tiles, n, n_iter, ckpt, n_warp, w);
                if (wt != nullptr && done >= a_last) {
                    // the last segment: the same launch, also writing the state after each
                    // iteration into the warp's tape (slot j: the state after iteration a_last + j)
                    PdTapeLaunch L;
                    L.H = H;
                    L.W = W;
                    L.mask = lv->mask;
                    L.abc = lv->abc;
                    L.eta = lv->eta;
                    L.deriv = ws.deriv;
                    const float* ins[5] = {sp[cur], sq[cur], first ? lv->u : su[cur], sub[cur], svb[cur]};
                    float* outs[5] = {sp[1 - cur], sq[1 - cur], su[1 - cur], sub[1 - cur], svb[1 - cur]};
                    for (int c = 0; c < 5; ++c) {
                        L.in[c] = ins[c];
                        L.out[c] = outs[c];
                    }
                    L.taus = taus + done;
                    L.sigmas = sigmas + done;
                    L.mus = mus + done;
                    L.n_iter = kc;
                    L.tile_x = tiles->fwd_tile_x;
                    L.tile_y = tiles->fwd_tile_y;
                    L.k = fk;
                    L.threads = tiles->fwd_threads;
                    L.tape = wt + static_cast<int64_t>(done - a_last) * state_floats(n);
                    L.tape_stride = state_floats(n);
                    L.tape_off[0] = al(2 * n);
                    L.tape_off[1] = al(2 * n) + al(4 * n);
                    L.tape_off[2] = L.tape_off[1] + al(n);
                    L.tape_off[3] = L.tape_off[2] + al(n);
                    L.tape_count = kc;
                    L.first = first ? 1 : 0;
                    VFS_TRY(pd_blocks_tape(s, L, prm));
                } else {
                    VFS_TRY(vfs_pd_blocks(stream, H, W, lv->mask, lv->abc, lv->eta, ws.deriv,
                                          first ? lv->u : su[cur], sub[cur], svb[cur], sp[cur], sq[cur], su[1 - cur],
                                          sub[1 - cur], svb[1 - cur], sp[1 - cur], sq[1 - cur], first ? 1 : 0,
                                          taus + done, sigmas + done, mus + done, kc, tiles->fwd_tile_x,
                                          tiles->fwd_tile_y, fk, tiles->fwd_threads, prm));
                }
                cur = 1 - cur;
                done += kc;
                first = false;
            }
            const float* uth = n_iter > 0 ? su[cur] : lv->u;
            // median, clamp, and the next warp's warping vector and warped image
            VFS_TRY(median_clamp_prep_launch(s, prm->median_k, lv->u, lv->flow, lv->tv, lv->warped, uth, lv->mask,
                                             lv->tvf, lv->other, H, W, prm->limit_range, w + 1 < n_warp ? 1 : 0));
        }
        if (cur != 0) VFS_TRY(vfs_copy(stream, lv->u_bar, sub[cur], n));
        return 0;
    }
    for (int w = 0; w < n_warp; ++w) {
        const WarpSave ws = warp_save_at(save, n, n_iter, ckpt, w);
        VFS_TRY(vfs_find_warping_vector(stream, lv->tv, lv->flow, H, W, lv->tvf, H, W));
        VFS_TRY(vfs_warp_image(stream, lv->warped, lv->flow, H, W, lv->other, H, W));
        VFS_TRY(vfs_compute_derivatives(stream, ws.deriv, lv->ref, lv->warped, lv->tvf, H, W));
        hipLaunchKernelGGL(k_warp_reset_save, dim3(nb1(n)), dim3(256), 0, s, reinterpret_cast<float2*>(lv->p),
                           reinterpret_cast<float4*>(lv->q), reinterpret_cast<float2*>(lv->v_bar), lv->u0, lv->u,
                           lv->u_bar, reinterpret_cast<const float2*>(lv->flow), ws.u, ws.ub,
                           reinterpret_cast<float2*>(ws.flow), static_cast<int>(n));
        if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
        for (int it = 0; it < n_iter; ++it) {
            if (it % ckpt == 0) {
                const State c = state_at(ws.ckpt0 + static_cast<int64_t>(it / ckpt) * state_floats(n), n);
                hipLaunchKernelGGL(k_save_state, dim3(nb1(n)), dim3(256), 0, s, reinterpret_cast<float2*>(c.p),
                                   reinterpret_cast<float4*>(c.q), c.u, c.ub, reinterpret_cast<float2*>(c.vb),
                                   reinterpret_cast<const float2*>(lv->p), reinterpret_cast<const float4*>(lv->q),
                                   lv->u, lv->u_bar, reinterpret_cast<const float2*>(lv->v_bar), static_cast<int>(n),
                                   0);
                if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
            }
            VFS_TRY(vfs_update_dual(stream, lv->p, lv->tp, lv->q, sigmas[it], lv->u_bar, lv->v_bar, lv->abc, lv->mask,
                                    H, W, prm));
            VFS_TRY(vfs_l1_thresholding(stream, lv->u, taus[it], lv->tp, lv->u, lv->u_bar, ws.deriv, lv->eta,
                                        lv->mask, H, W, prm));
            VFS_TRY(vfs_update_primal(stream, lv->u_bar, lv->v_bar, taus[it], mus[it], lv->p, lv->q, lv->u,
                                      lv->u_bar, lv->v_bar, lv->abc, lv->eta, lv->mask, H, W, prm));
        }
        VFS_TRY(vfs_median_filter(stream, lv->uf, lv->u, lv->mask, H, W, prm->median_k));
        hipLaunchKernelGGL(k_flow_clamp_u, dim3(nb1(n)), dim3(256), 0, s, reinterpret_cast<float2*>(lv->flow), lv->u,
                           lv->u0, lv->uf, reinterpret_cast<const float2*>(lv->tv), lv->mask, static_cast<int>(n),
                           prm->limit_range);
        if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
    }
    return 0;
}

int vfs_adj_reverse_level(void* stream, const vfs_level* lv, const vfs_params* prm, const float* taus,
                          const float* sigmas, const float* mus, int n_iter, int n_warp, int ckpt, const float* save,
                          float* workspace, float* gu, float* gub, float* gf, float* gref, float* gpar,
                          const vfs_adj_opts* tiles)
{
    const int mode = tiles ? tiles->sweep : 0;
    const int rtx = tiles ? tiles->rec_tile_x : 0, rty = tiles ? tiles->rec_tile_y : 0;
    const int rk = tiles ? tiles->rec_k : 0;
    if (rtx > 0 && rk < 1) return VFS_EINVAL;
    VFS_TRY(check_args(lv, prm, taus, sigmas, mus, n_iter, n_warp, ckpt));
    const void* ptrs[] = {workspace, gu, gub, gf, gref, gpar};
    for (const void* p : ptrs)
        if (p == nullptr) return VFS_EINVAL;
    if (save == nullptr && n_warp > 0) return VFS_EINVAL;
    const int tw = tiles ? tiles->tape_warps : 0;
    if (tw < 0 || tw > n_warp || (tw > 0 && tiles->tape_save == nullptr)) return VFS_EINVAL;
    const RevVariant* rv = find_rev_variant(mode);
    if (mode != 0 && mode != 1 && mode != 11 && mode != 12 && mode != 14 && rv == nullptr) return VFS_EINVAL;
    const int H = lv->H, W = lv->W;
    const int64_t n = static_cast<int64_t>(H) * W;
    hipStream_t s = S(stream);
    float* sv = const_cast<float*>(save);
    const int nbx = (W + FT_X - 1) / FT_X;
    const AdjWs A = adj_ws_at(workspace, n);
    const int nck = n_ckpt(n_iter, ckpt);
    VFS_TRY(vfs_fill(stream, A.gabc, 0.0f, 3 * n));
    VFS_TRY(vfs_fill(stream, A.geta, 0.0f, 3 * n));
    // blocked sweep tiles (rv)
    const int rtiles_x = rv ? (W + rv->tx - 1) / rv->tx : 0, rtiles_y = rv ? (H + rv->ty - 1) / rv->ty : 0;
    const int rtiles = rtiles_x * rtiles_y;
    FusedArgs fa;
    fa.gu = A.guth;
    fa.gubB = A.gubB;
    fa.gvbB = reinterpret_cast<float2*>(A.gvbB);
    fa.gabc = A.gabc;
    fa.geta = A.geta;
    fa.gderiv = reinterpret_cast<float2*>(A.gderiv);
    fa.gpar = gpar;
    fa.gub_out = gub;
    fa.abc = lv->abc;
    fa.eta = lv->eta;
    fa.deriv = reinterpret_cast<const float2*>(lv->deriv);
    fa.mask = lv->mask;
    fa.H = H;
    fa.W = W;
    fa.alpha0 = prm->alpha0;
    fa.alpha1 = prm->alpha1;
    fa.eta_p = prm->eta_p;
    fa.eta_q = prm->eta_q;
    fa.lambda = prm->lambda;
    const int rows = (mode == 0 || rv != nullptr) ? fused_rows(n) : mode - 10;
    const int tile_y = FT_Y * rows, fnby = (H + tile_y - 1) / tile_y, fblocks = nbx * fnby;
    fa.nbx = nbx;
    fa.nby = fnby;
    const bool fused = mode != 1;
    const int rw = tiles ? tiles->reverse_warps : 0;
    if (rw < 0) return VFS_EINVAL;
    const int w_end = (rw > 0 && rw < n_warp) ? n_warp - rw : 0;
    // optional timing probe of the blocked reverse launches
    const int probe_max = (tiles && tiles->probe_events) ? tiles->probe_max : 0;
    if (probe_max < 0 || (probe_max > 0 && (tiles->probe_units == nullptr || tiles->probe_count == nullptr)))
        return VFS_EINVAL;
    int probed = 0;
    if (probe_max > 0) *tiles->probe_count = 0;
    for (int w = n_warp - 1; w >= w_end; --w) {
        const WarpSave ws = warp_save_at(sv, n, n_iter, ckpt, w);
        // the warp's warping vector at its start flow; its derivatives are in the save record
        VFS_TRY(vfs_find_warping_vector(stream, lv->tv, ws.flow, H, W, lv->tvf, H, W));
        vfs_level lw = *lv;
        lw.deriv = ws.deriv;
        fa.deriv = reinterpret_cast<const float2*>(ws.deriv);
        // the forward's tape of this warp's last segment, if it kept one
        float* wt = warp_tape_at(tiles, n, n_iter, ckpt, n_warp, w);
        // state i of segment seg = [a, b]: the checkpoint for i == a, else tape slot i - a - 1
        // (of the warp's kept tape for the last segment, of the recompute tape otherwise)
        auto state_of = [&](int seg, int i) {
            const int a = seg * ckpt;
            if (i == a) return state_at(ws.ckpt0 + static_cast<int64_t>(seg) * state_floats(n), n);
            float* tape = (wt != nullptr && seg == nck - 1) ? wt : A.tape;
            return state_at(tape + static_cast<int64_t>(i - a - 1) * state_floats(n), n);
        };
        // states a + 1 .. b of segment seg into the tape: blocked launches of up to rk
        // iterations that also write their intermediate states, or per-kernel steps
        auto recompute = [&](int seg) {
            if (wt != nullptr && seg == nck - 1) return 0;  // kept by the forward
            const int a = seg * ckpt, b = (seg + 1) * ckpt < n_iter ? (seg + 1) * ckpt : n_iter;
            if (rtx <= 0) {
                for (int i = a; i < b; ++i)
                    VFS_TRY(pd_step(stream, &lw, prm, state_of(seg, i), state_of(seg, i + 1), A.tp, taus[i],
                                    sigmas[i], mus[i]));
                return 0;
            }
            for (int i = a; i < b;) {
                const int kc = b - i < rk ? b - i : rk;
                const State in = state_of(seg, i), out = state_of(seg, i + kc);
                PdTapeLaunch L;
                L.H = H;
                L.W = W;
                L.mask = lv->mask;
                L.abc = lv->abc;
                L.eta = lv->eta;
                L.deriv = ws.deriv;
                const float* ins[5] = {in.p, in.q, in.u, in.ub, in.vb};
                float* outs[5] = {out.p, out.q, out.u, out.ub, out.vb};
                for (int c = 0; c < 5; ++c) {
                    L.in[c] = ins[c];
                    L.out[c] = outs[c];
                }
                L.taus = taus + i;
                L.sigmas = sigmas + i;
                L.mus = mus + i;
                L.n_iter = kc;
                L.tile_x = rtx;
                L.tile_y = rty;
                L.k = rk;
                L.threads = tiles->rec_threads;
                L.tape = kc > 1 ? state_of(seg, i + 1).p : nullptr;
                L.tape_stride = state_floats(n);
                L.tape_off[0] = al(2 * n);
                L.tape_off[1] = al(2 * n) + al(4 * n);
                L.tape_off[2] = L.tape_off[1] + al(n);
                L.tape_off[3] = L.tape_off[2] + al(n);
                L.tape_count = kc - 1;
                VFS_TRY(pd_blocks_tape(s, L, prm));
                i += kc;
            }
            return 0;
        };
        const float* u_th = ws.u;
        if (nck > 0) {
            VFS_TRY(recompute(nck - 1));
            u_th = state_of(nck - 1, n_iter).u;
        }
        VFS_TRY(vfs_adj_clamp(stream, gu, A.guf, A.pos, A.gtv, gf, u_th, ws.u, lv->tv, lv->mask, H, W,
                              prm->limit_range, prm->median_k));
        VFS_TRY(vfs_adj_median(stream, A.guth, A.guf, A.pos, H, W, prm->median_k));
        hipLaunchKernelGGL(k_warp_reset, dim3(nb1(n)), dim3(256), 0, s, A.gubB, gub, reinterpret_cast<float2*>(A.gvbB),
                           reinterpret_cast<float2*>(A.gp), reinterpret_cast<float4*>(A.gq),
                           reinterpret_cast<float2*>(A.gderiv), A.gow, n);
        if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
        // the adjoint of the warp-start u after the sweep (the blocked sweep may end in its
        // second buffer set, which the tail then reads in place)
        float* fin_gu = A.guth;
        if (nck > 0 && fused) {
            // fused sweep: head PL(n-1), F(i) for i = n-1 .. 1, tail D(0) + finish
            int cs = nck - 1;
            // ping-pong of (gp, gq, qmsg, tmsg): set 0 is A.gp ..., set 1 the second copies;
            // the head reads set 0 (gp, gq cleared above)
            int cur = 0;
            auto set_bufs = [&]() {
                float* gpi[2] = {A.gp, A.gp2};
                float* gqi[2] = {A.gq, A.gq2};
                float* qmi[2] = {A.qmsg, A.qmsg2};
                float* tmi[2] = {A.tmsg, A.tmsg2};
                fa.gp = reinterpret_cast<const float2*>(gpi[cur]);
                fa.gq = reinterpret_cast<const float4*>(gqi[cur]);
                fa.qmsg = reinterpret_cast<const float2*>(qmi[cur]);
                fa.tmsg = tmi[cur];
                fa.gp_out = reinterpret_cast<float2*>(gpi[1 - cur]);
                fa.gq_out = reinterpret_cast<float4*>(gqi[1 - cur]);
                fa.qmsg_out = reinterpret_cast<float2*>(qmi[1 - cur]);
                fa.tmsg_out = tmi[1 - cur];
                cur = 1 - cur;
            };
            auto set_states = [&](const State& si, const State& sim1) {
                fa.p1 = reinterpret_cast<const float2*>(si.p);
                fa.q1 = reinterpret_cast<const float4*>(si.q);
                fa.ub1 = si.ub;
                fa.vb1 = reinterpret_cast<const float2*>(si.vb);
                fa.u2 = sim1.u;
                fa.ub2 = sim1.ub;
            };
            set_states(state_of(cs, n_iter), state_of(cs, n_iter - 1));
            fa.st = dual_steps(0.0f, prm->alpha0, prm->alpha1, prm->eta_p, prm->eta_q);
            fa.tau_pl = taus[n_iter - 1];
            fa.mu_pl = mus[n_iter - 1];
            set_bufs();
            VFS_TRY(launch_fused_rows(rows, 0, fblocks, s, fa));
            if (rv == nullptr) {
                for (int i = n_iter - 1; i >= 1; --i) {
                    if (i - 1 < cs * ckpt) {
                        --cs;
                        VFS_TRY(recompute(cs));
                    }
                    set_states(state_of(cs, i), state_of(cs, i - 1));
                    fa.st = dual_steps(sigmas[i], prm->alpha0, prm->alpha1, prm->eta_p, prm->eta_q);
                    fa.tau_pl = taus[i - 1];
                    fa.mu_pl = mus[i - 1];
                    set_bufs();
                    VFS_TRY(launch_fused_rows(rows, 1, fblocks, s, fa));
                }
            } else {
                // F(i), F(i - 1), ... in launches of up to rv->kr units, each within one
                // checkpoint segment; (gu, g u_bar, g v_bar) ping-pong like (gp, gq, messages)
                float* gus[2] = {A.guth, A.gu2};
                float* gubs[2] = {A.gubB, A.gub2};
                float* gvbs[2] = {A.gvbB, A.gvb2};
                float* gpi[2] = {A.gp, A.gp2};
                float* gqi[2] = {A.gq, A.gq2};
                float* qmi[2] = {A.qmsg, A.qmsg2};
                float* tmi[2] = {A.tmsg, A.tmsg2};
                int scur = 0;
                int i = n_iter - 1;
                while (i >= 1) {
                    if (i - 1 < cs * ckpt) {
                        --cs;
                        VFS_TRY(recompute(cs));
                    }
                    int kc = rv->kr;
                    if (kc > i - cs * ckpt) kc = i - cs * ckpt;
                    BlockRevArgs ba;
                    ba.gu_in = gus[scur];
                    ba.gub_in = gubs[scur];
                    ba.gvb_in = reinterpret_cast<const float2*>(gvbs[scur]);
                    ba.gp_in = reinterpret_cast<const float2*>(gpi[cur]);
                    ba.gq_in = reinterpret_cast<const float4*>(gqi[cur]);
                    ba.qm_in = reinterpret_cast<const float2*>(qmi[cur]);
                    ba.tm_in = tmi[cur];
                    ba.gu_out = gus[1 - scur];
                    ba.gub_out = gubs[1 - scur];
                    ba.gvb_out = reinterpret_cast<float2*>(gvbs[1 - scur]);
                    ba.gp_out = reinterpret_cast<float2*>(gpi[1 - cur]);
                    ba.gq_out = reinterpret_cast<float4*>(gqi[1 - cur]);
                    ba.qm_out = reinterpret_cast<float2*>(qmi[1 - cur]);
                    ba.tm_out = tmi[1 - cur];
                    ba.gabc = A.gabc;
                    ba.geta = A.geta;
                    ba.gderiv = reinterpret_cast<float2*>(A.gderiv);
                    ba.gpar = gpar;
                    ba.abc = lv->abc;
                    ba.eta = lv->eta;
                    ba.deriv = reinterpret_cast<const float2*>(ws.deriv);
                    ba.mask = lv->mask;
                    for (int u = 0; u < KRMAX; ++u) {
                        const int iu = u < kc ? i - u : i;  // unused slots repeat the first unit
                        const State si = state_of(cs, iu), sm = state_of(cs, iu - 1);
                        ba.p1[u] = reinterpret_cast<const float2*>(si.p);
                        ba.q1[u] = reinterpret_cast<const float4*>(si.q);
                        ba.ub1[u] = si.ub;
                        ba.vb1[u] = reinterpret_cast<const float2*>(si.vb);
                        ba.u2[u] = sm.u;
                        ba.ub2[u] = sm.ub;
                        ba.st[u] = dual_steps(sigmas[iu], prm->alpha0, prm->alpha1, prm->eta_p, prm->eta_q);
                        ba.tau_pl[u] = taus[iu - 1];
                        ba.mu_pl[u] = mus[iu - 1];
                    }
                    ba.n_units = kc;
                    ba.H = H;
                    ba.W = W;
                    ba.tiles_x = rtiles_x;
                    ba.tiles_y = rtiles_y;
                    ba.alpha0 = prm->alpha0;
                    ba.alpha1 = prm->alpha1;
                    ba.eta_p = prm->eta_p;
                    ba.eta_q = prm->eta_q;
                    ba.lambda = prm->lambda;
                    if (probed < probe_max) {
                        hipEvent_t e0 = static_cast<hipEvent_t>(tiles->probe_events[2 * probed]);
                        const hipError_t er = hipEventRecord(e0, s);
                        if (er != hipSuccess) return static_cast<int>(er);
                    }
                    VFS_TRY(launch_adj_block(*rv, s, rtiles, ba));
                    if (probed < probe_max) {
                        hipEvent_t e1 = static_cast<hipEvent_t>(tiles->probe_events[2 * probed + 1]);
                        const hipError_t er = hipEventRecord(e1, s);
                        if (er != hipSuccess) return static_cast<int>(er);
                        tiles->probe_units[probed] = kc;
                        *tiles->probe_count = ++probed;
                    }
                    cur = 1 - cur;
                    scur = 1 - scur;
                    i -= kc;
                }
                fin_gu = gus[scur];
                fa.gu = gus[scur];
                fa.gubB = gubs[scur];
                fa.gvbB = reinterpret_cast<float2*>(gvbs[scur]);
            }
            const State s0 = state_at(ws.ckpt0, n);
            set_states(s0, s0);
            fa.st = dual_steps(sigmas[0], prm->alpha0, prm->alpha1, prm->eta_p, prm->eta_q);
            set_bufs();
            VFS_TRY(launch_fused_rows(rows, 2, fblocks, s, fa));
            fa.gu = A.guth;
            fa.gubB = A.gubB;
            fa.gvbB = reinterpret_cast<float2*>(A.gvbB);
        } else if (nck > 0) {
            // one kernel per step (primal^T, threshold^T, dual^T), messages through memory
            int has_msgs = 0;
            for (int seg = nck - 1; seg >= 0; --seg) {
                if (seg != nck - 1) VFS_TRY(recompute(seg));
                const int a = seg * ckpt, b = (seg + 1) * ckpt < n_iter ? (seg + 1) * ckpt : n_iter;
                for (int it = b - 1; it >= a; --it) {
                    const State after = state_of(seg, it + 1), before = state_of(seg, it);
                    vfs_adj_pd ad;
                    ad.H = H;
                    ad.W = W;
                    ad.gu = A.guth;
                    ad.gub = A.gubB;
                    ad.gvb = A.gvbB;
                    ad.gp = A.gp;
                    ad.gq = A.gq;
                    ad.um = A.um;
                    ad.vm = A.vm;
                    ad.qmsg = A.qmsg;
                    ad.tmsg = A.tmsg;
                    ad.gabc = A.gabc;
                    ad.geta = A.geta;
                    ad.gderiv = A.gderiv;
                    ad.gpar = gpar;
                    ad.p = after.p;
                    ad.q = after.q;
                    ad.u = after.u;
                    ad.p_prev = before.p;
                    ad.q_prev = before.q;
                    ad.u_prev = before.u;
                    ad.u_bar_prev = before.ub;
                    ad.v_bar_prev = before.vb;
                    ad.abc = lv->abc;
                    ad.eta = lv->eta;
                    ad.deriv = ws.deriv;
                    ad.mask = lv->mask;
                    VFS_TRY(vfs_adj_pd_iteration(stream, &ad, prm, taus[it], sigmas[it], mus[it], has_msgs));
                    has_msgs = 1;
                }
            }
            // adjoint of u_bar at the warp start; v_bar, p, q start at zero
            VFS_TRY(vfs_adj_pd_finish(stream, gub, A.gvb_fin, A.gubB, A.gvbB, A.um, A.vm, lv->mask, H, W));
        }
        hipLaunchKernelGGL(k_add, dim3(nb1(n)), dim3(256), 0, s, gu, fin_gu, n);
        if (hipGetLastError() != hipSuccess) return VFS_EINVAL;
        VFS_TRY(vfs_adj_derivatives(stream, gref, A.gow, A.gderiv, lv->tvf, H, W));
        VFS_TRY(vfs_adj_warp(stream, gf, A.gow, A.gtv, ws.flow, lv->other, lv->tvf, H, W));
    }
    if (rw > 0) {
        // truncated: the level's start state is held fixed
        VFS_TRY(vfs_fill(stream, gu, 0.0f, n));
        VFS_TRY(vfs_fill(stream, gub, 0.0f, n));
        VFS_TRY(vfs_fill(stream, gf, 0.0f, 2 * n));
    }
    VFS_TRY(vfs_fill(stream, A.gIs, 0.0f, n));
    VFS_TRY(vfs_adj_tensor(stream, A.gIs, gpar, A.gabc, A.geta, lv->ref_smoothed, lv->mask, H, W, prm));
    VFS_TRY(vfs_adj_gaussian(stream, gref, A.gIs, lv->mask, H, W));
    return 0;
}

}  // extern "C"
