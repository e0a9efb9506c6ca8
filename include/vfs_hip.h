/*
 * vfs_hip.h -- C ABI of libvfs_hip.so, the gfx950 implementation of the VFS
 * coarse-to-fine variational fisheye stereo solver.
 *
 * Conventions
 *   - Every function returns 0 on success, VFS_EINVAL (-22) for bad sizes or
 *     null pointers, or a positive hipError_t from the launch.
 *   - `stream` is a hipStream_t passed as void* (0 = the null stream).  Work is
 *     enqueued and the function returns without synchronising.
 *   - All arrays are device pointers to C-contiguous float32 (H, W[, C]) or
 *     uint8 (H, W) buffers owned by the caller.  The library never allocates
 *     device memory for the caller's data and keeps no pointer after returning
 *     (the plan objects below keep the pointers they were built with).
 *   - "Sampled" inputs, which the reference passes as CUDA texture objects
 *     (Mirror addressing, normalised coordinates), are passed as a raw array
 *     plus its own size (Hs, Ws); the sampling semantics are applied in the
 *     kernel: bilinear with exact weights, half-sample symmetric mirror edges.
 *   - Mask values are uint8; 0 = outside.  Stencil neighbours outside the image
 *     are read through the mirror (index -1 -> 0, N -> N-1).
 *
 * Each entry point names the reference interface it replaces (file:line in
 * eshibusawa/VFS-Python).
 */
#ifndef VFS_HIP_H_
#define VFS_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFS_EINVAL (-22)
#define VFS_ETOOBIG (-7) /* a size beyond a fixed capacity of the library */

/* Solver parameters: runtime values of the constants that Python/vfs.py:52-68
 * substitutes into the CUDA source before NVRTC compilation. */
typedef struct vfs_params {
    float beta;          /* VFS_BETA          diffusion weight exp(-beta |g|^gamma) */
    float gamma;         /* VFS_GAMMA */
    float alpha0;        /* VFS_ALPHA0        TGV second-order weight */
    float alpha1;        /* VFS_ALPHA1        TGV first-order weight */
    float eta_p;         /* VFS_ETA_P = 3 */
    float eta_q;         /* VFS_ETA_Q = 2 */
    float lambda;        /* VFS_LAMBDA        data term weight */
    float limit_range;   /* VFS_UPPER_LIMIT   upper clamp of the per-warp increment */
    float f_scale;       /* VFS_FSCALE        pyramid ratio, flow scale on upsampling */
    float min_grad;      /* VFS_MIN_NORMALIZED_GRAD = 1e-8 */
    float min_tensor;    /* VFS_MIN_TESOR_VALUE = 1e-8 */
    float a11;           /* float(alpha1 * alpha1) evaluated in double (Python/vfs.cu:231) */
    float a004;          /* float(4 * alpha0 * alpha0) evaluated in double (Python/vfs.cu:230) */
    int32_t median_k;    /* VFS_MEDIAN_FILTER_KERNEL_SIZE, 3 or 5 */
} vfs_params;

/* ---- pyramids (Python/vfs.py:93-218) ------------------------------------------ */

/* castAndScale, Python/vfs.cu:145-156: out = float(in) / 256. */
int vfs_cast_and_scale(void* stream, float* out, const uint8_t* in, int H, int W);

/* downSamplingImage, Python/vfs.cu:171-174 (Bilinear sampler :35-49). */
int vfs_downsample_image(void* stream, float* out, int Hd, int Wd,
                         const float* src, int Hs, int Ws);

/* downSamplingVectorField, Python/vfs.cu:181-184; float2 elements. */
int vfs_downsample_vector_field(void* stream, float* out, int Hd, int Wd,
                                const float* src, int Hs, int Ws);

/* downSamplingMask, Python/vfs.cu:176-179 (NearestNeighbor :54-59):
 * out[Y,X] = src[((2Y+1)Hs) / (2Hd), ((2X+1)Ws) / (2Wd)] (integer division). */
int vfs_downsample_mask(void* stream, uint8_t* out, int Hd, int Wd,
                        const uint8_t* src, int Hs, int Ws);

/* ---- per-level and per-warp kernels ------------------------------------------ */

/* warpingImage, Python/vfs.cu:186-202: out = S_lin(src, X + warp.x, Y + warp.y).
 * warp is (H, W, 2); src is (Hs, Ws). */
int vfs_warp_image(void* stream, float* out, const float* warp, int H, int W,
                   const float* src, int Hs, int Ws);

/* gaussianMasked, Python/vfs.cu:204-226: binomial 3x3 / 16, 0 where mask == 0. */
int vfs_gaussian_masked(void* stream, float* out, const float* src, const uint8_t* mask,
                        int H, int W);

/* computeTensorEtaMasked, Python/vfs.cu:239-281: abc, eta are (H, W, 3). */
int vfs_compute_tensor_eta(void* stream, float* abc, float* eta, const float* img,
                           const uint8_t* mask, int H, int W, const vfs_params* prm);

/* findWarpingVector, Python/vfs.cu:283-302: out = S_lin(tvf, X + warp.x, Y + warp.y);
 * out, warp (H, W, 2); tvf (Hs, Ws, 2). */
int vfs_find_warping_vector(void* stream, float* out, const float* warp, int H, int W,
                            const float* tvf, int Hs, int Ws);

/* computeDerivatives, Python/vfs.cu:304-324: out (H, W, 2) = (Iu, Iz); ref, other
 * (H, W); tvf (H, W, 2). */
int vfs_compute_derivatives(void* stream, float* out, const float* ref, const float* other,
                            const float* tvf, int H, int W);

/* updateDualVariables, Python/vfs.cu:326-419.  p (H,W,2) and q (H,W,4) are
 * updated in place; tp (H,W,2) is written; u_bar (H,W), v_bar (H,W,2),
 * abc (H,W,3) are read. */
int vfs_update_dual(void* stream, float* p, float* tp, float* q, float sigma,
                    const float* u_bar, const float* v_bar, const float* abc,
                    const uint8_t* mask, int H, int W, const vfs_params* prm);

/* l1Thresholding, Python/vfs.cu:421-505: out (H, W) = thresholded u.  out may
 * alias u (each pixel reads only its own u). */
int vfs_l1_thresholding(void* stream, float* out, float tau, const float* tp,
                        const float* u, const float* u_bar, const float* deriv,
                        const float* eta, const uint8_t* mask, int H, int W,
                        const vfs_params* prm);

/* updatePrimalVariables, Python/vfs.cu:507-581: writes u_bar_out (H, W) and
 * v_bar_out (H, W, 2); they may alias u_bar / v_bar. */
int vfs_update_primal(void* stream, float* u_bar_out, float* v_bar_out, float tau, float mu,
                      const float* p, const float* q, const float* u, const float* u_bar,
                      const float* v_bar, const float* abc, const float* eta,
                      const uint8_t* mask, int H, int W, const vfs_params* prm);

/* medianFilter, Python/vfs.cu:583-639: k x k median, clamp-to-edge, masked
 * pixels copy their input.  ksize is 3 or 5; out must not alias in. */
int vfs_median_filter(void* stream, float* out, const float* in, const uint8_t* mask,
                      int H, int W, int ksize);

/* computeOpticalFlowWithClamp, Python/vfs.cu:641-676: in place on flow (H,W,2)
 * and u0 (H,W): d = min(uf - u0, limit); u0 += d; flow += d * tv. */
int vfs_optical_flow_with_clamp(void* stream, float* flow, float* u0, const float* uf,
                                const float* tv, const uint8_t* mask, int H, int W,
                                float limit_range);

/* upSamplingVectorField1 / upSamplingVectorField, Python/vfs.cu:678-705:
 * out = f_scale * S_lin(src, (X + 0.5) Ws / Wd - 0.5, ...); channels is 1 or 2. */
int vfs_upsample(void* stream, float* out, int Hd, int Wd, const float* src, int Hs, int Ws,
                 int channels, float f_scale);

/* triangulateRays, Python/vfs.cu:707-745: out (H,W,3); rays (H,W,3); t (3) is a
 * device pointer. */
int vfs_triangulate_rays(void* stream, float* out, const float* rays_ref,
                         const float* rays_other, const float* t, const uint8_t* mask,
                         int H, int W);

/* Fill helpers used by the solver loop (the reference uses cp.zeros / cp.copy /
 * flow[mask == 0] = 0 at Python/vfs.py:489-497,531). */
int vfs_fill(void* stream, float* out, float value, int64_t n);
int vfs_copy(void* stream, float* out, const float* in, int64_t n);
int vfs_zero_masked(void* stream, float* flow, const uint8_t* mask, int H, int W, int channels);

/* ---- equidistant camera (Python/equidistant_cam.cuh, Python/template_cam.cu) --- */

/* getRays, Python/template_cam.cu:25-37: out (H,W,3) = unproject((X, Y)). */
int vfs_cam_get_rays(void* stream, float* out, int H, int W, float f, float u0, float v0);
/* getUVFromXYZ, Python/template_cam.cu:56-70: uv (H,W,2) = project(xyz (H,W,3)). */
int vfs_cam_get_uv_from_xyz(void* stream, float* uv, const float* xyz, int H, int W,
                            float f, float u0, float v0);
/* getXYZFromUV, Python/template_cam.cu:72-86: xyz (H,W,3) = unproject(uv (H,W,2)). */
int vfs_cam_get_xyz_from_uv(void* stream, float* xyz, const float* uv, int H, int W,
                            float f, float u0, float v0);

/* getThetas, Python/template_cam.cu:39-54: out (H,W,3) = (atan2(x,z), atan2(y,z),
 * atan2(|xy|,z)) of rays (H,W,3). */
int vfs_cam_get_thetas(void* stream, float* out, const float* rays, int H, int W);
/* getXYZFromDistance, Python/template_cam.cu:88-112: out (H,W,3) = dist * ray, or
 * (0,0,-1) where mask == 0. */
int vfs_cam_xyz_from_distance(void* stream, float* out, const float* rays, const float* dist,
                              const uint8_t* mask, int H, int W);

/* get_vector_fields after the normaliser, Python/vector_field.py:34-54, fused per
 * pixel: tvf = unit((project_ref(xyz + st0) - project_ref(xyz - st1)) / 2) and
 * cvf = project_other(R xyz) - uv.  tvf, cvf, uv are (H,W,2); xyz is (H,W,3).
 * cam_ref / cam_other = (f, u0, v0), R row-major 3x3, st0 / st1 the scaled
 * translations (3): host pointers. */
int vfs_vector_fields(void* stream, float* tvf, float* cvf, const float* xyz, const float* uv,
                      int H, int W, const float* cam_ref, const float* cam_other,
                      const float* R, const float* st0, const float* st1);

/* ---- one pyramid level of VFS.solve_stereo (Python/vfs.py:473-521) ------------- */

/* Buffers of one pyramid level.  Inputs: ref, other (H,W), tvf (H,W,2), mask.
 * State (in/out): flow (H,W,2), u (H,W), u_bar (H,W).  The remaining fields are
 * caller-provided workspace of the stated shapes. */
typedef struct vfs_level {
    int32_t H, W;
    const float* ref;
    const float* other;
    const float* tvf;
    const uint8_t* mask;
    float* flow;
    float* u;
    float* u_bar;
    float* ref_smoothed;  /* (H,W)   */
    float* abc;           /* (H,W,3) */
    float* eta;           /* (H,W,3) */
    float* tv;            /* (H,W,2) */
    float* warped;        /* (H,W)   */
    float* deriv;         /* (H,W,2) */
    float* p;             /* (H,W,2) */
    float* q;             /* (H,W,4) */
    float* tp;            /* (H,W,2) */
    float* v_bar;         /* (H,W,2) */
    float* u0;            /* (H,W)   */
    float* uf;            /* (H,W)   */
} vfs_level;

/* Runs the warp loop of one level with the reference's kernel decomposition
 * (Gaussian + tensor once, then per warp: warping vector, warp, derivatives,
 * reset, n_iter x (dual, threshold, primal), median, clamp), enqueued kernel by
 * kernel on `stream`.  taus/sigmas/mus are host arrays of n_iter values. */
int vfs_solve_level_plain(void* stream, const vfs_level* lv, const vfs_params* prm,
                          const float* taus, const float* sigmas, const float* mus,
                          int n_iter, int n_warp);

/* The same level loop recorded once into a HIP graph and replayed.  A plan
 * keeps the pointers of `lv` and the schedule; launching it re-runs the level
 * on `stream`.  Returns an opaque handle in *plan. */
int vfs_plan_create_level_plain(void** plan, const vfs_level* lv, const vfs_params* prm,
                                const float* taus, const float* sigmas, const float* mus,
                                int n_iter, int n_warp);
int vfs_plan_launch(void* plan, void* stream);

/* Generic recording: vfs_capture_begin creates a stream in capture mode
 * (thread-local) and returns it; every vfs_* launch enqueued on it by this
 * thread is recorded.  vfs_capture_end stops the capture, destroys the stream
 * and instantiates the recording as a plan for vfs_plan_launch.  Used to
 * record a whole solve_stereo (pyramids, all levels, upsampling) as one graph.
 * No reference counterpart: the reference synchronises after every launch. */
int vfs_capture_begin(void** capture_stream);
int vfs_capture_end(void* capture_stream, void** plan);
int vfs_plan_destroy(void* plan);

/* ---- temporally blocked primal-dual loop (no reference counterpart; the reference
 * launches the three kernels of Python/vfs.py:505-511 once per iteration) ------- */

/* Second state set for the ping-pong of the blocked loop, shapes as in vfs_level:
 * u_th0 / u_th1 (H,W) thresholded u, u_bar2 (H,W), v_bar2 (H,W,2), p2 (H,W,2),
 * q2 (H,W,4). */
typedef struct vfs_pd_ws {
    float* u_th0;
    float* u_th1;
    float* u_bar2;
    float* v_bar2;
    float* p2;
    float* q2;
} vfs_pd_ws;

/* Tile variants compiled into the library: writes (tile_x, tile_y, k, threads)
 * quadruples to out (4 ints each) and returns their number; out == NULL returns
 * the count only. */
int vfs_pd_variants(int* out, int max_variants);

/* n_iter (1 <= n_iter <= k) primal-dual iterations of Python/vfs.py:498-515 in one
 * launch of the (tile_x, tile_y, k, threads) variant (threads 0: the first listed
 * thread count of that tile and k; likewise below): reads the state (u_in, ub_in,
 * vb_in, p_in, q_in; vb/p/q taken as zero when first != 0) and writes it to the
 * out buffers, which must not alias the in buffers.  taus/sigmas/mus: n_iter
 * host values. */
int vfs_pd_blocks(void* stream, int H, int W, const uint8_t* mask, const float* abc,
                  const float* eta, const float* deriv, const float* u_in, const float* ub_in,
                  const float* vb_in, const float* p_in, const float* q_in, float* u_out,
                  float* ub_out, float* vb_out, float* p_out, float* q_out, int first,
                  const float* taus, const float* sigmas, const float* mus, int n_iter,
                  int tile_x, int tile_y, int k, int threads, const vfs_params* prm);

/* One level of Python/vfs.py:473-521 with the blocked loop: Gaussian + tensor,
 * then per warp: derivatives, ceil(n_iter / k) blocked launches, and one kernel
 * for median + clamp + the next warp's warping vector and warped image.  Same
 * results as vfs_solve_level_plain; lv->p, q, v_bar and ws are scratch, lv->tp,
 * lv->u0 and lv->uf are unused. */
int vfs_solve_level_blocked(void* stream, const vfs_level* lv, const vfs_pd_ws* ws,
                            const vfs_params* prm, const float* taus, const float* sigmas,
                            const float* mus, int n_iter, int n_warp, int tile_x, int tile_y,
                            int k, int threads);

/* ---- batches of pairs (no reference counterpart: the reference solves one pair
 * per VFS object) ---------------------------------------------------------------- */

/* Bytes of device workspace that one level of an n_pairs batch needs for its
 * descriptor tables. */
size_t vfs_batch_workspace_bytes(int n_pairs);

/* Writes the descriptor tables of one pyramid level of a batch (lv[i], ws[i] for
 * pair i, any sizes) into d_workspace with a synchronous copy.  Call it before
 * recording a graph that uses the level; the tables hold the buffer pointers. */
int vfs_batch_prepare_level(void* d_workspace, size_t workspace_bytes, const vfs_level* lv,
                            const vfs_pd_ws* ws, int n_pairs, int tile_x, int tile_y, int k,
                            int threads);

/* vfs_solve_level_blocked for every pair of the batch at once: per warp, one
 * batched derivatives launch, ceil(n_iter / k) batched blocked launches over the
 * tiles of all pairs, and one batched median / clamp / next-warp kernel.
 * d_workspace must have been prepared with the same lv, ws, tile_x, tile_y, k, threads. */
int vfs_solve_level_blocked_batch(void* stream, const void* d_workspace, const vfs_level* lv,
                                  const vfs_pd_ws* ws, int n_pairs, const vfs_params* prm,
                                  const float* taus, const float* sigmas, const float* mus,
                                  int n_iter, int n_warp, int tile_x, int tile_y, int k,
                                  int threads);

/* The blocked primal-dual launches alone of one warp of the batch's level
 * (ceil(n_iter / k) batched launches over the tiles of all pairs, from the warp's
 * first launch, state sets alternating as in vfs_solve_level_blocked_batch), for
 * timing the batched kernel; it overwrites the pairs' primal-dual state.
 * d_workspace as for vfs_solve_level_blocked_batch. */
int vfs_pd_blocks_batch(void* stream, const void* d_workspace, const vfs_level* lv, const vfs_pd_ws* ws,
                        int n_pairs, const vfs_params* prm, const float* taus, const float* sigmas,
                        const float* mus, int n_iter, int tile_x, int tile_y, int k, int threads);

/* ---- adjoint (reverse mode) of the solve (no reference counterpart) -------------- */

/* Buffers of one reverse primal-dual iteration at one pyramid level.  The forward
 * iteration t maps (p, q, u, u_bar, v_bar)_{t-1} to (p, q, u, u_bar, v_bar)_t
 * (Python/vfs.py:505-511); its reverse maps the adjoint of the state after t to
 * the adjoint of the state before t.  Adjoint contributions addressed to a
 * neighbouring pixel travel as per-pixel messages (um, vm, qmsg, tmsg) that the
 * receiving pixel collects in the next reverse kernel; the adjoints of u_bar and
 * v_bar therefore stay split into an own part (gub, gvb) and the messages until
 * vfs_adj_pd_finish combines them. */
typedef struct vfs_adj_pd {
    int32_t H, W;
    float* gu;      /* (H,W)   adjoint of u_t on entry, of u_{t-1} on exit          */
    float* gub;     /* (H,W)   own part of the adjoint of u_bar_t / u_bar_{t-1}       */
    float* gvb;     /* (H,W,2) own part of the adjoint of v_bar_t / v_bar_{t-1}       */
    float* gp;      /* (H,W,2) adjoint of p_t on entry, of p_{t-1} on exit          */
    float* gq;      /* (H,W,4) adjoint of q_t on entry, of q_{t-1} on exit          */
    float* um;      /* (H,W,3) dual messages for u_bar (read on entry if has_msgs)   */
    float* vm;      /* (H,W,6) dual messages for v_bar                              */
    float* qmsg;    /* (H,W,4) scratch: primal messages for q                        */
    float* tmsg;    /* (H,W,4) scratch: threshold messages for tp                    */
    float* gabc;    /* (H,W,3) accumulated adjoint of abc                            */
    float* geta;    /* (H,W,3) accumulated adjoint of eta                            */
    float* gderiv;  /* (H,W,2) accumulated adjoint of deriv (Iu, Iz)                 */
    float* gpar;    /* (H,W,5) accumulated lambda, alpha0, alpha1, beta, gamma       */
    const float* p;          /* p_t   (H,W,2) */
    const float* q;          /* q_t   (H,W,4) */
    const float* u;          /* u_t   (H,W)   */
    const float* p_prev;     /* p_{t-1} */
    const float* q_prev;     /* q_{t-1} */
    const float* u_prev;     /* u_{t-1} */
    const float* u_bar_prev; /* u_bar_{t-1} */
    const float* v_bar_prev; /* v_bar_{t-1} */
    const float* abc;
    const float* eta;
    const float* deriv;
    const uint8_t* mask;
} vfs_adj_pd;

/* Reverse of one primal-dual iteration with the schedule values of that iteration;
 * has_msgs = 0 when no later iteration left messages in um / vm. */
int vfs_adj_pd_iteration(void* stream, const vfs_adj_pd* a, const vfs_params* prm, float tau, float sigma,
                         float mu, int has_msgs);

/* Complete adjoints of u_bar and v_bar before the first iteration of a loop. */
int vfs_adj_pd_finish(void* stream, float* gub_out, float* gvb_out, const float* gub, const float* gvb,
                      const float* um, const float* vm, const uint8_t* mask, int H, int W);

/* Reverse of computeOpticalFlowWithClamp after medianFilter (Python/vfs.cu:583-666):
 * gf (H,W,2) is the adjoint of the flow after the warp (it is also the adjoint of
 * the flow before); gu (H,W) the adjoint of u after the warp, replaced by the
 * clamp's share of the adjoint of the warp-start u0.  Writes guf (adjoint of the
 * median), pos (int32 window position of the median, -1 outside the mask), gtv
 * (H,W,2, adjoint of the warping vector). */
int vfs_adj_clamp(void* stream, float* gu, float* guf, int* pos, float* gtv, const float* gf,
                  const float* u_th, const float* u0, const float* tv, const uint8_t* mask, int H, int W,
                  float limit, int ksize);
/* Reverse of medianFilter: gu_th (H,W) := adjoint of the thresholded u. */
int vfs_adj_median(void* stream, float* gu_th, const float* guf, const int* pos, int H, int W, int ksize);
/* Reverse of computeDerivatives (Python/vfs.cu:120-142): gref, gow (H,W) += the texels'
 * tap shares (each a float product, summed exactly in int64 fixed point per block and
 * rounded once, so the result does not depend on the order of the additions). */
int vfs_adj_derivatives(void* stream, float* gref, float* gow, const float* gderiv, const float* tvf,
                        int H, int W);
/* Reverse of warpingImage / findWarpingVector at the warp-start flow:
 * gf (H,W,2) += gow dS(other)/dx + gtv . dS(tvf)/dx. */
int vfs_adj_warp(void* stream, float* gf, const float* gow, const float* gtv, const float* flow,
                 const float* other, const float* tvf, int H, int W);
/* Reverse of computeTensorEtaMasked (Python/vfs.cu:228-276): gimg (H,W) += (gather);
 * gpar (H,W,5) += beta, gamma, alpha0, alpha1 shares. */
int vfs_adj_tensor(void* stream, float* gimg, float* gpar, const float* gabc, const float* geta,
                   const float* img, const uint8_t* mask, int H, int W, const vfs_params* prm);
/* Reverse of gaussianMasked: gimg (H,W) += (gather). */
int vfs_adj_gaussian(void* stream, float* gimg, const float* gout, const uint8_t* mask, int H, int W);
/* Reverse of downSamplingImage: gsrc (Hs,Ws) += (gather). */
int vfs_adj_downsample(void* stream, float* gsrc, int Hs, int Ws, const float* gout, int Hd, int Wd);
/* Reverse of upSamplingVectorField(1): gsrc (Hs,Ws,channels) += f_scale S^T gout (gather). */
int vfs_adj_upsample(void* stream, float* gsrc, int Hs, int Ws, const float* gout, int Hd, int Wd,
                     int channels, float f_scale);
/* Depth loss |log z - log z_gt| per pixel from the flow (z of triangulateRays,
 * Python/vfs.cu:707-735, camera (f, u0, v0), translation x / z components):
 * term (H,W,2) = (loss, valid); gflow (H,W,2) = d loss / d flow (not divided by
 * the number of valid pixels). */
int vfs_depth_loss(void* stream, float* term, float* gflow, const float* flow, const float* z_gt,
                   const uint8_t* mask, int H, int W, float f, float u0, float v0, float tx, float tz);
/* out[c] = sum_i in[i * channels + c] in double, deterministic (wavefront, then
 * block, then one final block); partial holds nparts * channels doubles. */
int vfs_reduce_sum(void* stream, double* out, double* partial, int nparts, const float* in, int64_t n,
                   int channels);

/* Residual norm for convergence monitoring (no reference counterpart):
 * out[0] = sum over pixels with mask != 0 (every pixel when mask is NULL) of
 * sum_c (a - b)^2 over `channels` interleaved channels, out[1] = the number of
 * those pixels, both in double and deterministic (wavefront shuffles, then one
 * block); partial holds 2 * nparts doubles (1 <= nparts <= 65535). */
int vfs_residual_norm(void* stream, double* out, double* partial, int nparts, const float* a, const float* b,
                      const uint8_t* mask, int64_t n, int channels);

/* vfs_residual_norm without the caller's `partial` workspace: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It allocates
 * nothing: the partials live in a static device array of VFS_NOALLOC_MAX_PARTS pairs
 * of doubles (16 KiB) that each GPU holds until the process ends.  nparts beyond that
 * maximum returns VFS_ETOOBIG; the result is the same bits as vfs_residual_norm with
 * the same nparts.  `stream` must belong to the current device.  Calls on one GPU
 * share its array, so calls that overlap on one GPU (other streams or threads) race. */
#define VFS_NOALLOC_MAX_PARTS 1024
int vfs_residual_norm_noalloc(void* stream, double* out, int nparts, const float* a, const float* b,
                              const uint8_t* mask, int64_t n, int channels);

/* ---- adjoint level driver ---------------------------------------------------------
 * The forward warp loop of one level that records what the reverse needs, and the
 * reverse sweep of one level, each enqueued on one stream without host
 * synchronisation or memsets (both can be recorded into a HIP graph).
 *
 * Save area of one level (vfs_adj_save_bytes): per warp the warp-start u, u_bar and
 * flow, the warp's derivatives (Iu, Iz), and ceil(n_iter / ckpt) checkpoints of
 * (p, q, u, u_bar, v_bar), taken before iterations 0, ckpt, 2 ckpt, ...  Workspace (vfs_adj_workspace_bytes, one for the
 * largest level serves every level): a tape of ckpt states that a checkpoint
 * segment is recomputed into, plus the adjoint buffers.  Both must be 256-byte
 * aligned (torch allocations are). */
size_t vfs_adj_save_bytes(int H, int W, int n_warp, int n_iter, int ckpt);
size_t vfs_adj_workspace_bytes(int H, int W, int ckpt);
/* Bytes of the optional last-segment tapes of one level (vfs_adj_opts.tape_save): for
 * each of tape_warps warps, the states after every iteration of the warp's last
 * checkpoint segment (iterations ckpt (ceil(n_iter / ckpt) - 1) .. n_iter - 1). */
size_t vfs_adj_tape_bytes(int H, int W, int n_iter, int ckpt, int tape_warps);

/* Options of the adjoint level driver: kernel choices (a zero variant selects the
 * per-kernel path) and the depth of the reverse sweep.
 *   fwd_tile_x, fwd_tile_y, fwd_k, fwd_threads: blocked primal-dual variant
 *     (vfs_pd_variants; threads 0 = the first listed) of the forward; its launches
 *     end at every checkpoint;
 *   rec_tile_x, rec_tile_y, rec_k, rec_threads: blocked variant of the tape recompute;
 *     each launch runs up to rec_k iterations and writes every intermediate state to
 *     the tape;
 *   sweep: reverse primal-dual sweep.  0: one fused kernel per iteration (dual^T of
 *     iteration i with primal^T and threshold^T of iteration i - 1; the dual
 *     messages of a 32 x 8 tile and its one-pixel ring exchanged in LDS); 1: one
 *     kernel per step (vfs_adj_pd_iteration), messages through memory; 11, 12, 14:
 *     fused with tiles of 32 x 8, 32 x 16, 32 x 32 (1, 2, 4 rows per thread);
 *     21-26: temporally blocked, K = 1-6 consecutive fused units per launch on a
 *     32 x 32 extended tile (1024 threads), the adjoint state in registers and the
 *     messages in LDS; 31-33: the same with K = 1-3 on 32 x 16 (512 threads).
 *     The blocked modes keep the fused kernel for the first and last unit of a warp;
 *   reverse_warps: 0 = the exact reverse of every warp; r > 0 = only the last r warps
 *     of the level are reversed and their start state is held fixed (the adjoints of
 *     the level's initial u, u_bar, flow are returned as zero): a truncated gradient;
 *   tape_save, tape_warps: optional (NULL / 0 = none) tapes of the last checkpoint
 *     segment of the level's last tape_warps warps (vfs_adj_tape_bytes; 256-byte
 *     aligned).  The forward writes every state of those segments (it computes them
 *     anyway) and the reverse reads them instead of recomputing the segment; earlier
 *     segments are recomputed from their checkpoints as without it.  The gradients are
 *     the same bits either way.  Both calls of a level must pass the same tapes;
 *   probe_events, probe_units, probe_count, probe_max: an optional timing probe of the
 *     temporally blocked reverse launches (NULL / 0 = none).  For the j-th such launch
 *     of the call, j < probe_max, probe_events[2 j] and probe_events[2 j + 1] (hipEvent_t,
 *     created by the caller with timing enabled) are recorded on the stream just before
 *     and after it and probe_units[j] (host memory) receives its unit count; on return
 *     *probe_count (host) holds the number of launches probed. */
typedef struct vfs_adj_opts {
    int32_t fwd_tile_x, fwd_tile_y, fwd_k, fwd_threads;
    int32_t rec_tile_x, rec_tile_y, rec_k, rec_threads;
    int32_t sweep;
    int32_t reverse_warps;
    void* const* probe_events;
    int32_t* probe_units;
    int32_t* probe_count;
    int32_t probe_max;
    float* tape_save;
    int32_t tape_warps;
} vfs_adj_opts;

/* Same results as vfs_solve_level_plain (the per-kernel warp loop of
 * Python/vfs.py:473-521; the blocked variants give the same values), plus the
 * save area.  workspace (vfs_adj_workspace_bytes) holds the second state set of
 * the blocked path (unused by the per-kernel path); opts may be NULL (per-kernel). */
int vfs_adj_forward_level(void* stream, const vfs_level* lv, const vfs_params* prm, const float* taus,
                          const float* sigmas, const float* mus, int n_iter, int n_warp, int ckpt, float* save,
                          float* workspace, const vfs_adj_opts* tiles);

/* Reverse of vfs_adj_forward_level (opts: see vfs_adj_opts; NULL = per-kernel
 * recompute, sweep 0, every warp).  lv must still hold the forward's ref_smoothed,
 * abc and eta; its tv, p, q, tp, v_bar, u0, uf may be overwritten.
 * In: gu, gub (H,W) and gf (H,W,2) are the adjoints of the level's final u, u_bar
 * and flow; on return they are the adjoints of the level's initial u, u_bar and
 * flow.  gref (H,W) += adjoint of the level's reference image; gpar (H,W,5) +=
 * shares of lambda, alpha0, alpha1, beta, gamma whose sum over all pixels is the
 * level's contribution to those gradients. */
int vfs_adj_reverse_level(void* stream, const vfs_level* lv, const vfs_params* prm, const float* taus,
                          const float* sigmas, const float* mus, int n_iter, int n_warp, int ckpt,
                          const float* save, float* workspace, float* gu, float* gub, float* gf, float* gref,
                          float* gpar, const vfs_adj_opts* tiles);

/* Library version string (static storage). */
const char* vfs_version(void);

#ifdef __cplusplus
}
#endif

#endif /* VFS_HIP_H_ */
