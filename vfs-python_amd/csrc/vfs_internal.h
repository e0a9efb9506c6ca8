// Entry points shared between the library's translation units (not part of the
// C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vfs_hip.h"

namespace vfs {

// One launch of the blocked primal-dual kernel (tile_x x tile_y, k iterations per
// launch at most, `threads` per workgroup or 0 for the first listed) that runs n_iter iterations from the state `in` to the state
// `out` and also writes the state after each of the first tape_count iterations
// into tape + i * tape_stride.  State arrays: p (H,W,2), q (H,W,4), u, u_bar (H,W),
// v_bar (H,W,2); tape_off = offsets (floats) of q, u, u_bar, v_bar in a tape slot
// (p at 0).  taus / sigmas / mus: host arrays of n_iter values.
struct PdTapeLaunch {
    int H, W;
    const uint8_t* mask;
    const float *abc, *eta, *deriv;
    const float* in[5];  // p, q, u, u_bar, v_bar
    float* out[5];
    const float *taus, *sigmas, *mus;
    int n_iter, tile_x, tile_y, k, threads;
    float* tape;
    int64_t tape_stride, tape_off[4];
    int tape_count;  // states after the first tape_count iterations (<= n_iter) into the tape
    int first = 0;   // 1: a warp's first launch (p = q = v_bar = 0 on input)
};

int pd_blocks_tape(hipStream_t s, const PdTapeLaunch& a, const vfs_params* prm);

// The blocked loop's per-warp kernels (vfs_pd_block.hip): the warping vector and warped
// image at a flow (findWarpingVector + warpingImage), and median + clamp of a warp's
// thresholded u followed, with prep, by the next warp's warping vector and warped image.
int warp_prep_launch(hipStream_t s, float* tv, float* warped, const float* flow, const float* tvf, const float* other,
                     int H, int W);
int median_clamp_prep_launch(hipStream_t s, int median_k, float* u_ws, float* flow, float* tv, float* warped,
                             const float* u_th, const uint8_t* mask, const float* tvf, const float* other, int H,
                             int W, float limit, int prep);

}  // namespace vfs
