// Median of a 3 x 3 or 5 x 5 window: the selection shared by every median filter of
// the solver (medianFilter, Python/vfs.cu:584-639) and its adjoint, so that all of them
// select the same element.
//
// 5 x 5: a selection network of 113 comparators (101 min and 101 max operations).  It is
// Batcher's odd-even merge sort on 32 wires (Knuth, TAOCP vol. 3, 5.3.4, Algorithm M)
// with wires 25-31 held at +infinity, reduced to the comparators that output 12 (the
// median) depends on: comparators with a constant +infinity input are no-ops and drop
// out, then a backward pass over the rest keeps a comparator only if one of its
// outputs is still used, and only that output.  By the 0-1 principle a comparator
// network selects the k-th smallest of all inputs if it does so for all 0/1 inputs;
// all 2^25 of them were checked when the list below was generated.  The partial
// selection sort it replaces took 234 comparators.
#pragma once

#include <hip/hip_runtime.h>

namespace vfs {

template <int KM>
__device__ __forceinline__ float median_select(float (&v)[KM * KM]);

template <>
__device__ __forceinline__ float median_select<5>(float (&v)[25])
{
    auto ce = [&](int i, int j) {
        const float a = v[i], b = v[j];
        v[i] = fminf(a, b);
        v[j] = fmaxf(a, b);
    };
    auto mn = [&](int i, int j) { v[i] = fminf(v[i], v[j]); };
    auto mx = [&](int i, int j) { v[j] = fmaxf(v[i], v[j]); };
    ce(0, 1); ce(2, 3); ce(0, 2); ce(1, 3); ce(1, 2); ce(4, 5); ce(6, 7); ce(4, 6);
    ce(5, 7); ce(5, 6); ce(0, 4); ce(2, 6); ce(2, 4); ce(1, 5); ce(3, 7); ce(3, 5);
    ce(1, 2); ce(3, 4); ce(5, 6); ce(8, 9); ce(10, 11); ce(8, 10); ce(9, 11); ce(9, 10);
    ce(12, 13); ce(14, 15); ce(12, 14); ce(13, 15); ce(13, 14); ce(8, 12); ce(10, 14); ce(10, 12);
    ce(9, 13); ce(11, 15); ce(11, 13); ce(9, 10); ce(11, 12); ce(13, 14); ce(0, 8); ce(4, 12);
    ce(4, 8); ce(2, 10); ce(6, 14); ce(6, 10); ce(2, 4); ce(6, 8); ce(10, 12); ce(1, 9);
    ce(5, 13); ce(5, 9); ce(3, 11); mn(7, 15); ce(7, 11); ce(3, 5); ce(7, 9); ce(11, 13);
    ce(1, 2); ce(3, 4); ce(5, 6); ce(7, 8); ce(9, 10); ce(11, 12); mn(13, 14); ce(16, 17);
    ce(18, 19); ce(16, 18); ce(17, 19); ce(17, 18); ce(20, 21); ce(22, 23); ce(20, 22); ce(21, 23);
    ce(21, 22); ce(16, 20); ce(18, 22); ce(18, 20); ce(17, 21); ce(19, 23); ce(19, 21); ce(17, 18);
    ce(19, 20); ce(21, 22); ce(16, 24); ce(20, 24); ce(18, 20); ce(22, 24); ce(19, 21); ce(17, 18);
    ce(19, 20); ce(21, 22); ce(23, 24); mx(0, 16); mn(8, 24); mx(8, 16); mx(4, 20); mn(12, 20);
    mn(12, 16); mx(2, 18); mn(10, 18); mn(6, 22); mx(6, 10); mx(10, 12); mx(1, 17); mx(9, 17);
    mx(5, 21); mn(13, 21); mn(13, 17); mx(3, 19); mn(11, 19); mn(7, 23); mx(7, 11); mn(11, 13);
    mx(11, 12);
    return v[12];
}

// 3 x 3: the partial selection sort (ascending order of the five smallest).
template <>
__device__ __forceinline__ float median_select<3>(float (&v)[9])
{
#pragma unroll
    for (int j = 0; j <= 4; ++j) {
#pragma unroll
        for (int i = j + 1; i < 9; ++i) {
            const float lo = fminf(v[j], v[i]);
            const float hi = fmaxf(v[j], v[i]);
            v[j] = lo;
            v[i] = hi;
        }
    }
    return v[4];
}

}  // namespace vfs
