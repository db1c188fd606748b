// Re-ordering particles by cell: space-filling-curve keys and the permutation
// kernels shared by the tracking launchers (kernels with internal linkage, one
// copy per translation unit).
#pragma once
#include "jt_common.h"
#include <hipcub/hipcub.hpp>

namespace jt {

// 10 bits -> every third bit of a 30-bit word.
__device__ __forceinline__ unsigned spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Hilbert index of cell (x0, x1, x2), each < 2^bits (bits <= 10), by
// Skilling's transpose algorithm (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 381 (2004)): undo the excess work per bit
// plane, Gray-encode, and interleave the transposed coordinates (x most
// significant).  Consecutive Hilbert cells are face neighbours, so a run of
// consecutive keys spans a compact box even where it crosses between aligned
// sub-cubes (a Z-order run can jump across two or three axes there).
__device__ __forceinline__ unsigned hilbert3(unsigned x0, unsigned x1, unsigned x2, int bits) {
    unsigned X[3] = {x0, x1, x2};
    const unsigned M = 1u << (bits - 1);
    for (unsigned Q = M; Q > 1; Q >>= 1) {
        const unsigned P = Q - 1;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            if (X[i] & Q) {
                X[0] ^= P;                             // invert
            } else {
                const unsigned t = (X[0] ^ X[i]) & P;  // exchange
                X[0] ^= t;
                X[i] ^= t;
            }
        }
    }
    X[1] ^= X[0];                                      // Gray encode
    X[2] ^= X[1];
    unsigned t = 0;
    for (unsigned Q = M; Q > 1; Q >>= 1)
        if (X[2] & Q) t ^= Q - 1;
    X[0] ^= t; X[1] ^= t; X[2] ^= t;
    return (spread3(X[0]) << 2) | (spread3(X[1]) << 1) | spread3(X[2]);
}

namespace {

// xout[i] = xin[perm[i]]; idout[i] = the caller's index of that particle.
__global__ void __launch_bounds__(256)
k_permute(const float* __restrict__ xin, const int* __restrict__ idin, const int* __restrict__ perm, int n,
          float* __restrict__ xout, int* __restrict__ idout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    xout[3 * (size_t)i] = xin[3 * (size_t)j];
    xout[3 * (size_t)i + 1] = xin[3 * (size_t)j + 1];
    xout[3 * (size_t)i + 2] = xin[3 * (size_t)j + 2];
    idout[i] = idin ? idin[j] : j;
}

// x[ids[i]] = xs[i]: back to the caller's order.
__global__ void __launch_bounds__(256)
k_unpermute(const float* __restrict__ xs, const int* __restrict__ ids, int n, float* __restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t j = (size_t)ids[i];
    x[3 * j] = xs[3 * (size_t)i];
    x[3 * j + 1] = xs[3 * (size_t)i + 1];
    x[3 * j + 2] = xs[3 * (size_t)i + 2];
}

}  // namespace

inline size_t sort_temp_bytes(int n) {
    size_t temp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, n, 0, 30);
    return temp;
}

}  // namespace jt
