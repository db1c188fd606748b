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

// ---- counting sort by bucket (key >> shift) -------------------------------
// Consecutive lanes of a wave mostly carry equal buckets (the particles were
// sorted a few steps ago), so each run of equal buckets in lane order makes
// one atomic: its first lane adds the run length, and in the scatter hands
// its lanes consecutive slots.  Runs of one lane (unsorted input) are correct,
// only slower.

// The run of equal keys (over the valid lanes, a prefix of the wave) that
// this lane belongs to: its first lane and, on that lane, its length.
__device__ __forceinline__ void wave_run(unsigned key, bool valid, int lane, int& head, int& len) {
    const unsigned prev = __shfl_up(key, 1);
    const bool is_head = valid && (lane == 0 || key != prev);
    const unsigned long long heads = __ballot(is_head);
    const int nvalid = __popcll(__ballot(valid));
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    head = 63 - __clzll(heads & upto);                  // lane 0 of a valid prefix is always a head
    const unsigned long long above = (lane == 63) ? 0ull : (heads >> (lane + 1));
    const int next = above ? lane + 1 + __ffsll((long long)above) - 1 : nvalid;
    len = next - lane;
}

__global__ void __launch_bounds__(256) k_cs_zero(int* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void __launch_bounds__(256)
k_cs_count(const unsigned* __restrict__ keys, int n, int shift, int* __restrict__ counts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    const unsigned b = valid ? keys[i] >> shift : 0u;
    int head, len;
    wave_run(b, valid, lane, head, len);
    if (valid && head == lane) atomicAdd(&counts[b], len);
}

__global__ void __launch_bounds__(256)
k_cs_scatter(const unsigned* __restrict__ keys, int n, int shift, int* __restrict__ offs, int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const bool valid = i < n;
    const unsigned b = valid ? keys[i] >> shift : 0u;
    int head, len;
    wave_run(b, valid, lane, head, len);
    int base = 0;
    if (valid && head == lane) base = atomicAdd(&offs[b], len);
    base = __shfl(base, head);
    if (valid) out[base + (lane - head)] = i;
}

}  // namespace

// Scan scratch of bucket_sort for nb buckets.
inline size_t bucket_scan_bytes(int nb) {
    size_t temp = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (const int*)nullptr, (int*)nullptr, nb);
    return temp;
}

// out[j] = the index i of the j-th of n keys ordered by bucket keys[i] >>
// shift (nb buckets; the order within a bucket is not defined).
inline hipError_t bucket_sort(hipStream_t s, const unsigned* keys, int n, int shift, int nb, int* counts, int* offs,
                              char* tmp, size_t temp, int* out) {
    const dim3 b(256), gn((unsigned)((n + 255) / 256)), gb((unsigned)((nb + 255) / 256));
    hipLaunchKernelGGL(k_cs_zero, gb, b, 0, s, counts, nb);
    hipLaunchKernelGGL(k_cs_count, gn, b, 0, s, keys, n, shift, counts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, temp, counts, offs, nb, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cs_scatter, gn, b, 0, s, keys, n, shift, offs, out);
    return hipGetLastError();
}

inline size_t sort_temp_bytes(int n) {
    size_t temp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, n, 0, 30);
    return temp;
}

}  // namespace jt
