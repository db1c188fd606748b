// Gaussian KDE (jaxtrace/density/kde.py:244-270):
//   dens(q) = mean_n (2 pi)^(-d/2) exp(-0.5 |q - p_n|^2 * f32(1/h^2)).
//
// k_kde_exhaustive: the reference's all-pairs algorithm.  One thread per
//   query; particles stream through LDS in tiles of 256 and are reused by all
//   256 queries of the block.  Terms are accumulated in float64.
// k_kde_tile_mfma (below; jt_kde_grid_cutoff_ws and jt_kde_grid_cutoff run
//   it): the cut-off sum on the structured grid on the matrix cores, 16 x 16
//   x 32 tiles.  Each dropped particle is farther than R from the grid point,
//   so each dropped term is below norm*exp(-R^2/(2h^2)).
// k_kde_reduce: the sum of a tile's per-block slices in a fixed order.
#include "jt_common.h"
#include <algorithm>


namespace jt {

__global__ void __launch_bounds__(256)
k_kde_exhaustive(const float* __restrict__ Q, int m, int D, const float* __restrict__ P, int n, int chunk,
                 float inv_h2, double scale, float* __restrict__ out, double* __restrict__ partial) {
    __shared__ float sp[256 * 3];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 256 + tid;
    // this block's particles: chunk blockIdx.y (all of them with one chunk)
    const int p0 = blockIdx.y * chunk, p1 = min(n, p0 + chunk);
    float q[3] = {0.f, 0.f, 0.f};
    if (i < m)
        for (int a = 0; a < D; ++a) q[a] = Q[(size_t)i * D + a];
    double acc = 0.0;
    for (int base = p0; base < p1; base += 256) {
        const int j = base + tid;
        for (int a = 0; a < D; ++a) sp[tid * 3 + a] = (j < p1) ? P[(size_t)j * D + a] : 0.0f;
        __syncthreads();
        const int cnt = min(256, p1 - base);
        for (int k = 0; k < cnt; ++k) {
            float dx = q[0] - sp[k * 3], dy = q[1] - sp[k * 3 + 1];
            float r2 = dx * dx + dy * dy;
            if (D == 3) { float dz = q[2] - sp[k * 3 + 2]; r2 = r2 + dz * dz; }
            acc += (double)expf(-0.5f * (r2 * inv_h2));
        }
        __syncthreads();
    }
    if (i < m) {
        if (partial) partial[(size_t)blockIdx.y * m + i] = acc;
        else out[i] = (float)(acc * scale);
    }
}

// out[i] = scale * the sum of the chunks' partial sums of query i, in chunk
// order.
__global__ void __launch_bounds__(256)
k_kde_chunk_sum(const double* __restrict__ partial, int m, int nchunk, double scale, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    double s = 0.0;
    for (int c = 0; c < nchunk; ++c) s += partial[(size_t)c * m + i];
    out[i] = (float)(s * scale);
}

struct KGrid {
    float origin[3];
    float inv;
    int dims[3];
    const float* spos;
    const int* cstart;
    const int* cend;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// k_kde_tile_mfma: the cut-off sum with 16 x 16 x 32 tiles on the matrix
// cores.  For a tile and a pass of particles p the densities are a matrix
// product,
//   D[(x, y), z] += sum_p (ex[x, p] * ey[y, p]) * ez[p, z],
// with the separable factors e*[., p] = exp(-0.5 d^2 / h^2) per tile
// coordinate.  Each of the 4 waves owns 4 of the tile's 16 y rows.
//
// Precision: the products run on the bfloat16 matrix cores
// (v_mfma_f32_32x32x16_bf16, 16 times the float32 MFMA rate) with each
// float32 operand split into two bfloat16 halves, hi = bf16(v) and
// lo = bf16(v - hi) (both round-to-nearest; v - hi is exact), so
// v = hi + lo within 2^-18 |v|.  The four products hi*hi + hi*lo + lo*hi +
// lo*lo are exact in the float32 accumulator, so each term a*b carries a
// relative error of at most 2^-17 (7.6e-6), of either sign, and the
// accumulation over a pass is float32, then float64 across passes.
// A = ex * ey is formed (one float32 multiply) and split per wave; the z
// factors are split once per particle when the factors are computed.
//
// Culling: a particle is staged only if its distance to the tile's box is
// at most R (every dropped particle is farther than R from every tile
// point), and hash-cell columns whose (x, y) distance to the box exceeds R
// are not read at all; a column's z range is cut to the z half-width left
// over by its (x, y) distance.
typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
#define KDE_TX 16                                     // tile extent in x (16 of the 32 rows of an M block)
#define KDE_TY 16                                     // in y: 4 waves x 4 rows
#define KDE_TZ 32                                     // in z: the MFMA N dimension
#define KDE_BATCH 256                                 // particles read per round (one per thread)
#ifndef KDE_P
#define KDE_P 256                                     // particles per MFMA pass (a multiple of 16)
#endif
#define KDE_STAGE (KDE_P + KDE_BATCH - 1)             // staged positions at most
#define KDE_LDP (KDE_P + 8)                           // padded row length of the factor tables
#ifndef KDE_TARGET_BLOCKS
#define KDE_TARGET_BLOCKS 16384
#endif
#ifndef KDE_MAX_SPLIT
#define KDE_MAX_SPLIT 8
#endif

__global__ void __launch_bounds__(256) k_kde_zero(float* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = 0.0f;
}

// Two float32 values rounded to a packed bfloat16 pair, and back.
__device__ __forceinline__ unsigned pk_bf16(float a, float b) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = {(__bf16)a, (__bf16)b};
    return __builtin_bit_cast(unsigned, v);
}
__device__ __forceinline__ float bf16_lo_f32(unsigned w) { return __uint_as_float(w << 16); }
// The same as one shift instruction of the packed word: written as C the
// compiler rounds the element a second time instead, which costs the 2-D
// kernel's factor phase a conversion per pair (the 3-D kernel measured 4 %
// slower with this form, so it keeps the one above).
__device__ __forceinline__ float bf16_lo_f32_shift(unsigned w) {
    unsigned r;
    asm("v_lshlrev_b32 %0, 16, %1" : "=v"(r) : "v"(w));
    return __uint_as_float(r);
}
__device__ __forceinline__ float bf16_hi_f32(unsigned w) { return __uint_as_float(w & 0xffff0000u); }

// Eight float32 values split into their bfloat16 hi and lo halves.
__device__ __forceinline__ void split8(const float v[8], bf16x8& hi, bf16x8& lo) {
    unsigned h[4], l[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        h[k] = pk_bf16(v[2 * k], v[2 * k + 1]);
        l[k] = pk_bf16(v[2 * k] - bf16_lo_f32(h[k]), v[2 * k + 1] - bf16_hi_f32(h[k]));
    }
    typedef unsigned u32x4v __attribute__((ext_vector_type(4)));
    hi = __builtin_bit_cast(bf16x8, (u32x4v){h[0], h[1], h[2], h[3]});
    lo = __builtin_bit_cast(bf16x8, (u32x4v){l[0], l[1], l[2], l[3]});
}

// nsplit blocks share a tile: block b takes tile b / nsplit and, of the
// tile's cell columns (cx, cy), those with index = b % nsplit modulo nsplit;
// its partial densities go to its own slice of partial (or, with no
// workspace, are added into out, zeroed first, with atomics).  The particle
// counts of the tiles' boxes differ widely on clustered data, so one block
// per tile leaves a tail of heavy blocks running alone.
// Tiles are KDE_TX x KDE_TY x KDE_TZ grid points; each pass of at most
// KDE_P staged particles is D[(x, y), z] += A[(x, y), p] B[p, z] with
// A = ex * ey (formed and split per wave) and B = ez (split once per
// particle), on v_mfma_f32_32x32x16_bf16: the 32 z of the tile are the N
// dimension, so each formed and split A value feeds 32 output columns.
// Wave w owns the y rows 4w .. 4w + 3, as two 32-row M blocks (two y rows
// of 16 x each).
__global__ void __launch_bounds__(256)
k_kde_tile_mfma(const float* __restrict__ gx, int nx, const float* __restrict__ gy, int ny,
                const float* __restrict__ gz, int nz, KGrid g, float R, float inv_h2, double scale,
                int nsplit, float* __restrict__ out, float* __restrict__ partial,
                const float* __restrict__ pall, int n_all, int chunk) {
    __shared__ __attribute__((aligned(16))) float exs[KDE_TX][KDE_LDP];       // x factors [x][p]
    __shared__ __attribute__((aligned(16))) float eys[KDE_TY][KDE_LDP];       // y factors [y][p]
    __shared__ __attribute__((aligned(16))) __bf16 ezh[KDE_TZ][KDE_LDP];      // z factors [z][p], hi half
    __shared__ __attribute__((aligned(16))) __bf16 ezl[KDE_TZ][KDE_LDP];      // lo half
    __shared__ __attribute__((aligned(16))) float tcs[KDE_TX + KDE_TY + KDE_TZ];   // the tile's x | y | z
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tiles_z = (nz + KDE_TZ - 1) / KDE_TZ, tiles_x = (nx + KDE_TX - 1) / KDE_TX;
    const int part = blockIdx.x % nsplit;
    int b = blockIdx.x / nsplit;
    const int bz = b % tiles_z; b /= tiles_z;
    const int bx = b % tiles_x; b /= tiles_x;
    const int by = b;
    // the tile's coordinates (clamped at the grid's end; clamped points are not written)
    if (tid < KDE_TX) tcs[tid] = gx[min(bx * KDE_TX + tid, nx - 1)];
    else if (tid < KDE_TX + KDE_TY) tcs[tid] = gy[min(by * KDE_TY + tid - KDE_TX, ny - 1)];
    else if (tid < KDE_TX + KDE_TY + KDE_TZ) tcs[tid] = gz[min(bz * KDE_TZ + tid - KDE_TX - KDE_TY, nz - 1)];
    __syncthreads();
    const float* xs = tcs;
    const float* ys = tcs + KDE_TX;
    const float* zs = tcs + KDE_TX + KDE_TY;
    float lo[3] = {xs[0], ys[0], zs[0]}, hi[3] = {xs[0], ys[0], zs[0]};
    for (int i = 1; i < KDE_TX; ++i) { lo[0] = fminf(lo[0], xs[i]); hi[0] = fmaxf(hi[0], xs[i]); }
    for (int i = 1; i < KDE_TY; ++i) { lo[1] = fminf(lo[1], ys[i]); hi[1] = fmaxf(hi[1], ys[i]); }
    for (int i = 1; i < KDE_TZ; ++i) { lo[2] = fminf(lo[2], zs[i]); hi[2] = fmaxf(hi[2], zs[i]); }
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = clampi((int)floorf((lo[a] - R - g.origin[a]) * g.inv), 0, g.dims[a] - 1);
        c1[a] = clampi((int)floorf((hi[a] + R - g.origin[a]) * g.inv), 0, g.dims[a] - 1);
    }
    const float R2 = R * R;
    const float cell = 1.0f / g.inv;
    const float nl2 = -0.5f * inv_h2 * 1.44269504f;    // exp(-0.5 d^2 / h^2) = 2^(nl2 d^2)
    const int row = lane & 31, kg = lane >> 5;   // MFMA operand lane: A row / B column, and k group
    __shared__ int wcount[2][4];
    __shared__ float sx[KDE_STAGE], sy[KDE_STAGE], sz[KDE_STAGE];
    double accd[2][16];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int q = 0; q < 16; ++q) accd[mb][q] = 0.0;
    // Sum the first n (<= KDE_P) staged particles: their factors (one
    // particle per thread), then the MFMA passes over them.  Block-uniform.
    auto consume = [&](int n) {
        const int n16 = (n + 15) & ~15;            // padded with zero x factors to whole 16-particle steps
        if (tid < n16) {
            const bool real = tid < n;
            // a padding slot sits at 1e30: its factors are exactly 0
            const float px = real ? sx[tid] : 1e30f, py = real ? sy[tid] : 1e30f, pz = real ? sz[tid] : 1e30f;
#pragma unroll
            for (int i = 0; i < KDE_TX; ++i) {
                const float dx = xs[i] - px;
                exs[i][tid] = __builtin_amdgcn_exp2f(nl2 * (dx * dx));
            }
#pragma unroll
            for (int i = 0; i < KDE_TY; ++i) {
                const float dy = ys[i] - py;
                eys[i][tid] = __builtin_amdgcn_exp2f(nl2 * (dy * dy));
            }
#pragma unroll
            for (int i = 0; i < KDE_TZ; ++i) {
                const float dz = zs[i] - pz;
                const float f = __builtin_amdgcn_exp2f(nl2 * (dz * dz));
                const __bf16 hz = (__bf16)f;
                ezh[i][tid] = hz;
                ezl[i][tid] = (__bf16)(f - (float)hz);
            }
        }
        __syncthreads();
        floatx16 acc[2];
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int q = 0; q < 16; ++q) acc[mb][q] = 0.0f;
        for (int p0 = 0; p0 < n16; p0 += 16) {
            const int pk = p0 + 8 * kg;               // this lane's 8 particles of the step
            const float4 xa = *reinterpret_cast<const float4*>(&exs[row & 15][pk]);
            const float4 xb = *reinterpret_cast<const float4*>(&exs[row & 15][pk + 4]);
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&ezh[row][pk]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&ezl[row][pk]);
            const float xv[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                const int y = 4 * wv + 2 * mb + (row >> 4);
                const float4 ya = *reinterpret_cast<const float4*>(&eys[y][pk]);
                const float4 yb = *reinterpret_cast<const float4*>(&eys[y][pk + 4]);
                const float yv[8] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y, yb.z, yb.w};
                float av[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) av[j] = xv[j] * yv[j];
                bf16x8 ah, al;
                split8(av, ah, al);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[mb], 0, 0, 0);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[mb], 0, 0, 0);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[mb], 0, 0, 0);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc[mb], 0, 0, 0);
            }
        }
#pragma unroll
        for (int mb = 0; mb < 2; ++mb)
#pragma unroll
            for (int q = 0; q < 16; ++q) accd[mb][q] += (double)acc[mb][q];
        __syncthreads();                               // factors and staged positions are rewritten next
    };
    int staged = 0, round = 0;                         // block-uniform
    // Stage the points [s, e) of src (x, y, z records), with cull those
    // within R of the tile's box, consuming every KDE_P of them.
    auto stage_run = [&](const float* __restrict__ src, int s, int e, bool cull) {
        // this thread's particle of the next round, loaded one round ahead
        float qx = 0.f, qy = 0.f, qz = 0.f;
        if ((long long)s + tid < (long long)e) {
            qx = src[(size_t)(s + tid) * 3]; qy = src[(size_t)(s + tid) * 3 + 1];
            qz = src[(size_t)(s + tid) * 3 + 2];
        }
        for (int base = s; base < e; base += KDE_BATCH) {
            const int j = base + tid;
            const float px = qx, py = qy, pz = qz;
            if ((long long)j + KDE_BATCH < (long long)e) {
                qx = src[(size_t)(j + KDE_BATCH) * 3]; qy = src[(size_t)(j + KDE_BATCH) * 3 + 1];
                qz = src[(size_t)(j + KDE_BATCH) * 3 + 2];
            }
            // keep the particles within R of the tile's box: each one
            // dropped is farther than R from every tile point
            const float dx = fmaxf(fmaxf(lo[0] - px, px - hi[0]), 0.0f);
            const float dy = fmaxf(fmaxf(lo[1] - py, py - hi[1]), 0.0f);
            const float dz = fmaxf(fmaxf(lo[2] - pz, pz - hi[2]), 0.0f);
            const bool keep = j < e && (!cull || (dx * dx + dy * dy) + dz * dz <= R2);
            // staging slots in thread order (deterministic): wave counts, then
            // offsets; the counts alternate between two rows, as a wave may
            // write the next round's count while another still reads these
            const unsigned long long kb = __ballot(keep);
            if (lane == 0) wcount[round & 1][wv] = __popcll(kb);
            __syncthreads();
            int off = 0, cnt = 0;
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int c = wcount[round & 1][w];
                off += w < wv ? c : 0;
                cnt += c;
            }
            ++round;
            if (keep) {
                const int slot = staged + off + __popcll(kb & ((1ull << lane) - 1ull));
                sx[slot] = px; sy[slot] = py; sz[slot] = pz;
            }
            staged += cnt;
            while (staged >= KDE_P) {
                __syncthreads();                       // the staged positions are visible
                consume(KDE_P);
                // move the rest to the front
                const int rest = staged - KDE_P;
                float a = 0.f, b2 = 0.f, c2 = 0.f;
                if (tid < rest) { a = sx[KDE_P + tid]; b2 = sy[KDE_P + tid]; c2 = sz[KDE_P + tid]; }
                __syncthreads();
                if (tid < rest) { sx[tid] = a; sy[tid] = b2; sz[tid] = c2; }
                staged = rest;
            }
        }
    };
    if (pall) {
        // every pair: this block's chunk of the particles, in their order
        const long long s0 = (long long)part * chunk;
        const long long e0 = min(s0 + chunk, (long long)n_all);
        if (s0 < e0) stage_run(pall, (int)s0, (int)e0, false);
    } else {
        const int ncy = c1[1] - c0[1] + 1, ncol = (c1[0] - c0[0] + 1) * ncy;
        for (int k = part; k < ncol; k += nsplit) {
            const int cx = c0[0] + k / ncy, cy = c0[1] + k % ncy;
            // the column's (x, y) distance to the box, from its cells' bounds
            // (widened by a thousandth of a cell against rounding)
            const float slack = 1e-3f * cell;
            const float xl = g.origin[0] + (float)cx * cell - slack, xh = g.origin[0] + (float)(cx + 1) * cell + slack;
            const float yl = g.origin[1] + (float)cy * cell - slack, yh = g.origin[1] + (float)(cy + 1) * cell + slack;
            const float gx_ = fmaxf(fmaxf(lo[0] - xh, xl - hi[0]), 0.0f);
            const float gy_ = fmaxf(fmaxf(lo[1] - yh, yl - hi[1]), 0.0f);
            const float rem = R2 - (gx_ * gx_ + gy_ * gy_);
            if (rem < 0.0f) continue;                      // every point of the column is farther than R
            const float rz = sqrtf(rem) + slack;
            const int z0 = max(c0[2], clampi((int)floorf((lo[2] - rz - g.origin[2]) * g.inv), 0, g.dims[2] - 1));
            const int z1 = min(c1[2], clampi((int)floorf((hi[2] + rz - g.origin[2]) * g.inv), 0, g.dims[2] - 1));
            if (z0 > z1) continue;
            // cells (cx, cy, z0..z1) are consecutive ids: their points are one run
            const int id0 = (cx * g.dims[1] + cy) * g.dims[2] + z0;
            const int id1 = (cx * g.dims[1] + cy) * g.dims[2] + z1;
            // the run's ends: the lanes of each wave read the cells' ranges at
            // once (64 cells a round) and reduce them, so every wave has them
            int s = 0x7fffffff, e = 0;
            for (int base = id0; base <= id1; base += 64) {
                const int id = base + lane;
                int s2 = 0x7fffffff, e2 = 0;
                if (id <= id1) {
                    const int cs = g.cstart[id], ce = g.cend[id];
                    if (ce > cs) { s2 = cs; e2 = ce; }
                }
                for (int o = 32; o > 0; o >>= 1) {
                    s2 = min(s2, __shfl_xor(s2, o));
                    e2 = max(e2, __shfl_xor(e2, o));
                }
                s = min(s, s2);
                e = max(e, e2);
            }
            if (s >= e) continue;                          // no particles in these cells (s is still INT_MAX)
            stage_run(g.spos, s, e, true);
        }
    }
    if (staged > 0) {
        __syncthreads();
        consume(staged);
    }
    // D of M block mb: lane holds z = lane % 32 and, for q = 0 .. 15, the
    // block row 8 (q / 4) + 4 (lane / 32) + q % 4 = (x = row % 16,
    // y = 4 wv + 2 mb + row / 16)
    // partial != nullptr: this block's sums go to its own slice of partial
    // (tile-local order ((y * KDE_TX + x) * KDE_TZ + z), every point), and
    // k_kde_reduce adds the slices of a tile in a fixed order; else they are
    // added into out with float atomics
    const int iz = bz * KDE_TZ + row;
    float* slice = partial ? partial + (size_t)blockIdx.x * (KDE_TX * KDE_TY * KDE_TZ) : nullptr;
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int r = 8 * (q >> 2) + 4 * kg + (q & 3);
            const int lx = r & 15, ly = 4 * wv + 2 * mb + (r >> 4);
            const float v = (float)(accd[mb][q] * scale);
            if (slice) {
                slice[(ly * KDE_TX + lx) * KDE_TZ + row] = v;
            } else {
                const int ix = bx * KDE_TX + lx, iy = by * KDE_TY + ly;
                if (ix < nx && iy < ny && iz < nz && accd[mb][q] != 0.0)
                    unsafeAtomicAdd(&out[((size_t)iy * nx + ix) * nz + iz], v);
            }
        }
    }
}

// out (ny, nx, nz) = the sum over a tile's nsplit slices of partial, in
// slice order (float64), so the result does not depend on the order the
// blocks ran in.
__global__ void __launch_bounds__(256)
k_kde_reduce(const float* __restrict__ partial, int nx, int ny, int nz, int nsplit, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nx * ny * nz) return;
    const int iz = (int)(i % nz);
    const size_t t = i / nz;
    const int ix = (int)(t % nx), iy = (int)(t / nx);
    const int tiles_z = (nz + KDE_TZ - 1) / KDE_TZ, tiles_x = (nx + KDE_TX - 1) / KDE_TX;
    const size_t tile = ((size_t)(iy / KDE_TY) * tiles_x + ix / KDE_TX) * tiles_z + iz / KDE_TZ;
    const int local = ((iy % KDE_TY) * KDE_TX + ix % KDE_TX) * KDE_TZ + iz % KDE_TZ;
    const float* p = partial + tile * nsplit * (size_t)(KDE_TX * KDE_TY * KDE_TZ) + local;
    double s = 0.0;
    for (int k = 0; k < nsplit; ++k) s += (double)p[(size_t)k * (KDE_TX * KDE_TY * KDE_TZ)];
    out[i] = (float)s;
}


// ---- 2-D grid, all pairs, on the matrix cores --------------------------------
// The reference's all-pairs sum on evaluate_2d's grid (kde.py:205-221,
// 244-270) is a matrix product: with the separable Gaussian,
//   D[y, x] = sum_p ey[y, p] * ex[x, p],  e*[., p] = exp(-0.5 d^2 / h^2),
// so both operands are per-particle factors and no product has to be formed
// per wave.  A block takes a 64 x 128 (y x x) tile of the grid and a chunk of
// the particles; per pass of 64 particles its threads compute the 192 factor
// rows (two particles a thread, so one packed conversion rounds both), split
// into bfloat16 hi and lo halves, into LDS, then run 4 MFMA steps:
// wave w owns the tile's x columns 32w .. 32w + 31 against both 32-row y
// blocks (v_mfma_f32_32x32x16_bf16, the four hi/lo products of each term).
// Pass sums are float32 over K2_FOLD passes, then float64; each block writes
// its partial tile to its own slice and k_kde2d_reduce adds a tile's slices in
// chunk order.
#define K2_TY 64                                      // tile rows (y): two 32-row M blocks
#define K2_TX 128                                     // tile columns (x): one 32-column N block per wave
#define K2_P 64                                       // particles per pass (4 MFMA steps of 16)
#ifndef K2_LDP
#define K2_LDP (K2_P + 4)                             // padded factor row (bf16 elements): 52 KiB, 3 blocks a CU
#endif
#define K2_FOLD 4                                     // passes summed in float32 before the float64 fold
#define K2_TARGET_BLOCKS 4096
#ifndef K2_FU
#define K2_FU 8                                       // factor rows unrolled per loop trip
#endif

__global__ void __launch_bounds__(256, 3)
k_kde2d_tile(const float* __restrict__ P, int n, int chunk, int nchunk, const float* __restrict__ gx, int nx,
             const float* __restrict__ gy, int ny, float nh, double scale, float* __restrict__ partial) {
    __shared__ __attribute__((aligned(16))) __bf16 eyh[K2_TY][K2_LDP];
    __shared__ __attribute__((aligned(16))) __bf16 eyl[K2_TY][K2_LDP];
    __shared__ __attribute__((aligned(16))) __bf16 exh[K2_TX][K2_LDP];
    __shared__ __attribute__((aligned(16))) __bf16 exl[K2_TX][K2_LDP];
    __shared__ float tcs[K2_TY + K2_TX];                                      // the tile's y | x
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int c = blockIdx.x % nchunk, t = blockIdx.x / nchunk;
    const int tiles_x = (nx + K2_TX - 1) / K2_TX;
    const int bx = t % tiles_x, by = t / tiles_x;
    // coordinates clamped at the grid's end (those points are not read back)
    if (tid < K2_TY) tcs[tid] = gy[min(by * K2_TY + tid, ny - 1)];
    else if (tid < K2_TY + K2_TX) tcs[tid] = gx[min(bx * K2_TX + tid - K2_TY, nx - 1)];
    __syncthreads();
    const int row = lane & 31, kg = lane >> 5;         // MFMA operand lane: A row / B column, and k group
    const float nl2 = nh * 1.44269504f;                // exp(nh d^2) = 2^(nl2 d^2)
    const int pq = tid & 31, grp = tid >> 5;           // particle pair of a pass, factor rows grp mod 8
    const long long p0 = (long long)c * chunk;
    const long long p1 = min(p0 + chunk, (long long)n);
    floatx16 acc[2];
    double accd[2][16];
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int q = 0; q < 16; ++q) { acc[mb][q] = 0.0f; accd[mb][q] = 0.0; }
    // two factors of one row (particles 2 pq and 2 pq + 1) split into bf16 hi
    // and lo pairs, one 32-bit store each
    auto put2 = [&](__bf16 (*th)[K2_LDP], __bf16 (*tl)[K2_LDP], int r, float fa, float fb) {
        const unsigned hp = pk_bf16(fa, fb);
        const unsigned lp = pk_bf16(fa - bf16_lo_f32_shift(hp), fb - bf16_hi_f32(hp));
        *reinterpret_cast<unsigned*>(&th[r][2 * pq]) = hp;
        *reinterpret_cast<unsigned*>(&tl[r][2 * pq]) = lp;
    };
    int since_fold = 0;
    for (long long base = p0; base < p1; base += K2_P) {
        // factors: thread (pair pq, group grp) takes rows grp, grp + 8, ... of
        // the 64 y rows and the 128 x rows for particles 2 pq and 2 pq + 1; a
        // particle past the chunk sits at 1e30, where its factors are exactly 0
        const long long ja = base + 2 * pq;
        const bool ra = ja < p1, rb = ja + 1 < p1;
        const float xa = ra ? P[2 * ja] : 1e30f, ya = ra ? P[2 * ja + 1] : 1e30f;
        const float xb = rb ? P[2 * ja + 2] : 1e30f, yb = rb ? P[2 * ja + 3] : 1e30f;
#pragma unroll
        for (int i = 0; i < K2_TY / 8; ++i) {
            const int r = 8 * i + grp;
            const float cy = tcs[r];
            const float da = cy - ya, db = cy - yb;
            put2(eyh, eyl, r, __builtin_amdgcn_exp2f(nl2 * (da * da)), __builtin_amdgcn_exp2f(nl2 * (db * db)));
        }
#pragma unroll K2_FU
        for (int i = 0; i < K2_TX / 8; ++i) {
            const int r = 8 * i + grp;
            const float cx = tcs[K2_TY + r];
            const float da = cx - xa, db = cx - xb;
            put2(exh, exl, r, __builtin_amdgcn_exp2f(nl2 * (da * da)), __builtin_amdgcn_exp2f(nl2 * (db * db)));
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < K2_P / 16; ++s) {
            const int pk = 16 * s + 8 * kg;            // this lane's 8 particles of the step
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&exh[32 * wv + row][pk]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&exl[32 * wv + row][pk]);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb) {
                const bf16x8 ah = *reinterpret_cast<const bf16x8*>(&eyh[32 * mb + row][pk]);
                const bf16x8 al = *reinterpret_cast<const bf16x8*>(&eyl[32 * mb + row][pk]);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bh, acc[mb], 0, 0, 0);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bl, acc[mb], 0, 0, 0);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bh, acc[mb], 0, 0, 0);
                acc[mb] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bl, acc[mb], 0, 0, 0);
            }
        }
        __syncthreads();                               // the factor tables are rewritten next
        if (++since_fold == K2_FOLD) {
            since_fold = 0;
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int q = 0; q < 16; ++q) { accd[mb][q] += (double)acc[mb][q]; acc[mb][q] = 0.0f; }
        }
    }
    // D of y block mb: lane holds x = 32 wv + lane % 32 and, for q = 0 .. 15,
    // y = 32 mb + 8 (q / 4) + 4 (lane / 32) + q % 4
    float* slice = partial + (size_t)blockIdx.x * (K2_TY * K2_TX);
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int ly = 32 * mb + 8 * (q >> 2) + 4 * kg + (q & 3);
            slice[ly * K2_TX + 32 * wv + row] = (float)((accd[mb][q] + (double)acc[mb][q]) * scale);
        }
}

// out (ny, nx) = the sum over a tile's nchunk slices of partial, in chunk
// order (float64).
__global__ void __launch_bounds__(256)
k_kde2d_reduce(const float* __restrict__ partial, int nx, int ny, int nchunk, float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= (size_t)nx * ny) return;
    const int ix = (int)(i % nx), iy = (int)(i / nx);
    const int tiles_x = (nx + K2_TX - 1) / K2_TX;
    const size_t t = (size_t)(iy / K2_TY) * tiles_x + ix / K2_TX;
    const int local = (iy % K2_TY) * K2_TX + ix % K2_TX;
    const float* p = partial + t * nchunk * (size_t)(K2_TY * K2_TX) + local;
    double s = 0.0;
    for (int k = 0; k < nchunk; ++k) s += (double)p[(size_t)k * (K2_TY * K2_TX)];
    out[i] = (float)s;
}

}  // namespace jt

using namespace jt;

extern "C" int jt_kde_eval(const float* P, int64_t n, int32_t dim, const float* Q, int64_t m, float h,
                           float* out, void* stream) {
    JT_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    JT_REQUIRE(n >= 1 && n < (1LL << 31) && m >= 0 && m < (1LL << 31), "sizes out of range");
    JT_REQUIRE(h > 0.0f, "bandwidth must be positive");
    if (m == 0) return JT_OK;
    JT_REQUIRE(P && Q && out, "NULL pointer");
    const double hh = (double)h;
    const float inv_h2 = (float)(1.0 / (hh * hh));                 // kde.py:251
    const double norm = pow(2.0 * M_PI, -0.5 * dim);                // kernels.py:91
    hipLaunchKernelGGL(k_kde_exhaustive, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       Q, (int)m, dim, P, (int)n, (int)n, inv_h2, norm / (double)n, out, (double*)nullptr);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

// The particle chunks of a per-point evaluation: enough blocks to fill the
// chip when the queries are few.
#define KDE_EVAL_TARGET_BLOCKS 2048
static void kde_eval_chunks(int64_t n, int64_t m, int& nchunk, long long& chunk) {
    const long long qblocks = std::max(1LL, (long long)((m + 255) / 256));
    const long long want = std::max(1LL, std::min((long long)KDE_EVAL_TARGET_BLOCKS / qblocks, (long long)(n + 255) / 256));
    chunk = ((n + want - 1) / want + 255) / 256 * 256;
    nchunk = (int)((n + chunk - 1) / chunk);
}

extern "C" size_t jt_kde_eval_workspace_size(int64_t n, int64_t m) {
    if (n < 1 || m < 1) return 0;
    int nchunk;
    long long chunk;
    kde_eval_chunks(n, m, nchunk, chunk);
    return nchunk > 1 ? (size_t)nchunk * m * sizeof(double) + 256 : 0;
}

extern "C" int jt_kde_eval_ws(const float* P, int64_t n, int32_t dim, const float* Q, int64_t m, float h,
                              float* out, void* ws, size_t ws_bytes, void* stream) {
    JT_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    JT_REQUIRE(n >= 1 && n < (1LL << 31) && m >= 0 && m < (1LL << 31), "sizes out of range");
    JT_REQUIRE(h > 0.0f, "bandwidth must be positive");
    if (m == 0) return JT_OK;
    JT_REQUIRE(P && Q && out, "NULL pointer");
    int nchunk;
    long long chunk;
    kde_eval_chunks(n, m, nchunk, chunk);
    if (nchunk == 1) return jt_kde_eval(P, n, dim, Q, m, h, out, stream);
    const size_t need = jt_kde_eval_workspace_size(n, m);
    Arena arena(ws, ws_bytes);
    double* partial = arena.take<double>((size_t)nchunk * m);
    if (!partial || !arena.ok()) return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", need, ws_bytes);
    const double hh = (double)h;
    const float inv_h2 = (float)(1.0 / (hh * hh));                 // kde.py:251
    const double norm = pow(2.0 * M_PI, -0.5 * dim);                // kernels.py:91
    hipLaunchKernelGGL(k_kde_exhaustive, dim3((unsigned)((m + 255) / 256), (unsigned)nchunk), dim3(256), 0,
                       (hipStream_t)stream, Q, (int)m, dim, P, (int)n, (int)chunk, inv_h2, 0.0, out, partial);
    JT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_kde_chunk_sum, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const double*)partial, (int)m, nchunk, norm / (double)n, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

// The tile count and the blocks per tile of a grid.
static void kde_tiling(int nx, int ny, int nz, long long& tiles, int& nsplit) {
    tiles = (long long)((ny + KDE_TY - 1) / KDE_TY) * ((nx + KDE_TX - 1) / KDE_TX) * ((nz + KDE_TZ - 1) / KDE_TZ);
    // about KDE_TARGET_BLOCKS blocks (the chip holds 512 at a time), at most KDE_MAX_SPLIT per tile
    nsplit = (int)std::max(1LL, std::min((long long)KDE_MAX_SPLIT, (long long)KDE_TARGET_BLOCKS / tiles));
}

static int kde_cutoff_run(const jt_hashgrid* grid, int64_t n_total, const float* gx, int32_t nx, const float* gy,
                          int32_t ny, const float* gz, int32_t nz, float h, float cutoff, float* out,
                          float* partial, void* stream) {
    JT_REQUIRE(grid && grid->dim == 3 && grid->key_mode == JT_KEY_X64, "needs a 3-D x64 hash grid");
    JT_REQUIRE(grid->inv > 0.0f && cutoff > 0.0f && (1.0f / grid->inv) >= cutoff * 0.124999f,
               "hash-grid cell size must be >= cutoff / 8");
    JT_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && n_total >= 1, "bad sizes");
    JT_REQUIRE(gx && gy && gz && out && grid->sorted_pos && grid->cell_start && grid->cell_end, "NULL pointer");
    KGrid g;
    for (int a = 0; a < 3; ++a) { g.origin[a] = grid->origin[a]; g.dims[a] = grid->dims[a]; }
    g.inv = grid->inv;
    g.spos = grid->sorted_pos; g.cstart = grid->cell_start; g.cend = grid->cell_end;
    const double hh = (double)h;
    const float inv_h2 = (float)(1.0 / (hh * hh));
    const double norm = pow(2.0 * M_PI, -1.5);
    long long tiles;
    int nsplit;
    kde_tiling(nx, ny, nz, tiles, nsplit);
    JT_REQUIRE(tiles * nsplit < (1LL << 31), "grid too large");
    const size_t npts = (size_t)nx * ny * nz;
    if (!partial) {
        hipLaunchKernelGGL(k_kde_zero, dim3((unsigned)std::min<size_t>((npts + 255) / 256, 4096)), dim3(256), 0,
                           (hipStream_t)stream, out, npts);
        JT_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_kde_tile_mfma, dim3((unsigned)(tiles * nsplit)), dim3(256), 0, (hipStream_t)stream,
                       gx, nx, gy, ny, gz, nz, g, cutoff, inv_h2, norm / (double)n_total, nsplit, out, partial,
                       (const float*)nullptr, 0, 0);
    JT_HIP(hipGetLastError());
    if (partial) {
        hipLaunchKernelGGL(k_kde_reduce, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                           (const float*)partial, nx, ny, nz, nsplit, out);
        JT_HIP(hipGetLastError());
    }
    return JT_OK;
}

extern "C" int jt_kde_grid_cutoff(const jt_hashgrid* grid, int64_t n_total, const float* gx, int32_t nx,
                                  const float* gy, int32_t ny, const float* gz, int32_t nz, float h,
                                  float cutoff, float* out, void* stream) {
    return kde_cutoff_run(grid, n_total, gx, nx, gy, ny, gz, nz, h, cutoff, out, nullptr, stream);
}

extern "C" size_t jt_kde_grid_cutoff_workspace_size(int32_t nx, int32_t ny, int32_t nz) {
    if (nx < 1 || ny < 1 || nz < 1) return 0;
    long long tiles;
    int nsplit;
    kde_tiling(nx, ny, nz, tiles, nsplit);
    return (size_t)tiles * nsplit * (KDE_TX * KDE_TY * KDE_TZ) * sizeof(float) + 256;
}

extern "C" int jt_kde_grid_cutoff_ws(const jt_hashgrid* grid, int64_t n_total, const float* gx, int32_t nx,
                                     const float* gy, int32_t ny, const float* gz, int32_t nz, float h, float cutoff,
                                     float* out, void* ws, size_t ws_bytes, void* stream) {
    const size_t need = jt_kde_grid_cutoff_workspace_size(nx, ny, nz);
    JT_REQUIRE(need > 0, "bad sizes");
    Arena arena(ws, ws_bytes);
    float* partial = arena.take<float>((need - 256) / sizeof(float));
    if (!partial || !arena.ok()) return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", need, ws_bytes);
    return kde_cutoff_run(grid, n_total, gx, nx, gy, ny, gz, nz, h, cutoff, out, partial, stream);
}

// The tile count and the particle chunks of an all-pairs 3-D evaluation.
static void kde3d_all_tiling(int64_t n, int nx, int ny, int nz, long long& tiles, int& nchunk, long long& chunk) {
    tiles = (long long)((ny + KDE_TY - 1) / KDE_TY) * ((nx + KDE_TX - 1) / KDE_TX) * ((nz + KDE_TZ - 1) / KDE_TZ);
    const long long rounds = (n + KDE_BATCH - 1) / KDE_BATCH;
    const long long want = std::max(1LL, std::min((long long)KDE_TARGET_BLOCKS / tiles, rounds));
    chunk = ((n + want - 1) / want + KDE_BATCH - 1) / KDE_BATCH * KDE_BATCH;
    nchunk = (int)((n + chunk - 1) / chunk);
}

extern "C" size_t jt_kde_grid3d_workspace_size(int64_t n, int32_t nx, int32_t ny, int32_t nz) {
    if (n < 1 || nx < 1 || ny < 1 || nz < 1) return 0;
    long long tiles, chunk;
    int nchunk;
    kde3d_all_tiling(n, nx, ny, nz, tiles, nchunk, chunk);
    return (size_t)tiles * nchunk * (KDE_TX * KDE_TY * KDE_TZ) * sizeof(float) + 256;
}

extern "C" int jt_kde_grid3d(const float* P, int64_t n, const float* gx, int32_t nx, const float* gy, int32_t ny,
                             const float* gz, int32_t nz, float h, float* out, void* ws, size_t ws_bytes,
                             void* stream) {
    JT_REQUIRE(n >= 1 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad grid sizes");
    JT_REQUIRE(h > 0.0f, "bandwidth must be positive");
    JT_REQUIRE(P && gx && gy && gz && out, "NULL pointer");
    long long tiles, chunk;
    int nchunk;
    kde3d_all_tiling(n, nx, ny, nz, tiles, nchunk, chunk);
    JT_REQUIRE(tiles * nchunk < (1LL << 31), "grid too large");
    const size_t need = jt_kde_grid3d_workspace_size(n, nx, ny, nz);
    Arena arena(ws, ws_bytes);
    float* partial = arena.take<float>((need - 256) / sizeof(float));
    if (!partial || !arena.ok()) return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", need, ws_bytes);
    KGrid g{};                                        // no hash grid: the blocks read particle chunks
    g.inv = 1.0f;
    g.dims[0] = g.dims[1] = g.dims[2] = 1;
    const double hh = (double)h;
    const float inv_h2 = (float)(1.0 / (hh * hh));                 // kde.py:251
    const double norm = pow(2.0 * M_PI, -1.5);                      // kernels.py:91, d = 3
    hipLaunchKernelGGL(k_kde_tile_mfma, dim3((unsigned)(tiles * nchunk)), dim3(256), 0, (hipStream_t)stream,
                       gx, nx, gy, ny, gz, nz, g, 0.0f, inv_h2, norm / (double)n, nchunk, (float*)nullptr, partial,
                       P, (int)n, (int)chunk);
    JT_HIP(hipGetLastError());
    const size_t npts = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_kde_reduce, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)partial, nx, ny, nz, nchunk, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

// The tile count and the particle chunks of a 2-D grid evaluation.
static void kde2d_tiling(int64_t n, int nx, int ny, long long& tiles, int& nchunk, long long& chunk) {
    tiles = (long long)((ny + K2_TY - 1) / K2_TY) * ((nx + K2_TX - 1) / K2_TX);
    const long long passes = (n + K2_P - 1) / K2_P;
    long long want = std::max(1LL, std::min((long long)K2_TARGET_BLOCKS / tiles, passes));
    chunk = ((n + want - 1) / want + K2_P - 1) / K2_P * K2_P;      // whole passes
    nchunk = (int)((n + chunk - 1) / chunk);
}

extern "C" size_t jt_kde_grid2d_workspace_size(int64_t n, int32_t nx, int32_t ny) {
    if (n < 1 || nx < 1 || ny < 1) return 0;
    long long tiles, chunk;
    int nchunk;
    kde2d_tiling(n, nx, ny, tiles, nchunk, chunk);
    return (size_t)tiles * nchunk * (K2_TY * K2_TX) * sizeof(float) + 256;
}

extern "C" int jt_kde_grid2d(const float* P, int64_t n, const float* gx, int32_t nx, const float* gy, int32_t ny,
                             float h, float* out, void* ws, size_t ws_bytes, void* stream) {
    JT_REQUIRE(n >= 1 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(nx >= 1 && ny >= 1 && (int64_t)nx * ny < (1LL << 31), "bad grid sizes");
    JT_REQUIRE(h > 0.0f, "bandwidth must be positive");
    JT_REQUIRE(P && gx && gy && out, "NULL pointer");
    long long tiles, chunk;
    int nchunk;
    kde2d_tiling(n, nx, ny, tiles, nchunk, chunk);
    JT_REQUIRE(tiles * nchunk < (1LL << 31), "grid too large");
    const size_t need = jt_kde_grid2d_workspace_size(n, nx, ny);
    Arena arena(ws, ws_bytes);
    float* partial = arena.take<float>((need - 256) / sizeof(float));
    if (!partial || !arena.ok()) return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", need, ws_bytes);
    const double hh = (double)h;
    const float inv_h2 = (float)(1.0 / (hh * hh));                 // kde.py:251
    const double norm = 1.0 / (2.0 * M_PI);                         // kernels.py:91, d = 2
    hipLaunchKernelGGL(k_kde2d_tile, dim3((unsigned)(tiles * nchunk)), dim3(256), 0, (hipStream_t)stream,
                       P, (int)n, (int)chunk, nchunk, gx, nx, gy, ny, -0.5f * inv_h2, norm / (double)n, partial);
    JT_HIP(hipGetLastError());
    const size_t npts = (size_t)nx * ny;
    hipLaunchKernelGGL(k_kde2d_reduce, dim3((unsigned)((npts + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float*)partial, nx, ny, nchunk, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
