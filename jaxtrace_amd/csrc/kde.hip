// Gaussian KDE (jaxtrace/density/kde.py:244-270):
//   dens(q) = mean_n (2 pi)^(-d/2) exp(-0.5 |q - p_n|^2 * f32(1/h^2)).
//
// k_kde_exhaustive: the reference's all-pairs algorithm.  One thread per
//   query; particles stream through LDS in tiles of 256 and are reused by all
//   256 queries of the block.  Terms are accumulated in float64.
// k_kde_tile_mfma (below; jt_kde_grid_cutoff runs it): the cut-off sum on
//   the structured grid on the matrix cores, 16 x 16 x 16 tiles.
//   Each dropped particle is farther than R from the grid point, so each
//   dropped term is below norm*exp(-R^2/(2h^2)).
#include "jt_common.h"
#include <algorithm>


namespace jt {

__global__ void __launch_bounds__(256)
k_kde_exhaustive(const float* __restrict__ Q, int m, int D, const float* __restrict__ P, int n,
                 float inv_h2, double scale, float* __restrict__ out) {
    __shared__ float sp[256 * 3];
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 256 + tid;
    float q[3] = {0.f, 0.f, 0.f};
    if (i < m)
        for (int a = 0; a < D; ++a) q[a] = Q[(size_t)i * D + a];
    double acc = 0.0;
    for (int base = 0; base < n; base += 256) {
        const int j = base + tid;
        for (int a = 0; a < D; ++a) sp[tid * 3 + a] = (j < n) ? P[(size_t)j * D + a] : 0.0f;
        __syncthreads();
        const int cnt = min(256, n - base);
        for (int k = 0; k < cnt; ++k) {
            float dx = q[0] - sp[k * 3], dy = q[1] - sp[k * 3 + 1];
            float r2 = dx * dx + dy * dy;
            if (D == 3) { float dz = q[2] - sp[k * 3 + 2]; r2 = r2 + dz * dz; }
            acc += (double)expf(-0.5f * (r2 * inv_h2));
        }
        __syncthreads();
    }
    if (i < m) out[i] = (float)(acc * scale);
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

// k_kde_tile_mfma: the cut-off sum with 16 x 16 x 16 tiles on the matrix
// cores.  For a tile and a pass of particles p the densities are a matrix
// product,
//   D[(x, y), z] += sum_p (ex[x, p] * ey[y, p]) * ez[p, z],
// with the separable factors e*[., p] = exp(-0.5 d^2 / h^2) per tile
// coordinate.  Each of the 4 waves owns 4 of the tile's 16 y rows.
//
// Precision: the products run on the bfloat16 matrix cores
// (v_mfma_f32_16x16x32_bf16, 16 times the float32 MFMA rate) with each
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
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
#define KDE_T 16
#define KDE_BATCH 256                                 // particles read per round (one per thread)
#ifndef KDE_P
#define KDE_P 256                                     // particles per MFMA pass (a multiple of 32)
#endif
#define KDE_STAGE (KDE_P + KDE_BATCH - 1)             // staged positions at most
#define KDE_LDP (KDE_P + 8)                           // padded row length of the factor tables
#ifndef KDE_TARGET_BLOCKS
#define KDE_TARGET_BLOCKS 32768
#endif
#ifndef KDE_MAX_SPLIT
#define KDE_MAX_SPLIT 16
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
// tile's cell columns (cx, cy), those with index = b % nsplit modulo nsplit,
// adding its partial densities into out (zeroed first).  The particle
// counts of the tiles' boxes differ widely on clustered data, so one block
// per tile leaves a tail of heavy blocks running alone.
__global__ void __launch_bounds__(256)
k_kde_tile_mfma(const float* __restrict__ gx, int nx, const float* __restrict__ gy, int ny,
                const float* __restrict__ gz, int nz, KGrid g, float R, float inv_h2, double scale,
                int nsplit, float* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) float exs[KDE_T][KDE_LDP];        // x factors [x][p]
    __shared__ __attribute__((aligned(16))) float eys[KDE_T][KDE_LDP];        // y factors [y][p]
    __shared__ __attribute__((aligned(16))) __bf16 ezh[KDE_T][KDE_LDP];       // z factors [z][p], hi half
    __shared__ __attribute__((aligned(16))) __bf16 ezl[KDE_T][KDE_LDP];       // lo half
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int tiles_z = (nz + KDE_T - 1) / KDE_T, tiles_x = (nx + KDE_T - 1) / KDE_T;
    const int part = blockIdx.x % nsplit;
    int b = blockIdx.x / nsplit;
    const int bz = b % tiles_z; b /= tiles_z;
    const int bx = b % tiles_x; b /= tiles_x;
    const int by = b;
    // the tile's coordinates (clamped at the grid's end; clamped points are not written)
    float xs[KDE_T], ys[KDE_T], zs[KDE_T];
#pragma unroll
    for (int i = 0; i < KDE_T; ++i) {
        xs[i] = gx[min(bx * KDE_T + i, nx - 1)];
        ys[i] = gy[min(by * KDE_T + i, ny - 1)];
        zs[i] = gz[min(bz * KDE_T + i, nz - 1)];
    }
    float lo[3] = {xs[0], ys[0], zs[0]}, hi[3] = {xs[0], ys[0], zs[0]};
#pragma unroll
    for (int i = 1; i < KDE_T; ++i) {
        lo[0] = fminf(lo[0], xs[i]); hi[0] = fmaxf(hi[0], xs[i]);
        lo[1] = fminf(lo[1], ys[i]); hi[1] = fmaxf(hi[1], ys[i]);
        lo[2] = fminf(lo[2], zs[i]); hi[2] = fmaxf(hi[2], zs[i]);
    }
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = clampi((int)floorf((lo[a] - R - g.origin[a]) * g.inv), 0, g.dims[a] - 1);
        c1[a] = clampi((int)floorf((hi[a] + R - g.origin[a]) * g.inv), 0, g.dims[a] - 1);
    }
    const float R2 = R * R;
    const float cell = 1.0f / g.inv;
    const float nh = -0.5f * inv_h2;
    const int col = lane & 15, kg = lane >> 4;   // MFMA operand lane: A row / B column, and k group
    __shared__ int wcount[2][4];
    __shared__ float sx[KDE_STAGE], sy[KDE_STAGE], sz[KDE_STAGE];
    double accd[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) accd[r][q] = 0.0;
    // Sum the first n (<= KDE_P) staged particles: their factors (one
    // particle per thread), then the MFMA passes over them.  Block-uniform.
    auto consume = [&](int n) {
        const int n32 = (n + 31) & ~31;            // padded with zero x factors to whole 32-particle steps
        if (tid < n32) {
            const bool real = tid < n;
            const float px = real ? sx[tid] : 0.0f, py = real ? sy[tid] : 0.0f, pz = real ? sz[tid] : 0.0f;
#pragma unroll
            for (int i = 0; i < KDE_T; i += 2) {
                const float dx0 = xs[i] - px, dx1 = xs[i + 1] - px;
                const float dy0 = ys[i] - py, dy1 = ys[i + 1] - py;
                const float dz0 = zs[i] - pz, dz1 = zs[i + 1] - pz;
                exs[i][tid] = real ? __expf(nh * (dx0 * dx0)) : 0.0f;
                exs[i + 1][tid] = real ? __expf(nh * (dx1 * dx1)) : 0.0f;
                eys[i][tid] = __expf(nh * (dy0 * dy0));
                eys[i + 1][tid] = __expf(nh * (dy1 * dy1));
                const float f0 = __expf(nh * (dz0 * dz0)), f1 = __expf(nh * (dz1 * dz1));
                const __bf16 h0 = (__bf16)f0, h1 = (__bf16)f1;
                ezh[i][tid] = h0;
                ezh[i + 1][tid] = h1;
                ezl[i][tid] = (__bf16)(f0 - (float)h0);
                ezl[i + 1][tid] = (__bf16)(f1 - (float)h1);
            }
        }
        __syncthreads();
        floatx4 acc[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int p0 = 0; p0 < n32; p0 += 32) {
            const int pk = p0 + 8 * kg;               // this lane's 8 particles of the step
            const float4 xa = *reinterpret_cast<const float4*>(&exs[col][pk]);
            const float4 xb = *reinterpret_cast<const float4*>(&exs[col][pk + 4]);
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&ezh[col][pk]);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&ezl[col][pk]);
            const float xv[8] = {xa.x, xa.y, xa.z, xa.w, xb.x, xb.y, xb.z, xb.w};
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int y = 4 * wv + r;
                const float4 ya = *reinterpret_cast<const float4*>(&eys[y][pk]);
                const float4 yb = *reinterpret_cast<const float4*>(&eys[y][pk + 4]);
                const float yv[8] = {ya.x, ya.y, ya.z, ya.w, yb.x, yb.y, yb.z, yb.w};
                float av[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) av[j] = xv[j] * yv[j];
                bf16x8 ah, al;
                split8(av, ah, al);
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[r], 0, 0, 0);
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[r], 0, 0, 0);
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[r], 0, 0, 0);
                acc[r] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bl, acc[r], 0, 0, 0);
            }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int q = 0; q < 4; ++q) accd[r][q] += (double)acc[r][q];
        __syncthreads();                               // factors and staged positions are rewritten next
    };
    int staged = 0, round = 0;                         // block-uniform
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
        int s = 0x7fffffff, e = 0;
        for (int id = id0; id <= id1; ++id) {
            const int cs = g.cstart[id], ce = g.cend[id];
            if (ce > cs) { s = min(s, cs); e = max(e, ce); }
        }
        if (s >= e) continue;                          // no particles in these cells (s is still INT_MAX)
        // this thread's particle of the next round, loaded one round ahead
        float qx = 0.f, qy = 0.f, qz = 0.f;
        if ((long long)s + tid < (long long)e) {
            qx = g.spos[(size_t)(s + tid) * 3]; qy = g.spos[(size_t)(s + tid) * 3 + 1];
            qz = g.spos[(size_t)(s + tid) * 3 + 2];
        }
        for (int base = s; base < e; base += KDE_BATCH) {
            const int j = base + tid;
            const float px = qx, py = qy, pz = qz;
            if ((long long)j + KDE_BATCH < (long long)e) {
                qx = g.spos[(size_t)(j + KDE_BATCH) * 3]; qy = g.spos[(size_t)(j + KDE_BATCH) * 3 + 1];
                qz = g.spos[(size_t)(j + KDE_BATCH) * 3 + 2];
            }
            // keep the particles within R of the tile's box: each one
            // dropped is farther than R from every tile point
            const float dx = fmaxf(fmaxf(lo[0] - px, px - hi[0]), 0.0f);
            const float dy = fmaxf(fmaxf(lo[1] - py, py - hi[1]), 0.0f);
            const float dz = fmaxf(fmaxf(lo[2] - pz, pz - hi[2]), 0.0f);
            const bool keep = j < e && (dx * dx + dy * dy) + dz * dz <= R2;
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
    }
    if (staged > 0) {
        __syncthreads();
        consume(staged);
    }
    // D of row r (y = 4 wv + r): lane holds x = 4 (lane / 16) + q, z = lane % 16
    const int iz = bz * KDE_T + col;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int iy = by * KDE_T + 4 * wv + r;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int ix = bx * KDE_T + 4 * kg + q;
            if (ix < nx && iy < ny && iz < nz && accd[r][q] != 0.0)
                unsafeAtomicAdd(&out[((size_t)iy * nx + ix) * nz + iz], (float)(accd[r][q] * scale));
        }
    }
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
                       Q, (int)m, dim, P, (int)n, inv_h2, norm / (double)n, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_kde_grid_cutoff(const jt_hashgrid* grid, int64_t n_total, const float* gx, int32_t nx,
                                  const float* gy, int32_t ny, const float* gz, int32_t nz, float h,
                                  float cutoff, float* out, void* stream) {
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
    const long long tiles = (long long)((ny + KDE_T - 1) / KDE_T) * ((nx + KDE_T - 1) / KDE_T) * ((nz + KDE_T - 1) / KDE_T);
    // about KDE_TARGET_BLOCKS blocks (the chip holds 768 at a time), at most KDE_MAX_SPLIT per tile
    const int nsplit = (int)std::max(1LL, std::min((long long)KDE_MAX_SPLIT, (long long)KDE_TARGET_BLOCKS / tiles));
    JT_REQUIRE(tiles * nsplit < (1LL << 31), "grid too large");
    const size_t npts = (size_t)nx * ny * nz;
    hipLaunchKernelGGL(k_kde_zero, dim3((unsigned)std::min<size_t>((npts + 255) / 256, 4096)), dim3(256), 0,
                       (hipStream_t)stream, out, npts);
    JT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_kde_tile_mfma, dim3((unsigned)(tiles * nsplit)), dim3(256), 0, (hipStream_t)stream,
                       gx, nx, gy, ny, gz, nz, g, cutoff, inv_h2, norm / (double)n_total, nsplit, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
