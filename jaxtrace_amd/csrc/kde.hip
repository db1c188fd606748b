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

// k_kde_tile_mfma: the same sum with 16 x 16 x 16 tiles on the matrix cores.
// For a tile and a batch of particles p the densities are a matrix product,
//   D[(x, y), z] += sum_p (ex[x, p] * ey[y, p]) * ez[p, z],
// with the separable factors e*[., p] = exp(-0.5 d^2 / h^2) per tile
// coordinate.  A 16^3 tile shares the 2R halo of its box among 4x the
// points of an 8 x 8 x 16 tile (the VALU kernel this replaced).  Each of the 4 waves owns 4 of the
// tile's 16 y rows: per 4 particles it issues one v_mfma_f32_16x16x4_f32 per
// row (A: 16 x by 4 particles, B: 4 particles by 16 z), so a pair costs an
// MFMA quarter-slot instead of a VALU multiply-add.  The fp32 MFMA sums run
// over one pass of staged particles and are then added in float64.
typedef float floatx4 __attribute__((ext_vector_type(4)));
#define KDE_T 16
#define KDE_BATCH 256                                 // particles read per round (one per thread)
#define KDE_SLOTS 240                                 // particles per MFMA pass (60 k-steps)
#define KDE_STAGE (KDE_SLOTS + KDE_BATCH - 1)         // staged positions at most
#ifndef KDE_TARGET_BLOCKS
#define KDE_TARGET_BLOCKS 32768
#endif
#ifndef KDE_MAX_SPLIT
#define KDE_MAX_SPLIT 16
#endif

__global__ void __launch_bounds__(256) k_kde_zero(float* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = 0.0f;
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
    __shared__ float ex[KDE_SLOTS][KDE_T];
    __shared__ float4 ey[KDE_SLOTS][KDE_T / 4];
    __shared__ float ez[KDE_SLOTS][KDE_T];
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
    const float nh = -0.5f * inv_h2;
    const int col = lane & 15, kq = lane >> 4;   // MFMA operand lane: A row / B column, and k
    __shared__ int wcount[2][4];
    __shared__ float sx[KDE_STAGE], sy[KDE_STAGE], sz[KDE_STAGE];
    double accd[4][4];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int q = 0; q < 4; ++q) accd[r][q] = 0.0;
    // Sum the first n (<= KDE_SLOTS) staged particles: their factors (one
    // particle per thread), then the MFMA k-steps over them.  Block-uniform.
    auto consume = [&](int n) {
        const int n4 = (n + 3) & ~3;              // padded with zero factors to whole k-steps
        if (tid < n4) {
            const bool real = tid < n;
            const float px = real ? sx[tid] : 0.0f, py = real ? sy[tid] : 0.0f, pz = real ? sz[tid] : 0.0f;
            float fx[KDE_T], fy[KDE_T], fz[KDE_T];
#pragma unroll
            for (int i = 0; i < KDE_T; ++i) {
                const float dx = xs[i] - px, dy = ys[i] - py, dz = zs[i] - pz;
                fx[i] = real ? __expf(nh * (dx * dx)) : 0.0f;
                fy[i] = __expf(nh * (dy * dy));
                fz[i] = __expf(nh * (dz * dz));
            }
#pragma unroll
            for (int i = 0; i < KDE_T; i += 4) {
                *reinterpret_cast<float4*>(&ex[tid][i]) = make_float4(fx[i], fx[i + 1], fx[i + 2], fx[i + 3]);
                ey[tid][i / 4] = make_float4(fy[i], fy[i + 1], fy[i + 2], fy[i + 3]);
                *reinterpret_cast<float4*>(&ez[tid][i]) = make_float4(fz[i], fz[i + 1], fz[i + 2], fz[i + 3]);
            }
        }
        __syncthreads();
        floatx4 acc[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[r] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        int p0 = 0;
        for (; p0 + 16 <= n4; p0 += 16) {             // 4 k-steps: their LDS reads first
            float ax[4], bz[4];
            float4 y4[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int p = p0 + 4 * u + kq;
                ax[u] = ex[p][col];
                bz[u] = ez[p][col];
                y4[u] = ey[p][wv];                     // this wave's rows 4 wv .. 4 wv + 3
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[u] * y4[u].x, bz[u], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[u] * y4[u].y, bz[u], acc[1], 0, 0, 0);
                acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[u] * y4[u].z, bz[u], acc[2], 0, 0, 0);
                acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax[u] * y4[u].w, bz[u], acc[3], 0, 0, 0);
            }
        }
        for (; p0 < n4; p0 += 4) {
            const int p = p0 + kq;
            const float ax = ex[p][col], bz4 = ez[p][col];
            const float4 y4 = ey[p][wv];
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax * y4.x, bz4, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax * y4.y, bz4, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax * y4.z, bz4, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f32_16x16x4f32(ax * y4.w, bz4, acc[3], 0, 0, 0);
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
        // cells (cx, cy, c0z..c1z) are consecutive ids: their points are one run
        const int id0 = (cx * g.dims[1] + cy) * g.dims[2] + c0[2];
        const int id1 = (cx * g.dims[1] + cy) * g.dims[2] + c1[2];
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
            // keep the particles within R of the tile's box on every axis
            // (the hash cells overshoot it); each one dropped is farther
            // than R from every tile point
            const bool keep = j < e && px >= lo[0] - R && px <= hi[0] + R && py >= lo[1] - R &&
                              py <= hi[1] + R && pz >= lo[2] - R && pz <= hi[2] + R;
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
            while (staged >= KDE_SLOTS) {
                __syncthreads();                       // the staged positions are visible
                consume(KDE_SLOTS);
                // move the rest to the front
                const int rest = staged - KDE_SLOTS;
                float a = 0.f, b2 = 0.f, c2 = 0.f;
                if (tid < rest) { a = sx[KDE_SLOTS + tid]; b2 = sy[KDE_SLOTS + tid]; c2 = sz[KDE_SLOTS + tid]; }
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
            const int ix = bx * KDE_T + 4 * kq + q;
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
