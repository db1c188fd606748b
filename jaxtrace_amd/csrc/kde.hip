// Gaussian KDE (jaxtrace/density/kde.py:244-270):
//   dens(q) = mean_n (2 pi)^(-d/2) exp(-0.5 |q - p_n|^2 * f32(1/h^2)).
//
// k_kde_exhaustive: the reference's all-pairs algorithm.  One thread per
//   query; particles stream through LDS in tiles of 256 and are reused by all
//   256 queries of the block.  Terms are accumulated in float64.
// k_kde_grid_cutoff: for the structured evaluation grid of evaluate_3d
//   (meshgrid 'xy' order, index (iy*nx + ix)*nz + iz).  A block owns an
//   8 x 8 x 16 tile of grid points (a thread: one (x, y) column, 4 z points)
//   and streams, through LDS, the particles of the hash-grid cells that
//   overlap the tile's box grown by R; every particle within R of any tile
//   point is summed (and some farther ones).  The Gaussian is separable on
//   the lattice: per particle the block evaluates exp(-0.5 d^2 / h^2) once
//   for each of the tile's 8 x, 8 y and 16 z coordinates, and a pair's term
//   is the product of its three factors (a few ulp from exp of the sum, far
//   inside the 1e-5 density bar).  Each dropped particle is farther than R
//   from the point, so each dropped term is below norm*exp(-R^2/(2h^2)).
#include "jt_common.h"

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

__global__ void __launch_bounds__(256)
k_kde_grid_cutoff(const float* __restrict__ gx, int nx, const float* __restrict__ gy, int ny,
                  const float* __restrict__ gz, int nz, KGrid g, float R, float inv_h2, double scale,
                  float* __restrict__ out) {
    __shared__ float ex[256][8];
    __shared__ float ey[256][8];
    __shared__ float4 ez[256][4];
    const int tid = threadIdx.x;
    const int tiles_z = (nz + 15) / 16, tiles_x = (nx + 7) / 8;
    int b = blockIdx.x;
    const int bz = b % tiles_z; b /= tiles_z;
    const int bx = b % tiles_x; b /= tiles_x;
    const int by = b;
    const int zq = tid & 3, txl = (tid >> 2) & 7, tyl = tid >> 5;
    const int ix = bx * 8 + txl, iy = by * 8 + tyl, iz0 = bz * 16 + zq * 4;
    // the tile's coordinates (clamped at the grid's end; clamped points are not written)
    float xs[8], ys[8], zs[16];
#pragma unroll
    for (int i = 0; i < 8; ++i) { xs[i] = gx[min(bx * 8 + i, nx - 1)]; ys[i] = gy[min(by * 8 + i, ny - 1)]; }
#pragma unroll
    for (int k = 0; k < 16; ++k) zs[k] = gz[min(bz * 16 + k, nz - 1)];
    float lo[3] = {xs[0], ys[0], zs[0]}, hi[3] = {xs[0], ys[0], zs[0]};
#pragma unroll
    for (int i = 1; i < 8; ++i) {
        lo[0] = fminf(lo[0], xs[i]); hi[0] = fmaxf(hi[0], xs[i]);
        lo[1] = fminf(lo[1], ys[i]); hi[1] = fmaxf(hi[1], ys[i]);
    }
#pragma unroll
    for (int k = 1; k < 16; ++k) { lo[2] = fminf(lo[2], zs[k]); hi[2] = fmaxf(hi[2], zs[k]); }
    int c0[3], c1[3];
    for (int a = 0; a < 3; ++a) {
        c0[a] = clampi((int)floorf((lo[a] - R - g.origin[a]) * g.inv), 0, g.dims[a] - 1);
        c1[a] = clampi((int)floorf((hi[a] + R - g.origin[a]) * g.inv), 0, g.dims[a] - 1);
    }
    const float nh = -0.5f * inv_h2;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    for (int cx = c0[0]; cx <= c1[0]; ++cx)
        for (int cy = c0[1]; cy <= c1[1]; ++cy) {
            // cells (cx, cy, c0z..c1z) are consecutive ids: their points are one run
            const int id0 = (cx * g.dims[1] + cy) * g.dims[2] + c0[2];
            const int id1 = (cx * g.dims[1] + cy) * g.dims[2] + c1[2];
            int s = 0x7fffffff, e = 0;
            for (int id = id0; id <= id1; ++id) {
                int cs = g.cstart[id], ce = g.cend[id];
                if (ce > cs) { s = min(s, cs); e = max(e, ce); }
            }
            for (int base = s; base < e; base += 256) {
                const int j = base + tid;
                // this thread's particle: its factor for each tile coordinate
                if (j < e) {
                    const float px = g.spos[(size_t)j * 3], py = g.spos[(size_t)j * 3 + 1];
                    const float pz = g.spos[(size_t)j * 3 + 2];
#pragma unroll
                    for (int i = 0; i < 8; ++i) {
                        const float dx = xs[i] - px, dy = ys[i] - py;
                        ex[tid][i] = __expf(nh * (dx * dx));
                        ey[tid][i] = __expf(nh * (dy * dy));
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        float4 v;
                        float dz;
                        dz = zs[4 * q] - pz;     v.x = __expf(nh * (dz * dz));
                        dz = zs[4 * q + 1] - pz; v.y = __expf(nh * (dz * dz));
                        dz = zs[4 * q + 2] - pz; v.z = __expf(nh * (dz * dz));
                        dz = zs[4 * q + 3] - pz; v.w = __expf(nh * (dz * dz));
                        ez[tid][q] = v;
                    }
                }
                __syncthreads();
                const int cnt = min(256, e - base);
                float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int k = 0; k < cnt; ++k) {
                    const float exy = ex[k][txl] * ey[k][tyl];
                    const float4 z4 = ez[k][zq];
                    part[0] = fmaf(exy, z4.x, part[0]);
                    part[1] = fmaf(exy, z4.y, part[1]);
                    part[2] = fmaf(exy, z4.z, part[2]);
                    part[3] = fmaf(exy, z4.w, part[3]);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) acc[q] += (double)part[q];
                __syncthreads();
            }
        }
    if (ix < nx && iy < ny)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (iz0 + q < nz) out[((size_t)iy * nx + ix) * nz + iz0 + q] = (float)(acc[q] * scale);
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
    long long blocks = (long long)((ny + 7) / 8) * ((nx + 7) / 8) * ((nz + 15) / 16);
    JT_REQUIRE(blocks < (1LL << 31), "grid too large");
    hipLaunchKernelGGL(k_kde_grid_cutoff, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       gx, nx, gy, ny, gz, nz, g, cutoff, inv_h2, norm / (double)n_total, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
