// Uniform-grid seed positions (jaxtrace/tracking/seeding.py:185-247).
//
// The reference builds three float32 coordinate arrays with np.linspace,
// takes np.meshgrid(..., indexing='ij') and stacks the flattened grids as
// (N, 3) columns: seed (i, j, k) sits at row (i * ny + j) * nz + k and holds
// (xc[i], yc[j], zc[k]).  The coordinate arrays are a few hundred floats and
// come from the host (their arithmetic is NumPy's); the N x 12 B expansion is
// the HBM-bound part and runs here, one thread per output float so a wave's
// stores are 256 consecutive bytes.  Two launch shapes: one block per z line
// for long lines, a grid-stride loop over all values otherwise.
#include "jt_common.h"

namespace {

__global__ void k_grid_seeds(const float* __restrict__ xc, const float* __restrict__ yc,
                             const float* __restrict__ zc, unsigned ny, unsigned nz, unsigned long long n_vals,
                             float* __restrict__ out) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long e = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; e < n_vals;
         e += stride) {
        const unsigned long long row = e / 3u;
        const unsigned comp = (unsigned)(e - row * 3u);
        const unsigned long long ij = row / nz;
        const unsigned k = (unsigned)(row - ij * nz);
        const unsigned long long i = ij / ny;
        const unsigned j = (unsigned)(ij - i * ny);
        out[e] = comp == 0 ? xc[i] : (comp == 1 ? yc[j] : zc[k]);
    }
}

// One (i, j) line of nz seeds (3 * nz consecutive floats) per block: the
// block splits its line index once, and a lane's only per-value arithmetic is
// a division by the constant 3.  Used when lines are long enough to fill a
// block (nz >= 64).
// V4: the line length is a multiple of 4 floats and out is 16-byte aligned,
// so each lane writes 4 consecutive values with one 16-byte store.
template <bool V4>
__global__ void k_grid_seed_lines(const float* __restrict__ xc, const float* __restrict__ yc,
                                  const float* __restrict__ zc, unsigned ny, unsigned nz, unsigned long long lines,
                                  float* __restrict__ out) {
    const unsigned lv = 3u * nz;
    for (unsigned long long line = blockIdx.x; line < lines; line += gridDim.x) {
        const unsigned long long i = line / ny;
        const unsigned j = (unsigned)(line - i * ny);
        const float x = xc[i], y = yc[j];
        float* o = out + line * lv;
        if constexpr (V4) {
            for (unsigned q = threadIdx.x; q < lv / 4u; q += blockDim.x) {
                float v[4];
#pragma unroll
                for (unsigned u = 0; u < 4u; ++u) {
                    const unsigned e = 4u * q + u;
                    const unsigned k = e / 3u;
                    const unsigned comp = e - 3u * k;
                    v[u] = comp == 0 ? x : (comp == 1 ? y : zc[k]);
                }
                ((float4*)o)[q] = make_float4(v[0], v[1], v[2], v[3]);
            }
        } else {
            for (unsigned e = threadIdx.x; e < lv; e += blockDim.x) {
                const unsigned k = e / 3u;
                const unsigned comp = e - 3u * k;
                o[e] = comp == 0 ? x : (comp == 1 ? y : zc[k]);
            }
        }
    }
}

}  // namespace

extern "C" int jt_grid_seeds(const float* xc, int32_t nx, const float* yc, int32_t ny, const float* zc, int32_t nz,
                             float* out, void* stream) {
    JT_REQUIRE(nx >= 0 && ny >= 0 && nz >= 0, "negative resolution (%d, %d, %d)", nx, ny, nz);
    const long long n = (long long)nx * ny * nz;
    JT_REQUIRE(n < (1LL << 40), "too many seeds");
    if (n == 0) return JT_OK;
    JT_REQUIRE(xc && yc && zc && out, "NULL pointer");
    const unsigned long long n_vals = (unsigned long long)n * 3u;
    if (nz >= 64) {
        const unsigned long long lines = (unsigned long long)nx * ny;
        const unsigned blocks = (unsigned)(lines < (1ull << 20) ? lines : (1ull << 20));
        if (nz % 4 == 0 && ((uintptr_t)out & 15u) == 0)
            hipLaunchKernelGGL(k_grid_seed_lines<true>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, xc, yc, zc,
                               (unsigned)ny, (unsigned)nz, lines, out);
        else
            hipLaunchKernelGGL(k_grid_seed_lines<false>, dim3(blocks), dim3(256), 0, (hipStream_t)stream, xc, yc,
                               zc, (unsigned)ny, (unsigned)nz, lines, out);
        JT_HIP(hipGetLastError());
        return JT_OK;
    }
    // about 8 waves per SIMD-slot's worth of blocks over 256 CUs, grid-stride beyond
    const unsigned long long want = (n_vals + 255) / 256;
    const unsigned blocks = (unsigned)(want < 65536ull ? want : 65536ull);
    hipLaunchKernelGGL(k_grid_seeds, dim3(blocks), dim3(256), 0, (hipStream_t)stream, xc, yc, zc, (unsigned)ny,
                       (unsigned)nz, n_vals, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
