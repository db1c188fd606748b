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

// ---- random_seeds: jax.random.uniform bits, then lo + u (hi - lo) -----------
// Threefry-2x32 with 20 rounds (Salmon et al., SC11): rotations 13 15 26 6 /
// 17 29 16 24, key parity constant 0x1BD11BDA, a key injection after every
// four rounds.
__device__ __forceinline__ unsigned rotl32(unsigned x, int r) { return (x << r) | (x >> (32 - r)); }

__device__ __forceinline__ void threefry2x32(unsigned k0, unsigned k1, unsigned& x0, unsigned& x1) {
    const unsigned ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
    const int rot[8] = {13, 15, 26, 6, 17, 29, 16, 24};
    x0 += ks[0];
    x1 += ks[1];
#pragma unroll
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl32(x1, rot[r & 7]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            const int s = (r + 1) >> 2;
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + (unsigned)s;
        }
    }
}

// The m = 3n words of jax.random.uniform(key, (n, 3)) laid out as JAX's
// non-partitionable random_bits: iota(m) (one 0 appended when m is odd) split
// into halves x0 | x1 of h words; block i hashes (x0[i], x1[i]) and gives
// word i and word h + i.  A word becomes the float in [1, 2) with its top 23
// bits as mantissa, minus 1 (seeding.py:167-172 then scale to the bounds).
// m < 2^32 (checked on the host), so word indices are 32-bit: the component
// i % 3 is a multiply-high, not a 64-bit division.
__global__ void __launch_bounds__(256)
k_random_seeds(unsigned k0, unsigned k1, unsigned m, unsigned h, float lo0, float lo1, float lo2, float w0,
               float w1, float w2, float* __restrict__ out) {
    const float lo[3] = {lo0, lo1, lo2}, w[3] = {w0, w1, w2};
    const unsigned stride = gridDim.x * blockDim.x;
    for (unsigned i = blockIdx.x * blockDim.x + threadIdx.x; i < h; i += stride) {
        const unsigned j = h + i;                       // < m + 1 <= 2^32 - 1 + 1: j < m or j == m (odd m)
        unsigned x0 = i, x1 = j < m ? j : 0u;
        threefry2x32(k0, k1, x0, x1);
        const float u0 = __uint_as_float((x0 >> 9) | 0x3f800000u) - 1.0f;
        const unsigned c0 = i % 3u;
        out[i] = lo[c0] + u0 * w[c0];
        if (j < m) {
            const float u1 = __uint_as_float((x1 >> 9) | 0x3f800000u) - 1.0f;
            const unsigned c1 = j % 3u;
            out[j] = lo[c1] + u1 * w[c1];
        }
    }
}

// ---- _generate_inlet_grid_particles: the inlet-face grid at given indices ----
// out[i] = (inlet_pos on the flow axis; c0[idx / n1], c1[idx % n1] on the
// two face axes in increasing axis order): np.meshgrid(c0, c1, 'ij') raveled.
__global__ void __launch_bounds__(256)
k_inlet_grid(const int* __restrict__ idx, int n, int axis, float inlet_pos, const float* __restrict__ c0, int n0,
             const float* __restrict__ c1, int n1, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int g = idx[i];
    const int a0 = axis == 0 ? 1 : 0, a1 = axis == 2 ? 1 : 2;
    float p[3];
    p[axis] = inlet_pos;
    if (g >= 0 && g < n0 * n1) {
        p[a0] = c0[g / n1];
        p[a1] = c1[g % n1];
    } else {                                            // an index off the face grid: NaN, no read
        p[a0] = p[a1] = __int_as_float(0x7fc00000);
    }
    out[3 * (size_t)i] = p[0];
    out[3 * (size_t)i + 1] = p[1];
    out[3 * (size_t)i + 2] = p[2];
}

}  // namespace

extern "C" int jt_random_seeds(int64_t n, uint64_t rng_seed, const float* lo, const float* hi, float* out,
                               void* stream) {
    JT_REQUIRE(n >= 0 && n < (1LL << 40), "n out of range");
    if (n == 0) return JT_OK;
    JT_REQUIRE(lo && hi && out, "NULL pointer");
    const unsigned long long m = (unsigned long long)n * 3u;
    JT_REQUIRE(m <= 0xffffffffull, "at most 2^32 - 1 random words per call (n <= 1431655765)");
    const unsigned long long h = (m + 1) / 2;
    float w[3];
    for (int a = 0; a < 3; ++a) w[a] = hi[a] - lo[a];                 // seeding.py:172, (hi - lo) first
    const unsigned k0 = (unsigned)(rng_seed >> 32), k1 = (unsigned)(rng_seed & 0xffffffffull);   // PRNGKey
    const unsigned long long want = (h + 255) / 256;
    const unsigned blocks = (unsigned)(want < 65536ull ? want : 65536ull);
    hipLaunchKernelGGL(k_random_seeds, dim3(blocks), dim3(256), 0, (hipStream_t)stream, k0, k1, (unsigned)m,
                       (unsigned)h, lo[0], lo[1], lo[2], w[0], w[1], w[2], out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_inlet_grid_particles(const int32_t* indices, int64_t n, int32_t axis, float inlet_position,
                                       const float* c0, int32_t n0, const float* c1, int32_t n1, float* out,
                                       void* stream) {
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(axis >= 0 && axis <= 2, "bad flow axis %d", axis);
    JT_REQUIRE(n0 >= 1 && n1 >= 1 && (long long)n0 * n1 < (1LL << 31), "bad inlet-face grid %d x %d", n0, n1);
    if (n == 0) return JT_OK;
    JT_REQUIRE(indices && c0 && c1 && out, "NULL pointer");
    hipLaunchKernelGGL(k_inlet_grid, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, indices,
                       (int)n, axis, inlet_position, c0, n0, c1, n1, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

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
