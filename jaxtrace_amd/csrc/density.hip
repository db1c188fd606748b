// Hash-grid neighbour search and SPH density (jaxtrace/density/neighbors.py,
// jaxtrace/density/sph.py).
//
// Build: a stable counting sort of the points by dense cell key, by passes of
// at most 8 key bits (equal keys keep increasing point index, as the
// reference's stable argsort); the last pass writes the order and the
// positions in cell order, so the candidates of one cell are contiguous in
// memory; then the per-cell [start, end) table.
//
// Query: one 64-lane wavefront per query point.  The 27 (2-D: 9) stencil
// probes are visited in the reference's lexicographic order; lane s handles
// slot s of the probed cell (per_cell_cap slots, 64 at a time), so candidate
// c = probe * cap + slot exactly as neighbors.py:145-151 flattens them.
// Distances, validity (slot < count and d <= r) and the stable "first K by
// distance" selection follow neighbors.py:153-165 bit for bit.
#include "jt_common.h"
#include "cell_sort.h"
#include <algorithm>
#include <cmath>

namespace jt {

struct HGDev {
    int dim, key_mode;
    float origin[3];
    float inv;
    int dims[3];
    int n;
    const int* order;
    const float* spos;      // sorted positions (n, dim)
    const int* cstart;
    const int* cend;
};

// Order-preserving float <-> uint32 map for atomic min/max.
__device__ __forceinline__ unsigned f2ord(float f) {
    unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__host__ __device__ __forceinline__ float ord2f(unsigned u) {
    unsigned v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &v, 4);
    return f;
}

__global__ void k_fill_u32(unsigned* p, int n, unsigned v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Per-axis min and max of the points into out[0..2] / out[3..5] (ordered
// uint32 encodings): each thread folds a grid-stride slice (16-byte loads of
// 4 points (3-D) or 2 points (2-D) when the array is 16-byte aligned), each
// wave folds its lanes with shuffles, the block folds its waves in LDS and
// writes its six words to out[6 * block ...]; k_decode_bounds folds the
// blocks' words.  (Global atomics on the one shared cache line cost about
// 20 ns each, serialised: with 1,024 blocks they took longer than the read.)
__global__ void __launch_bounds__(256) k_bounds(const float* __restrict__ P, int n, int dim, int vec, unsigned* out) {
    __shared__ unsigned red[4][6];
    unsigned lmin[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, lmax[3] = {0u, 0u, 0u};
    auto fold = [&](int a, float v) {
        const unsigned u = f2ord(v);
        lmin[a] = min(lmin[a], u);
        lmax[a] = max(lmax[a], u);
    };
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    const int stride = gridDim.x * blockDim.x;
    int done = 0;                                       // points covered by the vector loop
    if (vec && dim == 3) {
        const float4* P4 = reinterpret_cast<const float4*>(P);
        const int ng = n >> 2;
#pragma unroll 4
        for (int gi = tid; gi < ng; gi += stride) {
            const float4 u = P4[3 * gi], v = P4[3 * gi + 1], w = P4[3 * gi + 2];
            fold(0, u.x); fold(1, u.y); fold(2, u.z); fold(0, u.w);
            fold(1, v.x); fold(2, v.y); fold(0, v.z); fold(1, v.w);
            fold(2, w.x); fold(0, w.y); fold(1, w.z); fold(2, w.w);
        }
        done = ng << 2;
    } else if (vec && dim == 2) {
        const float4* P4 = reinterpret_cast<const float4*>(P);
        const int ng = n >> 1;
#pragma unroll 4
        for (int gi = tid; gi < ng; gi += stride) {
            const float4 u = P4[gi];
            fold(0, u.x); fold(1, u.y); fold(0, u.z); fold(1, u.w);
        }
        done = ng << 1;
    }
    for (int i = done + tid; i < n; i += stride)
        for (int a = 0; a < dim; ++a) fold(a, P[(size_t)i * dim + a]);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        for (int o = 32; o > 0; o >>= 1) {
            lmin[a] = min(lmin[a], (unsigned)__shfl_xor((int)lmin[a], o));
            lmax[a] = max(lmax[a], (unsigned)__shfl_xor((int)lmax[a], o));
        }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) { red[wv][a] = lmin[a]; red[wv][3 + a] = lmax[a]; }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int a = threadIdx.x;
        unsigned mn = red[0][a], mx = red[0][3 + a];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { mn = min(mn, red[w][a]); mx = max(mx, red[w][3 + a]); }
        out[blockIdx.x * 6 + a] = mn;
        out[blockIdx.x * 6 + 3 + a] = mx;
    }
}

// Fold the blocks' partial bounds (6 words each) and decode them to floats.
__global__ void __launch_bounds__(256) k_decode_bounds(const unsigned* in, int nblocks, float* out, int dim) {
    __shared__ unsigned red[4][6];
    unsigned v[6] = {0xffffffffu, 0xffffffffu, 0xffffffffu, 0u, 0u, 0u};
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], in[b * 6 + a]);
            v[3 + a] = max(v[3 + a], in[b * 6 + 3 + a]);
        }
    for (int o = 32; o > 0; o >>= 1)
        for (int a = 0; a < 3; ++a) {
            v[a] = min(v[a], (unsigned)__shfl_xor((int)v[a], o));
            v[3 + a] = max(v[3 + a], (unsigned)__shfl_xor((int)v[3 + a], o));
        }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 6; ++k) red[wv][k] = v[k];
    __syncthreads();
    if (threadIdx.x < dim) {
        const int a = threadIdx.x;
        unsigned mn = red[0][a], mx = red[0][3 + a];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { mn = min(mn, red[w][a]); mx = max(mx, red[w][3 + a]); }
        out[a] = ord2f(mn);
        out[dim + a] = ord2f(mx);
    }
}

// neighbors.py:70: floor((P - origin) * inv) as int32.
__device__ __forceinline__ int cell_coord(float p, float o, float inv) {
    float f = floorf((p - o) * inv);
    if (f != f) return 0;
    f = fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);   // keep the int conversion defined
    return (int)f;
}

// Dense cell id of a point (neighbors.py:70-91 with the reference's packing
// replaced by a dense index over the grid's dims), clamped to the table for
// inputs outside the bounds the grid was sized from.
__device__ __forceinline__ unsigned hg_key(const HGDev& g, float x, float y, float z, unsigned nc) {
    const int c0 = cell_coord(x, g.origin[0], g.inv), c1 = cell_coord(y, g.origin[1], g.inv);
    long long k;
    if (g.dim == 3) {
        const int c2 = cell_coord(z, g.origin[2], g.inv);
        k = (g.key_mode == JT_KEY_X64) ? ((long long)c0 * g.dims[1] + c1) * g.dims[2] + c2
                                       : (long long)c1 * g.dims[2] + c2;
    } else {
        k = (g.key_mode == JT_KEY_X64) ? (long long)c0 * g.dims[1] + c1 : (long long)c1;
    }
    k = k < 0 ? 0 : k;
    return (unsigned)(k < (long long)nc ? k : (long long)nc - 1);
}

__global__ void __launch_bounds__(256)
k_cell_table(const unsigned* __restrict__ skeys, int n, int* __restrict__ cstart, int* __restrict__ cend) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) cstart[k] = i;
    if (i == n - 1 || skeys[i + 1] != k) cend[k] = i + 1;
}

struct alignas(4) Rec3 { float v[3]; };
struct alignas(4) Rec2 { float v[2]; };

// ---- build: stable LSD counting sort of the points by cell ------------------
// The reference's stable argsort of the cell keys (neighbors.py:93-95), by
// passes of at most 8 key bits (key bits = ceil(log2 n_cells)).  A pass:
//   k_hg_count    per tile of HG_TILE points, the histogram of the digit;
//   k_hg_scan_*   exclusive scan of the counts (digit-major, tile-minor), so
//                 entry (d, t) is where tile t's points of digit d start;
//   k_hg_scatter  per tile: each wave takes 512 consecutive points in
//                 chunks of 64; ballots on the digit bits give each lane the
//                 lanes with its digit, and a per-wave running count per digit
//                 in LDS carries the rank from chunk to chunk; the waves'
//                 counts are scanned per digit, the tile's records are
//                 reordered by digit in LDS, and each digit's run is written
//                 at its global offset with consecutive lanes on consecutive
//                 records.
// Points travel as (x, y, z, index) records and the key is recomputed from
// the position in every pass (the same float32 operations), so no key array
// is read.  The last pass writes order, the sorted positions and the sorted
// keys (for the cell table) directly, so there is no separate gather.
constexpr int HG_BLOCK = 512;
constexpr int HG_ITEMS = 8;                           // points per thread
constexpr int HG_TILE = HG_BLOCK * HG_ITEMS;          // 4,096 points a tile
constexpr int HG_WAVES = HG_BLOCK / 64;
constexpr int HG_WAVE_SPAN = HG_TILE / HG_WAVES;      // consecutive points a wave ranks
constexpr int HG_SCAN_ITEMS = 8;
constexpr int HG_SCAN_SPAN = HG_BLOCK * HG_SCAN_ITEMS;

template <int DIM, bool FIRST>
__device__ __forceinline__ float4 hg_load(const float* __restrict__ P, const float4* __restrict__ rin, int i) {
    if constexpr (!FIRST) {
        return rin[i];
    } else if constexpr (DIM == 3) {
        const Rec3 v = reinterpret_cast<const Rec3*>(P)[i];
        return make_float4(v.v[0], v.v[1], v.v[2], __int_as_float(i));
    } else {
        const Rec2 v = reinterpret_cast<const Rec2*>(P)[i];
        return make_float4(v.v[0], v.v[1], 0.0f, __int_as_float(i));
    }
}

// Lanes whose digit equals this lane's (among the active lanes), from one
// ballot per digit bit.
__device__ __forceinline__ unsigned long long hg_match(unsigned d, int nbits, bool act) {
    unsigned long long m = __ballot(act);
    for (int b = 0; b < nbits; ++b) {
        const bool bit = (d >> b) & 1u;
        const unsigned long long s = __ballot(act && bit);
        m &= bit ? s : ~s;
    }
    return m;
}

// Exclusive scan of one value per thread over the block; total = the sum.
__device__ __forceinline__ unsigned hg_block_excl(unsigned v, unsigned* red, unsigned& total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned x = v;
    for (int o = 1; o < 64; o <<= 1) {
        const unsigned y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) red[wv] = x;
    __syncthreads();
    unsigned pre = 0, tot = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        const unsigned r = red[w];
        pre += w < wv ? r : 0u;
        tot += r;
    }
    __syncthreads();
    total = tot;
    return pre + x - v;
}

template <int DIM, bool FIRST>
__global__ void __launch_bounds__(HG_BLOCK)
k_hg_count(HGDev g, const float* __restrict__ P, const float4* __restrict__ rin, int shift, int dbits, int ntiles,
           unsigned nc, unsigned* __restrict__ hist) {
    __shared__ unsigned h[256];
    const int tile = blockIdx.x;
    const unsigned dmask = (1u << dbits) - 1u;
    for (int b = threadIdx.x; b < 256; b += HG_BLOCK) h[b] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int base = tile * HG_TILE + wv * HG_WAVE_SPAN;
    float4 r[HG_ITEMS];
#pragma unroll
    for (int k = 0; k < HG_ITEMS; ++k) {
        const int i = base + k * 64 + lane;
        if (i < g.n) r[k] = hg_load<DIM, FIRST>(P, rin, i);
    }
#pragma unroll
    for (int k = 0; k < HG_ITEMS; ++k) {
        const bool act = base + k * 64 + lane < g.n;
        const unsigned d = act ? (hg_key(g, r[k].x, r[k].y, r[k].z, nc) >> shift) & dmask : 0u;
        const unsigned long long peers = hg_match(d, dbits, act);
        if (act && (peers & ((1ull << lane) - 1ull)) == 0ull) atomicAdd(&h[d], (unsigned)__popcll(peers));
    }
    __syncthreads();
    for (int b = threadIdx.x; b < (1 << dbits); b += HG_BLOCK) hist[(size_t)b * ntiles + tile] = h[b];
}

__global__ void __launch_bounds__(HG_BLOCK) k_hg_scan_sums(const unsigned* __restrict__ in, int E,
                                                            unsigned* __restrict__ bsum) {
    __shared__ unsigned red[HG_BLOCK / 64];
    const int lo = blockIdx.x * HG_SCAN_SPAN;
    unsigned s = 0;
    for (int k = 0; k < HG_SCAN_ITEMS; ++k) {
        const int e = lo + k * HG_BLOCK + threadIdx.x;
        if (e < E) s += in[e];
    }
    unsigned total;
    hg_block_excl(s, red, total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

__global__ void __launch_bounds__(HG_BLOCK) k_hg_scan_apply(unsigned* __restrict__ io, int E,
                                                             const unsigned* __restrict__ bsum) {
    __shared__ unsigned red[HG_BLOCK / 64];
    unsigned pre = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += HG_BLOCK) pre += bsum[b];
    unsigned before;
    hg_block_excl(pre, red, before);                   // before = the earlier blocks' total
    const int lo = blockIdx.x * HG_SCAN_SPAN + threadIdx.x * HG_SCAN_ITEMS;
    unsigned v[HG_SCAN_ITEMS], t = 0;
#pragma unroll
    for (int k = 0; k < HG_SCAN_ITEMS; ++k) {
        v[k] = lo + k < E ? io[lo + k] : 0u;
        t += v[k];
    }
    unsigned total;
    unsigned run = before + hg_block_excl(t, red, total);
#pragma unroll
    for (int k = 0; k < HG_SCAN_ITEMS; ++k) {
        if (lo + k < E) io[lo + k] = run;
        run += v[k];
    }
}

template <int DIM, bool FIRST, bool LAST>
__global__ void __launch_bounds__(HG_BLOCK)
k_hg_scatter(HGDev g, const float* __restrict__ P, const float4* __restrict__ rin, int shift, int dbits, int ntiles,
             unsigned nc, const unsigned* __restrict__ offs, float4* __restrict__ rout, int* __restrict__ order,
             float* __restrict__ spos, unsigned* __restrict__ skeys) {
    __shared__ float4 srec[HG_TILE];
    __shared__ unsigned char sdig[HG_TILE];
    __shared__ unsigned wcnt[HG_WAVES][256];
    __shared__ unsigned bst[256], goff[256], red[HG_WAVES];
    const int tile = blockIdx.x;
    const int nb = 1 << dbits;
    const unsigned dmask = (unsigned)nb - 1u;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long lt = (1ull << lane) - 1ull;
    for (int b = threadIdx.x; b < HG_WAVES * 256; b += HG_BLOCK) (&wcnt[0][0])[b] = 0u;
    __syncthreads();
    const int base = tile * HG_TILE + wv * HG_WAVE_SPAN;
    float4 r[HG_ITEMS];
    unsigned loc[HG_ITEMS];                            // digit << 16 | rank among the wave's points of the digit
#pragma unroll
    for (int k = 0; k < HG_ITEMS; ++k) {
        const int i = base + k * 64 + lane;
        if (i < g.n) r[k] = hg_load<DIM, FIRST>(P, rin, i);
    }
#pragma unroll
    for (int k = 0; k < HG_ITEMS; ++k) {
        const bool act = base + k * 64 + lane < g.n;
        const unsigned d = act ? (hg_key(g, r[k].x, r[k].y, r[k].z, nc) >> shift) & dmask : 0u;
        const unsigned long long peers = hg_match(d, dbits, act);
        // every lane of the group reads the running count before its lowest
        // lane advances it (one wave: LDS operations stay in program order)
        const unsigned before = wcnt[wv][d];
        loc[k] = (d << 16) | (before + (unsigned)__popcll(peers & lt));
        if (act && (peers & lt) == 0ull) wcnt[wv][d] = before + (unsigned)__popcll(peers);
    }
    __syncthreads();
    unsigned tot = 0;
    if ((int)threadIdx.x < nb)
        for (int w = 0; w < HG_WAVES; ++w) {
            const unsigned c = wcnt[w][threadIdx.x];
            wcnt[w][threadIdx.x] = tot;
            tot += c;
        }
    unsigned all;
    const unsigned st = hg_block_excl((int)threadIdx.x < nb ? tot : 0u, red, all);
    if ((int)threadIdx.x < nb) {
        bst[threadIdx.x] = st;
        goff[threadIdx.x] = offs[(size_t)threadIdx.x * ntiles + tile];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < HG_ITEMS; ++k) {
        if (base + k * 64 + lane < g.n) {
            const unsigned d = loc[k] >> 16;
            const unsigned p = bst[d] + wcnt[wv][d] + (loc[k] & 0xffffu);
            srec[p] = r[k];
            sdig[p] = (unsigned char)d;
        }
    }
    __syncthreads();
    const int cnt = min(HG_TILE, g.n - tile * HG_TILE);
    for (int j = threadIdx.x; j < cnt; j += HG_BLOCK) {
        const unsigned d = sdig[j];
        const unsigned dst = goff[d] + (unsigned)j - bst[d];
        const float4 v = srec[j];
        if constexpr (LAST) {
            order[dst] = __float_as_int(v.w);
            if constexpr (DIM == 3) reinterpret_cast<Rec3*>(spos)[dst] = Rec3{{v.x, v.y, v.z}};
            else reinterpret_cast<Rec2*>(spos)[dst] = Rec2{{v.x, v.y}};
            skeys[dst] = hg_key(g, v.x, v.y, v.z, nc);
        } else {
            rout[dst] = v;
        }
    }
}

// Decode a probed cell the way the reference's packed key would match a
// present key (neighbors.py:83-91, 138-143).  Returns the dense cell id or -1.
__device__ __forceinline__ int probe_cell(const HGDev& g, int cx, int cy, int cz) {
    if (g.dim == 3) {
        if (g.key_mode == JT_KEY_X64) {
            long long key = ((long long)cx << 42) | ((long long)cy << 21) | (long long)cz;
            if (key < 0) return -1;
            long long x = key >> 42, y = (key >> 21) & 0x1fffff, z = key & 0x1fffff;
            if (x >= g.dims[0] || y >= g.dims[1] || z >= g.dims[2]) return -1;
            return (int)((x * g.dims[1] + y) * g.dims[2] + z);
        }
        int key = (int)((unsigned)cy << 21) | cz;          // int32 key, x dropped
        if (key < 0) return -1;
        int y = key >> 21, z = key & 0x1fffff;
        if (y >= g.dims[1] || z >= g.dims[2]) return -1;
        return y * g.dims[2] + z;
    }
    if (g.key_mode == JT_KEY_X64) {
        long long key = ((long long)cx << 32) | (long long)(unsigned)cy;
        if (cy < 0) key = ((long long)cx << 32) | (long long)cy;   // sign-extended OR
        if (key < 0) return -1;
        long long x = key >> 32, y = key & 0xffffffffLL;
        if (x >= g.dims[0] || y >= g.dims[1]) return -1;
        return (int)(x * g.dims[1] + y);
    }
    if (cy < 0 || cy >= g.dims[1]) return -1;
    return cy;
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// SPH kernels (sph.py:227-262), float32, integer powers as x*(x*x).
__device__ __forceinline__ float sph_w(int kind, int dim, float r, float h) {
    const float hh = h < 1e-12f ? 1e-12f : h;
    const float PI = 3.14159265358979323846f;
    if (kind == JT_SPH_CUBIC) {
        float q = r / hh;
        float sigma = (dim == 2) ? 10.0f / (21.991148575128552f * (hh * hh)) : 1.0f / (PI * (hh * (hh * hh)));
        float q2 = q * q, q3 = q * q2;
        float w1 = sigma * ((1.0f - 1.5f * q2) + 0.75f * q3);
        float t = 2.0f - q;
        float w2 = (sigma * 0.25f) * (t * (t * t));
        return q <= 1.0f ? w1 : (q <= 2.0f ? w2 : 0.0f);
    } else if (kind == JT_SPH_WENDLAND) {
        float q = r / hh;
        float sigma = (dim == 2) ? 7.0f / (PI * (hh * hh)) : 21.0f / (6.283185307179586f * (hh * (hh * hh)));
        float t = 1.0f - q;
        t = t < 0.0f ? 0.0f : t;
        float t2 = t * t;
        return (sigma * (t2 * t2)) * (1.0f + 4.0f * q);
    } else {
        float u = r / hh;
        float q2 = u * u;
        float sigma = (dim == 2) ? 1.0f / (PI * (hh * hh)) : 1.0f / powf(PI * (hh * hh), 1.5f);
        return sigma * expf(-q2);
    }
}

// One wavefront per query.  mode 0: neighbour lists (idx, dist); mode 1: SPH.
// LDS per wave: nc_pad 64-bit keys (d_sel bits << 32 | candidate) + nc_pad
// candidate point indices, nc_pad = next pow2 of probes*cap.
// SPH densities whose valid neighbours exceed K are marked in out_rho with
// this NaN payload by k_sph_direct (arithmetic never produces it) and then
// recomputed by k_query<1> with the top-K selection.
#define JT_SPH_REDO 0x7fbadbadu

template <int MODE>
__global__ void __launch_bounds__(256)
k_query(HGDev g, const float* __restrict__ Q, int m, float r, int cap, int K, int nc_pad,
        int* __restrict__ out_idx, float* __restrict__ out_dist,
        const float* __restrict__ masses, const float* __restrict__ hs, float hmean, int kern,
        float* __restrict__ out_rho, int redo_only) {
    extern __shared__ unsigned long long lds[];
    const int lane = threadIdx.x & 63;
    const int wv = threadIdx.x >> 6;
    const int waves = blockDim.x >> 6;
    unsigned long long* keys = lds + (size_t)wv * nc_pad;
    int* cand = (int*)(lds + (size_t)waves * nc_pad) + (size_t)wv * nc_pad;
    const int qi = blockIdx.x * waves + wv;
    if (qi >= m) return;                       // whole wave exits together
    if (MODE == 1 && redo_only && __float_as_uint(out_rho[qi]) != JT_SPH_REDO) return;
    const int D = g.dim;
    float q[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < D; ++a) q[a] = Q[(size_t)qi * D + a];
    int cq[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) cq[a] = cell_coord(q[a], g.origin[a], g.inv);
    const int probes = (D == 3) ? 27 : 9;
    const int nc = probes * cap;
    const int first = g.order[0];
    int nvalid = 0;
    float acc = 0.0f;
    for (int p = 0; p < probes; ++p) {
        int di = (D == 3) ? p / 9 - 1 : p / 3 - 1;
        int dj = (D == 3) ? (p / 3) % 3 - 1 : p % 3 - 1;
        int dk = (D == 3) ? p % 3 - 1 : 0;
        int cell = (D == 3) ? probe_cell(g, cq[0] + di, cq[1] + dj, cq[2] + dk)
                            : probe_cell(g, cq[0] + di, cq[1] + dj, 0);
        int st = 0, cnt = 0;
        if (cell >= 0) { st = g.cstart[cell]; cnt = g.cend[cell] - st; }
        for (int s0 = 0; s0 < cap; s0 += 64) {
            int s = s0 + lane;
            if (s >= cap) continue;
            int c = p * cap + s;
            bool vs = s < cnt;
            int pt = vs ? g.order[st + s] : first;
            float d2;
            if (vs) {
                const float* pp = g.spos + (size_t)(st + s) * D;
                float dx = pp[0] - q[0], dy = pp[1] - q[1];
                d2 = dx * dx + dy * dy;
                if (D == 3) { float dz = pp[2] - q[2]; d2 = d2 + dz * dz; }
            } else {
                const float* pp = g.spos;       // sorted slot 0 holds order[0]
                float dx = pp[0] - q[0], dy = pp[1] - q[1];
                d2 = dx * dx + dy * dy;
                if (D == 3) { float dz = pp[2] - q[2]; d2 = d2 + dz * dz; }
            }
            float d = sqrtf(d2);
            bool valid = vs && (d <= r);
            float dsel = valid ? d : __int_as_float(0x7f800000);
            nvalid += valid ? 1 : 0;
            if (MODE == 1 && valid) {
                float mass = masses ? masses[pt] : 1.0f;
                float h = hs ? hs[pt] : hmean;
                acc += mass * sph_w(kern, D, d, h);
            }
            keys[c] = ((unsigned long long)__float_as_uint(dsel) << 32) | (unsigned)c;
            cand[c] = pt;
        }
    }
    for (int c = nc + lane; c < nc_pad; c += 64) { keys[c] = ~0ull; cand[c] = first; }
    nvalid = wave_sum_i(nvalid);
    // each wave owns its LDS slice: a wave-level fence orders its writes
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();

    if (MODE == 1 && nvalid <= K) {
        acc = wave_sum(acc);
        if (lane == 0) out_rho[qi] = acc;
        return;
    }
    if (MODE == 1 && masses == nullptr && hs == nullptr) {
        // The K smallest keys without sorting: a radix select, 8 bits a pass,
        // over the distance bits (non-negative floats order as their bits)
        // finds the K-th smallest distance Td and the rank kr of that key
        // among the keys of distance Td.  The K nearest are then every
        // candidate nearer than Td and kr + 1 candidates at Td, and with
        // uniform masses and smoothing the sum needs no more (nvalid > K, so
        // all of them are within r).  Each wave has 256 counters after the
        // candidate slices.
        unsigned* hist = (unsigned*)((int*)(lds + (size_t)waves * nc_pad) + (size_t)waves * nc_pad) + wv * 256;
        unsigned prefix = 0u, pmask = 0u;
        int kr = K - 1;                                 // rank sought among the keys under the prefix
        for (int shift = 24; shift >= 0; shift -= 8) {
            for (int b = lane; b < 256; b += 64) hist[b] = 0u;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            for (int c = lane; c < nc; c += 64) {
                const unsigned db = (unsigned)(keys[c] >> 32);
                if ((db & pmask) == prefix) atomicAdd(&hist[(db >> shift) & 255u], 1u);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
            const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                           h3 = hist[4 * lane + 3];
            const unsigned tot = h0 + h1 + h2 + h3;
            unsigned incl = tot;
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned v = __shfl_up(incl, o);
                if (lane >= o) incl += v;
            }
            const unsigned excl = incl - tot;
            const bool mine = (unsigned)kr >= excl && (unsigned)kr < incl;   // exactly one lane
            const int src = __builtin_ctzll(__ballot(mine));
            int bin = 4 * lane;
            unsigned below = excl, r2 = (unsigned)kr - excl;
            if (r2 >= h0) {
                r2 -= h0; below += h0; ++bin;
                if (r2 >= h1) {
                    r2 -= h1; below += h1; ++bin;
                    if (r2 >= h2) { below += h2; ++bin; }
                }
            }
            bin = __shfl(bin, src);
            below = __shfl(below, src);
            kr -= (int)below;
            prefix |= (unsigned)bin << shift;
            pmask |= 255u << shift;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // the counters are read before
            __builtin_amdgcn_wave_barrier();                           // the next pass clears them
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        float s = 0.0f;
        for (int c = lane; c < nc; c += 64) {
            const unsigned db = (unsigned)(keys[c] >> 32);
            if (db < prefix) s += sph_w(kern, D, __uint_as_float(db), hmean);
        }
        s = wave_sum(s);
        if (lane == 0) out_rho[qi] = s + (float)(kr + 1) * sph_w(kern, D, __uint_as_float(prefix), hmean);
        return;
    }
    // Bitonic sort of this wave's nc_pad keys in LDS (ascending).
    for (int k = 2; k <= nc_pad; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < nc_pad; i += 64) {
                int ixj = i ^ j;
                if (ixj > i) {
                    unsigned long long a = keys[i], b = keys[ixj];
                    bool up = (i & k) == 0;
                    if ((a > b) == up) { keys[i] = b; keys[ixj] = a; }
                }
            }
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        }
    }
    if (MODE == 0) {
        for (int k2 = lane; k2 < K; k2 += 64) {
            int id = first;
            float dd = __int_as_float(0x7f800000);
            if (k2 < nc) {
                unsigned long long key = keys[k2];
                id = cand[(int)(key & 0xffffffffu)];
                dd = __uint_as_float((unsigned)(key >> 32));
            } else {
                id = -1;
            }
            out_idx[(size_t)qi * K + k2] = id;
            out_dist[(size_t)qi * K + k2] = dd;
        }
    } else {
        float s = 0.0f;
        for (int k2 = lane; k2 < K && k2 < nc; k2 += 64) {
            unsigned long long key = keys[k2];
            float dd = __uint_as_float((unsigned)(key >> 32));
            if (dd <= r) {
                int pt = cand[(int)(key & 0xffffffffu)];
                float mass = masses ? masses[pt] : 1.0f;
                float h = hs ? hs[pt] : hmean;
                s += mass * sph_w(kern, D, dd, h);
            }
        }
        s = wave_sum(s);
        if (lane == 0) out_rho[qi] = s;
    }
}

static int make_hg(const jt_hashgrid* h, HGDev& g) {
    JT_REQUIRE(h != nullptr, "hashgrid descriptor is NULL");
    JT_REQUIRE(h->dim == 2 || h->dim == 3, "dim must be 2 or 3");
    JT_REQUIRE(h->key_mode == JT_KEY_X32 || h->key_mode == JT_KEY_X64, "bad key_mode");
    JT_REQUIRE(h->n_points >= 1 && h->n_points < (1LL << 31), "n_points out of range");
    long long nc = 1;
    for (int a = 0; a < h->dim; ++a) {
        JT_REQUIRE(h->dims[a] >= 1 && h->dims[a] < (1 << 21), "dims out of range");
        if (h->key_mode == JT_KEY_X64 || a > 0) nc *= h->dims[a];
    }
    JT_REQUIRE(nc < (1LL << 31), "too many cells (%lld)", nc);
    g.dim = h->dim; g.key_mode = h->key_mode;
    for (int a = 0; a < 3; ++a) { g.origin[a] = h->origin[a]; g.dims[a] = h->dims[a]; }
    g.inv = h->inv;
    g.n = (int)h->n_points;
    g.order = h->order; g.spos = h->sorted_pos; g.cstart = h->cell_start; g.cend = h->cell_end;
    return JT_OK;
}

}  // namespace jt

using namespace jt;

extern "C" int64_t jt_hashgrid_num_cells(const jt_hashgrid* h) {
    long long nc = 1;
    for (int a = 0; a < h->dim; ++a)
        if (h->key_mode == JT_KEY_X64 || a > 0) nc *= h->dims[a];
    return nc;
}

extern "C" int jt_points_bounds(const float* P, int64_t n, int32_t dim, float* out_minmax,
                                void* ws, size_t ws_bytes, void* stream) {
    JT_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
    JT_REQUIRE(n >= 1 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(P && out_minmax, "NULL pointer");
    JT_REQUIRE(ws && ws_bytes >= 6 * sizeof(unsigned), "workspace too small");
    hipStream_t s = (hipStream_t)stream;
    unsigned* part = (unsigned*)ws;                     // 6 words per block
    long long blocks = (n + 1023) / 1024;
    blocks = blocks < 1024 ? blocks : 1024;
    const long long fit = (long long)(ws_bytes / (6 * sizeof(unsigned)));
    blocks = blocks < fit ? blocks : fit;
    const int vec = ((uintptr_t)P & 15) == 0;
    hipLaunchKernelGGL(k_bounds, dim3((unsigned)blocks), dim3(256), 0, s, P, (int)n, dim, vec, part);
    hipLaunchKernelGGL(k_decode_bounds, dim3(1), dim3(256), 0, s, (const unsigned*)part, (int)blocks, out_minmax, dim);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

static int hg_scan_blocks(long long E) { return (int)((E + HG_SCAN_SPAN - 1) / HG_SCAN_SPAN); }

extern "C" size_t jt_hashgrid_workspace_size(int64_t n) {
    const size_t nn = n > 0 ? (size_t)n : 1;
    const size_t ntiles = (nn + HG_TILE - 1) / HG_TILE;
    Arena a(nullptr, 0);
    a.take<float4>(nn); a.take<float4>(nn);             // records between passes (two for 3-4 passes)
    a.take<unsigned>(nn);                               // sorted keys
    a.take<unsigned>(256 * ntiles);                     // digit counts, then offsets
    a.take<unsigned>(hg_scan_blocks(256LL * ntiles));   // scan block sums
    return a.used + 256;
}

template <int DIM>
static void hg_pass(hipStream_t s, const HGDev& g, const float* P, const float4* rin, float4* rout, bool first,
                    bool last, int shift, int db, int ntiles, unsigned nc, unsigned* hist, unsigned* bsum,
                    int* order, float* spos, unsigned* skeys) {
    const dim3 b(HG_BLOCK), gt((unsigned)ntiles);
    if (first) hipLaunchKernelGGL((k_hg_count<DIM, true>), gt, b, 0, s, g, P, rin, shift, db, ntiles, nc, hist);
    else hipLaunchKernelGGL((k_hg_count<DIM, false>), gt, b, 0, s, g, P, rin, shift, db, ntiles, nc, hist);
    const int E = (1 << db) * ntiles;
    const dim3 gs((unsigned)hg_scan_blocks(E));
    hipLaunchKernelGGL(k_hg_scan_sums, gs, b, 0, s, (const unsigned*)hist, E, bsum);
    hipLaunchKernelGGL(k_hg_scan_apply, gs, b, 0, s, hist, E, (const unsigned*)bsum);
#define JT_HG_SCATTER(F, L) \
    hipLaunchKernelGGL((k_hg_scatter<DIM, F, L>), gt, b, 0, s, g, P, rin, shift, db, ntiles, nc, \
                       (const unsigned*)hist, rout, order, spos, skeys)
    if (first && last) JT_HG_SCATTER(true, true);
    else if (first) JT_HG_SCATTER(true, false);
    else if (last) JT_HG_SCATTER(false, true);
    else JT_HG_SCATTER(false, false);
#undef JT_HG_SCATTER
}

extern "C" int jt_hashgrid_build(const jt_hashgrid* h, const float* P, void* ws, size_t ws_bytes, void* stream) {
    HGDev g;
    int rc = make_hg(h, g);
    if (rc) return rc;
    JT_REQUIRE(P && h->order && h->sorted_pos && h->cell_start && h->cell_end, "NULL pointer");
    JT_REQUIRE(h->n_points < (1LL << 31) - HG_TILE, "n_points out of range");
    hipStream_t s = (hipStream_t)stream;
    const int n = g.n;
    const long long nc = jt_hashgrid_num_cells(h);
    const int ntiles = (n + HG_TILE - 1) / HG_TILE;
    Arena a(ws, ws_bytes);
    float4* ra = a.take<float4>(n);
    float4* rb = a.take<float4>(n);
    unsigned* skeys = a.take<unsigned>(n);
    unsigned* hist = a.take<unsigned>((size_t)256 * ntiles);
    unsigned* bsum = a.take<unsigned>(hg_scan_blocks(256LL * ntiles));
    if (!ra || !rb || !skeys || !hist || !bsum || !a.ok())
        return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", jt_hashgrid_workspace_size(n), ws_bytes);
    int bits = 1;                                      // key bits: the cell ids are < nc <= 2^bits
    while (bits < 31 && (1LL << bits) < nc) ++bits;
    const int npass = (bits + 7) / 8;
    const int dbits = (bits + npass - 1) / npass;
    const float4* rin = nullptr;
    for (int p = 0; p < npass; ++p) {
        const int shift = p * dbits;
        const int db = dbits < bits - shift ? dbits : bits - shift;
        float4* rout = (p & 1) ? rb : ra;
        if (g.dim == 3)
            hg_pass<3>(s, g, P, rin, rout, p == 0, p == npass - 1, shift, db, ntiles, (unsigned)nc, hist, bsum,
                       h->order, h->sorted_pos, skeys);
        else
            hg_pass<2>(s, g, P, rin, rout, p == 0, p == npass - 1, shift, db, ntiles, (unsigned)nc, hist, bsum,
                       h->order, h->sorted_pos, skeys);
        rin = rout;
    }
    const dim3 b(256), gr((n + 255) / 256);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nc + 255) / 256)), b, 0, s, (unsigned*)h->cell_start, (int)nc, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nc + 255) / 256)), b, 0, s, (unsigned*)h->cell_end, (int)nc, 0u);
    hipLaunchKernelGGL(k_cell_table, gr, b, 0, s, (const unsigned*)skeys, n, h->cell_start, h->cell_end);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

// SPH density with no neighbour list: one wavefront per query sums m W(d, h)
// over its valid candidates.  The 27 (2-D: 9) probe cells' capped slot
// ranges are laid end to end (a wave scan of their counts, kept in LDS) and
// the lanes take consecutive candidates of that list, so cells of a few
// points do not leave most lanes idle.  Where the valid candidates number at
// most K (every one of them is among the K nearest) that sum is the density;
// otherwise it writes JT_SPH_REDO for the top-K pass.
// x32 keys drop the cell's x index (probe_cell), so the three x offsets of
// the reference's stencil probe the same cells: the candidate list is the
// 9 (2-D: 3) distinct cells' list three times over.  Those cells are read
// once and every candidate counts three times (3 nvalid against K, 3 times
// the sum: the same multiset of terms the reference adds).
__global__ void __launch_bounds__(256)
k_sph_direct(HGDev g, const float* __restrict__ Q, int m, float r, int cap, int K,
             const float* __restrict__ masses, const float* __restrict__ hs, float hmean, int kern,
             float* __restrict__ out_rho) {
    __shared__ int toff[4][32], tst[4][32];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int qi = blockIdx.x * 4 + wv;
    if (qi >= m) return;                       // whole wave exits together (no block barrier below)
    const int D = g.dim;
    float q[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < D; ++a) q[a] = Q[(size_t)qi * D + a];
    int cq[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) cq[a] = cell_coord(q[a], g.origin[a], g.inv);
    const bool x32 = g.key_mode == JT_KEY_X32;
    const int mult = x32 ? 3 : 1;              // copies of each distinct probe cell in the stencil
    const int probes = ((D == 3) ? 27 : 9) / mult;
    int st = 0, cnt = 0;
    if (lane < probes) {                       // lane p: probe p's cell, in the reference's offset order
        const int p = x32 ? lane + probes : lane;   // x32: the di = 0 third of the stencil
        const int di = (D == 3) ? p / 9 - 1 : p / 3 - 1;
        const int dj = (D == 3) ? (p / 3) % 3 - 1 : p % 3 - 1;
        const int dk = (D == 3) ? p % 3 - 1 : 0;
        const int cell = (D == 3) ? probe_cell(g, cq[0] + di, cq[1] + dj, cq[2] + dk)
                                  : probe_cell(g, cq[0] + di, cq[1] + dj, 0);
        if (cell >= 0) { st = g.cstart[cell]; cnt = min(g.cend[cell] - st, cap); }
    }
    int incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    if (lane < probes) { toff[wv][lane] = incl - cnt; tst[wv][lane] = st; }
    if (lane == 0) toff[wv][probes] = total;   // sentinel: the list's end
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    const bool per_point = masses != nullptr || hs != nullptr;
    int nvalid = 0;
    float acc = 0.0f;
    int p = 0;
    for (int f = lane; f < total; f += 64) {
        while (f >= toff[wv][p + 1]) ++p;      // stops at the sentinel at the latest
        const int slot = tst[wv][p] + (f - toff[wv][p]);
        const float* pp = g.spos + (size_t)slot * D;
        float dx = pp[0] - q[0], dy = pp[1] - q[1];
        float d2 = dx * dx + dy * dy;
        if (D == 3) { float dz = pp[2] - q[2]; d2 = d2 + dz * dz; }
        const float d = sqrtf(d2);
        if (d <= r) {
            ++nvalid;
            float mass = 1.0f, h = hmean;
            if (per_point) {
                const int pt = g.order[slot];
                if (masses) mass = masses[pt];
                if (hs) h = hs[pt];
            }
            acc += mass * sph_w(kern, D, d, h);
        }
    }
    nvalid = wave_sum_i(nvalid) * mult;
    acc = wave_sum(acc);
    if (x32) acc = acc * 3.0f;
    if (lane == 0) out_rho[qi] = nvalid <= K ? acc : __uint_as_float(JT_SPH_REDO);
}

// The top-K pass of jt_sph_eval for uniform masses and smoothing: one wave
// per query that k_sph_direct marked JT_SPH_REDO (more than K valid
// candidates), others exit at once.  The distinct probe cells' capped slots
// (x32: 9 cells, each standing for mult = 3 stencil copies) are laid end to
// end as in k_sph_direct and their distance bits (invalid: +inf) kept in
// LDS.  A radix select, 8 bits a pass, finds the distinct-list key of rank
// (K - 1) / mult: its distance Td, and b distinct candidates strictly nearer.
// The K nearest of the reference's multiset are then the mult copies of
// those b and K - mult b copies at Td (all within r, as more than K are), so
// the density is mult * sum_{d < Td} W(d) + (K - mult b) W(Td).
// LDS per wave: list_cap distance words, 256 counters and the 32-entry
// probe tables.
__device__ __forceinline__ void sph_select_query(const HGDev& g, const float* __restrict__ Q, int qi, float r,
                                                 int cap, int K, float hmean, int kern, unsigned* dist,
                                                 unsigned* hist, int* toff, int* tst, float* __restrict__ out_rho) {
    const int lane = threadIdx.x & 63;
    const int D = g.dim;
    float q[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < D; ++a) q[a] = Q[(size_t)qi * D + a];
    int cq[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) cq[a] = cell_coord(q[a], g.origin[a], g.inv);
    const bool x32 = g.key_mode == JT_KEY_X32;
    const int mult = x32 ? 3 : 1;
    const int probes = ((D == 3) ? 27 : 9) / mult;
    int st = 0, cnt = 0;
    if (lane < probes) {
        const int p = x32 ? lane + probes : lane;
        const int di = (D == 3) ? p / 9 - 1 : p / 3 - 1;
        const int dj = (D == 3) ? (p / 3) % 3 - 1 : p % 3 - 1;
        const int dk = (D == 3) ? p % 3 - 1 : 0;
        const int cell = (D == 3) ? probe_cell(g, cq[0] + di, cq[1] + dj, cq[2] + dk)
                                  : probe_cell(g, cq[0] + di, cq[1] + dj, 0);
        if (cell >= 0) { st = g.cstart[cell]; cnt = min(g.cend[cell] - st, cap); }
    }
    int incl = cnt;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o);
        if (lane >= o) incl += v;
    }
    const int total = __shfl(incl, 63);
    if (lane < probes) { toff[lane] = incl - cnt; tst[lane] = st; }
    if (lane == 0) toff[probes] = total;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    int p = 0;
    for (int f = lane; f < total; f += 64) {
        while (f >= toff[p + 1]) ++p;
        const int slot = tst[p] + (f - toff[p]);
        const float* pp = g.spos + (size_t)slot * D;
        float dx = pp[0] - q[0], dy = pp[1] - q[1];
        float d2 = dx * dx + dy * dy;
        if (D == 3) { float dz = pp[2] - q[2]; d2 = d2 + dz * dz; }
        const float d = sqrtf(d2);
        dist[f] = d <= r ? __float_as_uint(d) : 0x7f800000u;
    }
    unsigned prefix = 0u, pmask = 0u;
    const int kr0 = (K - 1) / mult;
    int kr = kr0;                                   // rank sought among the keys under the prefix
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int b = lane; b < 256; b += 64) hist[b] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        for (int c = lane; c < total; c += 64) {
            const unsigned db = dist[c];
            if ((db & pmask) == prefix) atomicAdd(&hist[(db >> shift) & 255u], 1u);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        const unsigned h0 = hist[4 * lane], h1 = hist[4 * lane + 1], h2 = hist[4 * lane + 2],
                       h3 = hist[4 * lane + 3];
        const unsigned tot = h0 + h1 + h2 + h3;
        unsigned inc = tot;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned v = __shfl_up(inc, o);
            if (lane >= o) inc += v;
        }
        const unsigned excl = inc - tot;
        const bool mine = (unsigned)kr >= excl && (unsigned)kr < inc;   // exactly one lane
        const int src = __builtin_ctzll(__ballot(mine));
        int bin = 4 * lane;
        unsigned below = excl, r2 = (unsigned)kr - excl;
        if (r2 >= h0) {
            r2 -= h0; below += h0; ++bin;
            if (r2 >= h1) {
                r2 -= h1; below += h1; ++bin;
                if (r2 >= h2) { below += h2; ++bin; }
            }
        }
        bin = __shfl(bin, src);
        below = __shfl(below, src);
        kr -= (int)below;
        prefix |= (unsigned)bin << shift;
        pmask |= 255u << shift;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // the counters are read before
        __builtin_amdgcn_wave_barrier();                           // the next pass clears them
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    float s = 0.0f;
    for (int c = lane; c < total; c += 64) {
        const unsigned db = dist[c];
        if (db < prefix) s += sph_w(kern, D, __uint_as_float(db), hmean);
    }
    s = wave_sum(s);
    const int nearer = kr0 - kr;                    // distinct candidates strictly nearer than Td
    if (lane == 0) {
        const float wt = sph_w(kern, D, __uint_as_float(prefix), hmean);
        out_rho[qi] = (x32 ? s * 3.0f : s) + (float)(K - mult * nearer) * wt;
    }
}

__global__ void __launch_bounds__(256)
k_sph_select(HGDev g, const float* __restrict__ Q, int m, float r, int cap, int K, int list_cap, float hmean,
             int kern, float* __restrict__ out_rho) {
    extern __shared__ unsigned sel_lds[];
    const int wv = threadIdx.x >> 6;
    const int qi = blockIdx.x * 4 + wv;
    if (qi >= m) return;                       // whole wave exits together (no block barrier below)
    if (__float_as_uint(out_rho[qi]) != JT_SPH_REDO) return;
    unsigned* dist = sel_lds + (size_t)wv * (list_cap + 256 + 64);
    unsigned* hist = dist + list_cap;
    int* toff = (int*)(hist + 256);
    sph_select_query(g, Q, qi, r, cap, K, hmean, kern, dist, hist, toff, toff + 32, out_rho);
}

// The same top-K pass over a list of the queries to redo (list[0 .. *count)),
// with a fixed grid: each wave takes list entries a grid-wave stride apart,
// so the launch costs nothing where no query needs it.
__global__ void __launch_bounds__(256)
k_sph_select_list(HGDev g, const float* __restrict__ Q, const int* __restrict__ list,
                  const unsigned* __restrict__ count, float r, int cap, int K, int list_cap, float hmean, int kern,
                  float* __restrict__ out_rho) {
    extern __shared__ unsigned sel_lds[];
    const int wv = threadIdx.x >> 6;
    unsigned* dist = sel_lds + (size_t)wv * (list_cap + 256 + 64);
    unsigned* hist = dist + list_cap;
    int* toff = (int*)(hist + 256);
    const int n = (int)*count;
    for (int e = blockIdx.x * 4 + wv; e < n; e += gridDim.x * 4)
        sph_select_query(g, Q, list[e], r, cap, K, hmean, kern, dist, hist, toff, toff + 32, out_rho);
}

// ---- SPH over query tiles (x32 keys, uniform masses and smoothing) ----------
// x32 keys drop the cell's x index, so a query's candidates depend only on
// its cell's (y, z) (2-D: y): every query of one such column key has the
// same 9 (2-D: 3) probe cells and the same capped slot lists.  The queries
// are ordered by that key (k_sph_tile_keys, then a counting sort), and a
// wave takes 64 consecutive ones: it stages the candidates of one key in
// LDS (up to 9 x per_cell_cap positions) and each lane sums its own query
// over them, reading the same LDS word as every other lane (a broadcast).
// A wave whose queries carry several keys repeats that for each key.  The
// candidate multiset, the distance arithmetic and the validity test are
// k_sph_direct's, so the terms summed are the same; only their order
// differs.

// The ordering key of a query: its column's (y, z) cell (2-D: y), then
// (with_x) its x cell, each clamped to two cells beyond the grid and packed
// densely.  Ordering by x within a column keeps a wave's 64 queries close in
// x, which lets k_sph_tiled stage only the candidates near them in x.
__global__ void __launch_bounds__(256)
k_sph_tile_keys(HGDev g, const float* __restrict__ Q, int m, int with_x, unsigned* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const int D = g.dim;
    int cy = cell_coord(Q[(size_t)i * D + 1], g.origin[1], g.inv);
    cy = min(max(cy, -2), g.dims[1] + 1) + 2;
    unsigned k = (unsigned)cy;
    if (D == 3) {
        int cz = cell_coord(Q[(size_t)i * D + 2], g.origin[2], g.inv);
        cz = min(max(cz, -2), g.dims[2] + 1) + 2;
        k = (unsigned)cy * (unsigned)(g.dims[2] + 4) + (unsigned)cz;
    }
    if (with_x) {
        int cx = cell_coord(Q[(size_t)i * D], g.origin[0], g.inv);
        cx = min(max(cx, -2), g.dims[0] + 1) + 2;
        k = k * (unsigned)(g.dims[0] + 4) + (unsigned)cx;
    }
    keys[i] = k;
}

#define SPH_TILE_CAP 576                        // 9 probe cells x per_cell_cap 64

// D, KERN: the dimension and the SPH kernel as constants.  d2max: the
// largest float whose correctly rounded square root is <= r, so d2 <= d2max
// is exactly k_sph_direct's test sqrtf(d2) <= r (the square root is monotone)
// and the root is taken only for candidates within r.
template <int D, int KERN>
__global__ void __launch_bounds__(256)
k_sph_tiled(HGDev g, const float* __restrict__ Q, const int* __restrict__ order, int m, float d2max, float xr,
            int cap, int K, float hmean, float* __restrict__ out_rho, int* __restrict__ redo,
            unsigned* __restrict__ redo_count) {
    __shared__ float cand[4][SPH_TILE_CAP * 3];
    __shared__ int toff[4][16], tst[4][16];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int base = (blockIdx.x * 4 + wv) * 64;
    if (base >= m) return;                     // whole wave exits together (no block barrier below)
    const int i = base + lane;
    const bool live = i < m;
    const int qi = live ? order[i] : 0;
    float q[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < D; ++a) q[a] = Q[(size_t)qi * D + a];
    const int cy = cell_coord(q[1], g.origin[1], g.inv);
    const int cz = D == 3 ? cell_coord(q[2], g.origin[2], g.inv) : 0;
    const int probes = D == 3 ? 9 : 3;
    unsigned long long pending = __ballot(live);
    int nvalid = 0;
    float acc = 0.0f;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int ky = __shfl(cy, leader), kz = __shfl(cz, leader);
        const bool mine = ((pending >> lane) & 1ull) && cy == ky && cz == kz;
        pending &= ~__ballot(mine);
        // the key's probe cells in the reference's order (the di = 0 third
        // of the stencil; x is dropped by the key)
        int st = 0, cnt = 0;
        if (lane < probes) {
            const int dj = D == 3 ? lane / 3 - 1 : lane - 1;
            const int dk = D == 3 ? lane % 3 - 1 : 0;
            const int cell = D == 3 ? probe_cell(g, 0, ky + dj, kz + dk) : probe_cell(g, 0, ky + dj, 0);
            if (cell >= 0) { st = g.cstart[cell]; cnt = min(g.cend[cell] - st, cap); }
        }
        int incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        const int total = __shfl(incl, 63);
        if (lane < probes) { toff[wv][lane] = incl - cnt; tst[wv][lane] = st; }
        if (lane == 0) toff[wv][probes] = total;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        // stage only the candidates within xr in x of this key's queries'
        // x range: for any other one every query's float dx (the kernel's
        // own subtraction, monotone in q) exceeds xr > r, so its d2 exceeds
        // d2max and dropping it changes no term
        float qlo = mine ? q[0] : INFINITY, qhi = mine ? q[0] : -INFINITY;
        for (int o = 32; o > 0; o >>= 1) {
            qlo = fminf(qlo, __shfl_xor(qlo, o));
            qhi = fmaxf(qhi, __shfl_xor(qhi, o));
        }
        int pr = 0, kept = 0;
        for (int f0 = 0; f0 < total; f0 += 64) {
            const int f = f0 + lane;
            float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f;
            bool keep = false;
            if (f < total) {
                while (f >= toff[wv][pr + 1]) ++pr;
                const float* pp = g.spos + (size_t)(tst[wv][pr] + (f - toff[wv][pr])) * D;
                c0 = pp[0]; c1 = pp[1];
                if constexpr (D == 3) c2 = pp[2];
                keep = c0 - qhi <= xr && qlo - c0 <= xr;
            }
            const unsigned long long km = __ballot(keep);
            if (keep) {
                const int at = kept + __popcll(km & ((1ull << lane) - 1ull));
                cand[wv][at * 3] = c0;
                cand[wv][at * 3 + 1] = c1;
                cand[wv][at * 3 + 2] = c2;
            }
            kept += __popcll(km);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (mine) {
            for (int f = 0; f < kept; ++f) {
                const float dx = cand[wv][f * 3] - q[0], dy = cand[wv][f * 3 + 1] - q[1];
                float d2 = dx * dx + dy * dy;
                if constexpr (D == 3) { const float dz = cand[wv][f * 3 + 2] - q[2]; d2 = d2 + dz * dz; }
                if (d2 <= d2max) {
                    ++nvalid;
                    acc += sph_w(KERN, D, sqrtf(d2), hmean);
                }
            }
        }
        // every lane is done with this key's candidates before the next key's
        // overwrite them
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    // more than K valid candidates: the top-K pass takes the query (listed,
    // one atomic per wave)
    const bool again = live && nvalid * 3 > K;
    if (live) out_rho[qi] = again ? __uint_as_float(JT_SPH_REDO) : acc * 3.0f;
    const unsigned long long am = __ballot(again);
    if (am) {
        const int first = __ffsll((long long)am) - 1;
        unsigned at = 0;
        if (lane == first) at = atomicAdd(redo_count, (unsigned)__popcll(am));
        at = __shfl(at, first);
        if (again) redo[at + __popcll(am & ((1ull << lane) - 1ull))] = qi;
    }
}

static int query_common(const jt_hashgrid* h, HGDev& g, int64_t m, int32_t cap, int32_t K, int& nc_pad, int& waves) {
    int rc = make_hg(h, g);
    if (rc) return rc;
    JT_REQUIRE(m >= 0 && m < (1LL << 31), "m out of range");
    JT_REQUIRE(cap >= 1 && cap <= 4096, "per_cell_cap out of range");
    JT_REQUIRE(K >= 1, "max_neighbors must be >= 1");
    int probes = g.dim == 3 ? 27 : 9;
    int nc = probes * cap;
    nc_pad = 64;
    while (nc_pad < nc) nc_pad <<= 1;
    size_t per_wave = (size_t)nc_pad * 12 + 256 * 4;      // keys, candidates, select counters
    waves = (int)(65536 / per_wave);
    waves = waves < 1 ? 1 : (waves > 4 ? 4 : waves);
    JT_REQUIRE(per_wave * waves <= 160 * 1024, "per_cell_cap too large for LDS");
    return JT_OK;
}

extern "C" int jt_hashgrid_query(const jt_hashgrid* h, const float* Q, int64_t m, float radius,
                                 int32_t per_cell_cap, int32_t max_neighbors, int32_t* out_idx,
                                 float* out_dist, void* stream) {
    HGDev g;
    int nc_pad, waves;
    int rc = query_common(h, g, m, per_cell_cap, max_neighbors, nc_pad, waves);
    if (rc) return rc;
    if (m == 0) return JT_OK;
    JT_REQUIRE(Q && out_idx && out_dist, "NULL pointer");
    size_t lds = (size_t)nc_pad * 12 * waves;
    hipLaunchKernelGGL(k_query<0>, dim3((unsigned)((m + waves - 1) / waves)), dim3(64 * waves), lds,
                       (hipStream_t)stream, g, Q, (int)m, radius, per_cell_cap, max_neighbors, nc_pad,
                       out_idx, out_dist, nullptr, nullptr, 0.0f, 0, nullptr, 0);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

// Per query: C_q = sum over the probed cells of min(count, cap) (candidates
// read), and the number of those within r.  The SPH/KDE byte and pair models.
// One wavefront per query, lanes striding over each cell's candidates.
__global__ void __launch_bounds__(256)
k_query_stats(HGDev g, const float* __restrict__ Q, int m, float r, int cap, int* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int qi = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (qi >= m) return;                       // whole wave exits together
    const int D = g.dim;
    float q[3] = {0.f, 0.f, 0.f};
    for (int a = 0; a < D; ++a) q[a] = Q[(size_t)qi * D + a];
    int cq[3] = {0, 0, 0};
    for (int a = 0; a < D; ++a) cq[a] = cell_coord(q[a], g.origin[a], g.inv);
    const int probes = (D == 3) ? 27 : 9;
    int cand = 0, within = 0;
    for (int p = 0; p < probes; ++p) {
        int di = (D == 3) ? p / 9 - 1 : p / 3 - 1;
        int dj = (D == 3) ? (p / 3) % 3 - 1 : p % 3 - 1;
        int dk = (D == 3) ? p % 3 - 1 : 0;
        int cell = (D == 3) ? probe_cell(g, cq[0] + di, cq[1] + dj, cq[2] + dk)
                            : probe_cell(g, cq[0] + di, cq[1] + dj, 0);
        if (cell < 0) continue;
        const int st = g.cstart[cell];
        const int cnt = min(g.cend[cell] - st, cap);
        cand += cnt;
        for (int s = lane; s < cnt; s += 64) {
            const float* pp = g.spos + (size_t)(st + s) * D;
            float dx = pp[0] - q[0], dy = pp[1] - q[1];
            float d2 = dx * dx + dy * dy;
            if (D == 3) { float dz = pp[2] - q[2]; d2 = d2 + dz * dz; }
            within += (sqrtf(d2) <= r) ? 1 : 0;
        }
    }
    for (int o = 32; o > 0; o >>= 1) within += __shfl_xor(within, o);
    if (lane == 0) {
        out[2 * (size_t)qi] = cand;
        out[2 * (size_t)qi + 1] = within;
    }
}

extern "C" int jt_hashgrid_query_stats(const jt_hashgrid* h, const float* Q, int64_t m, float radius,
                                       int32_t per_cell_cap, int32_t* stats, void* stream) {
    HGDev g;
    int rc = make_hg(h, g);
    if (rc) return rc;
    JT_REQUIRE(m >= 0 && m < (1LL << 31), "m out of range");
    JT_REQUIRE(per_cell_cap >= 1, "per_cell_cap must be >= 1");
    if (m == 0) return JT_OK;
    JT_REQUIRE(Q && stats, "NULL pointer");
    hipLaunchKernelGGL(k_query_stats, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, (hipStream_t)stream, g, Q,
                       (int)m, radius, per_cell_cap, stats);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_sph_eval(const jt_hashgrid* h, const float* Q, int64_t m, float radius, int32_t per_cell_cap,
                           int32_t max_neighbors, const float* masses, const float* smoothing, float h_mean,
                           int32_t kernel, float* out_rho, void* stream) {
    HGDev g;
    int nc_pad, waves;
    int rc = query_common(h, g, m, per_cell_cap, max_neighbors, nc_pad, waves);
    if (rc) return rc;
    JT_REQUIRE(kernel >= JT_SPH_CUBIC && kernel <= JT_SPH_GAUSSIAN, "bad kernel %d", kernel);
    if (m == 0) return JT_OK;
    JT_REQUIRE(Q && out_rho, "NULL pointer");
    hipLaunchKernelGGL(k_sph_direct, dim3((unsigned)((m + 3) / 4)), dim3(256), 0, (hipStream_t)stream, g, Q,
                       (int)m, radius, per_cell_cap, max_neighbors, masses, smoothing, h_mean, kernel, out_rho);
    JT_HIP(hipGetLastError());
    const int mult = g.key_mode == JT_KEY_X32 ? 3 : 1;   // stencil copies of each distinct probe cell
    const int list_cap = ((g.dim == 3 ? 27 : 9) / mult) * per_cell_cap;
    const size_t sel_lds = (size_t)(list_cap + 256 + 64) * 4 * 4;
    if (masses == nullptr && smoothing == nullptr && sel_lds <= 64 * 1024) {
        hipLaunchKernelGGL(k_sph_select, dim3((unsigned)((m + 3) / 4)), dim3(256), sel_lds, (hipStream_t)stream, g,
                           Q, (int)m, radius, per_cell_cap, max_neighbors, list_cap, h_mean, kernel, out_rho);
        JT_HIP(hipGetLastError());
        return JT_OK;
    }
    size_t lds = ((size_t)nc_pad * 12 + 256 * 4) * waves;      // + k_query<1>'s radix-select counters
    hipLaunchKernelGGL(k_query<1>, dim3((unsigned)((m + waves - 1) / waves)), dim3(64 * waves), lds,
                       (hipStream_t)stream, g, Q, (int)m, radius, per_cell_cap, max_neighbors, nc_pad,
                       nullptr, nullptr, masses, smoothing, h_mean, kernel, out_rho, 1);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

// The largest float x with sqrtf(x) <= r (correctly rounded square roots
// are monotone, so for every float d2, sqrtf(d2) <= r exactly when
// d2 <= that bound); -inf when no x qualifies (r negative or NaN).
static float sqrt_le_bound(float r) {
    if (!(r >= 0.0f)) return -INFINITY;
    if (std::isinf(r)) return INFINITY;
    float x = r * r;
    while (std::isfinite(x) && std::sqrt(std::nextafter(x, INFINITY)) <= r) x = std::nextafter(x, INFINITY);
    while (x > 0.0f && std::sqrt(x) > r) x = std::nextafter(x, -INFINITY);
    return x;
}

// Buckets of the tile ordering key of a grid, with the x cell in the key
// when that keeps the bucket count at most 2^24 (with_x), else without.
static int64_t sph_tile_buckets(const HGDev& g, bool& with_x) {
    const int64_t col = g.dim == 3 ? (int64_t)(g.dims[1] + 4) * (g.dims[2] + 4) : (int64_t)(g.dims[1] + 4);
    const int64_t full = col * (g.dims[0] + 4);
    with_x = full <= (1LL << 24);
    return with_x ? full : col;
}

extern "C" size_t jt_sph_tiled_workspace_size(const jt_hashgrid* h, int64_t m) {
    HGDev g;
    if (make_hg(h, g)) return 0;
    const size_t mm = (size_t)(m > 0 ? m : 1);
    bool with_x;
    const int64_t nb64 = sph_tile_buckets(g, with_x);
    if (nb64 >= (1LL << 31)) return 0;
    const int nb = (int)nb64;
    Arena a(nullptr, 0);
    a.take<unsigned>(mm);
    a.take<int>(mm);
    a.take<int>((size_t)nb);
    a.take<int>((size_t)nb);
    a.take<char>(bucket_scan_bytes(nb));
    a.take<int>(mm);                                // queries for the top-K pass
    a.take<unsigned>(1);                            // their count
    return a.used + 256;
}

extern "C" int jt_sph_eval_tiled(const jt_hashgrid* h, const float* Q, int64_t m, float radius, int32_t per_cell_cap,
                                 int32_t max_neighbors, float h_mean, int32_t kernel, float* out_rho, void* ws,
                                 size_t ws_bytes, void* stream) {
    HGDev g;
    int nc_pad, waves;
    int rc = query_common(h, g, m, per_cell_cap, max_neighbors, nc_pad, waves);
    if (rc) return rc;
    JT_REQUIRE(kernel >= JT_SPH_CUBIC && kernel <= JT_SPH_GAUSSIAN, "bad kernel %d", kernel);
    if (g.key_mode != JT_KEY_X32)
        return fail(JT_EUNSUPPORTED, "query tiles need x32 keys (the candidates of x64 keys depend on the x cell)");
    JT_REQUIRE(per_cell_cap <= SPH_TILE_CAP / 9, "per_cell_cap above %d for query tiles", SPH_TILE_CAP / 9);
    bool with_x;
    const int64_t nb64 = sph_tile_buckets(g, with_x);
    JT_REQUIRE(nb64 < (1LL << 31), "grid too large for query tiles");
    if (m == 0) return JT_OK;
    JT_REQUIRE(Q && out_rho, "NULL pointer");
    const int nb = (int)nb64;
    Arena arena(ws, ws_bytes);
    unsigned* keys = arena.take<unsigned>((size_t)m);
    int* order = arena.take<int>((size_t)m);
    int* counts = arena.take<int>((size_t)nb);
    int* offs = arena.take<int>((size_t)nb);
    const size_t temp = bucket_scan_bytes(nb);
    char* tmp = arena.take<char>(temp);
    int* redo = arena.take<int>((size_t)m);
    unsigned* redo_count = arena.take<unsigned>(1);
    if (!keys || !order || !counts || !offs || !tmp || !redo || !redo_count || !arena.ok())
        return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", jt_sph_tiled_workspace_size(h, m), ws_bytes);
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_sph_tile_keys, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, g, Q, (int)m,
                       (int)with_x, keys);
    JT_HIP(hipGetLastError());
    JT_HIP(bucket_sort(s, keys, (int)m, 0, nb, counts, offs, tmp, temp, order));
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, redo_count, 1, 0u);
    const float d2max = sqrt_le_bound(radius);
    // x half-width of the staging window: |dx| > xr means dx * dx > r * r by
    // a margin no rounding of the sum of squares can undo
    const float xr = std::isfinite(radius) && radius >= 0.0f ? radius * 1.001f : INFINITY;
    const dim3 gt((unsigned)((m + 255) / 256)), bt(256);
#define JT_SPH_TILED(DD, KK)                                                                               \
    hipLaunchKernelGGL((k_sph_tiled<DD, KK>), gt, bt, 0, s, g, Q, (const int*)order, (int)m, d2max, xr,   \
                       per_cell_cap, max_neighbors, h_mean, out_rho, redo, redo_count)
    if (g.dim == 3) {
        if (kernel == JT_SPH_CUBIC) JT_SPH_TILED(3, JT_SPH_CUBIC);
        else if (kernel == JT_SPH_WENDLAND) JT_SPH_TILED(3, JT_SPH_WENDLAND);
        else JT_SPH_TILED(3, JT_SPH_GAUSSIAN);
    } else {
        if (kernel == JT_SPH_CUBIC) JT_SPH_TILED(2, JT_SPH_CUBIC);
        else if (kernel == JT_SPH_WENDLAND) JT_SPH_TILED(2, JT_SPH_WENDLAND);
        else JT_SPH_TILED(2, JT_SPH_GAUSSIAN);
    }
#undef JT_SPH_TILED
    JT_HIP(hipGetLastError());
    // queries with more than max_neighbors valid candidates: the top-K pass
    // over the listed ones
    const int list_cap = (g.dim == 3 ? 9 : 3) * per_cell_cap;
    const size_t sel_lds = (size_t)(list_cap + 256 + 64) * 4 * 4;
    const unsigned sel_blocks = (unsigned)std::min<int64_t>((m + 3) / 4, 2048);
    hipLaunchKernelGGL(k_sph_select_list, dim3(sel_blocks), dim3(256), sel_lds, s, g, Q, (const int*)redo,
                       (const unsigned*)redo_count, radius, per_cell_cap, max_neighbors, list_cap, h_mean, kernel,
                       out_rho);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
