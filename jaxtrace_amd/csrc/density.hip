// Hash-grid neighbour search and SPH density (jaxtrace/density/neighbors.py,
// jaxtrace/density/sph.py).
//
// Build: dense cell key per point -> stable LSD radix sort of (key, index)
// (equal keys keep increasing point index, as the reference's stable argsort)
// -> per-cell [start, end) table -> positions gathered into cell order, so the
// candidates of one cell are contiguous in memory.
//
// Query: one 64-lane wavefront per query point.  The 27 (2-D: 9) stencil
// probes are visited in the reference's lexicographic order; lane s handles
// slot s of the probed cell (per_cell_cap slots, 64 at a time), so candidate
// c = probe * cap + slot exactly as neighbors.py:145-151 flattens them.
// Distances, validity (slot < count and d <= r) and the stable "first K by
// distance" selection follow neighbors.py:153-165 bit for bit.
#include "jt_common.h"
#include <hipcub/hipcub.hpp>

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
// uint32 encodings): each thread folds a grid-stride slice, each wave folds
// its lanes with shuffles, the block folds its waves in LDS, and one thread
// per block adds the result atomically (the six words share one cache line,
// so the atomics are kept to a few per block).
__global__ void __launch_bounds__(256) k_bounds(const float* __restrict__ P, int n, int dim, unsigned* out) {
    __shared__ unsigned red[4][6];
    unsigned lmin[3] = {0xffffffffu, 0xffffffffu, 0xffffffffu}, lmax[3] = {0u, 0u, 0u};
    const int stride = gridDim.x * blockDim.x;
    if (dim == 3) {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            const float* p = P + (size_t)i * 3;
            const unsigned a = f2ord(p[0]), b = f2ord(p[1]), c = f2ord(p[2]);
            lmin[0] = min(lmin[0], a); lmax[0] = max(lmax[0], a);
            lmin[1] = min(lmin[1], b); lmax[1] = max(lmax[1], b);
            lmin[2] = min(lmin[2], c); lmax[2] = max(lmax[2], c);
        }
    } else {
        for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
            const float* p = P + (size_t)i * 2;
            const unsigned a = f2ord(p[0]), b = f2ord(p[1]);
            lmin[0] = min(lmin[0], a); lmax[0] = max(lmax[0], a);
            lmin[1] = min(lmin[1], b); lmax[1] = max(lmax[1], b);
        }
    }
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
    if (threadIdx.x == 0)
        for (int a = 0; a < dim; ++a) {
            unsigned mn = red[0][a], mx = red[0][3 + a];
            for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { mn = min(mn, red[w][a]); mx = max(mx, red[w][3 + a]); }
            atomicMin(&out[a], mn);
            atomicMax(&out[3 + a], mx);
        }
}

__global__ void k_decode_bounds(const unsigned* in, float* out, int dim) {
    int a = threadIdx.x;
    if (a < dim) { out[a] = ord2f(in[a]); out[dim + a] = ord2f(in[3 + a]); }
}

// neighbors.py:70: floor((P - origin) * inv) as int32.
__device__ __forceinline__ int cell_coord(float p, float o, float inv) {
    float f = floorf((p - o) * inv);
    if (f != f) return 0;
    f = fminf(fmaxf(f, -1073741824.0f), 1073741824.0f);   // keep the int conversion defined
    return (int)f;
}

__global__ void __launch_bounds__(256)
k_point_keys(HGDev g, const float* __restrict__ P, unsigned* __restrict__ keys, int* __restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    int c[3] = {0, 0, 0};
    for (int a = 0; a < g.dim; ++a) c[a] = cell_coord(P[(size_t)i * g.dim + a], g.origin[a], g.inv);
    unsigned k;
    if (g.dim == 3)
        k = (g.key_mode == JT_KEY_X64) ? (unsigned)(((long long)c[0] * g.dims[1] + c[1]) * g.dims[2] + c[2])
                                       : (unsigned)((long long)c[1] * g.dims[2] + c[2]);
    else
        k = (g.key_mode == JT_KEY_X64) ? (unsigned)((long long)c[0] * g.dims[1] + c[1]) : (unsigned)c[1];
    keys[i] = k;
    idx[i] = i;
}

__global__ void __launch_bounds__(256)
k_cell_table(const unsigned* __restrict__ skeys, int n, int* __restrict__ cstart, int* __restrict__ cend) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned k = skeys[i];
    if (i == 0 || skeys[i - 1] != k) cstart[k] = i;
    if (i == n - 1 || skeys[i + 1] != k) cend[k] = i + 1;
}

// spos[i] = P[order[i]]: one 12-byte (3-D) or 8-byte (2-D) load and store
// per point, so a random source record costs one memory request.
struct alignas(4) Rec3 { float v[3]; };
struct alignas(4) Rec2 { float v[2]; };
__global__ void __launch_bounds__(256)
k_gather_sorted(const float* __restrict__ P, const int* __restrict__ order, int n, int dim, float* __restrict__ spos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = order[i];
    if (dim == 3) reinterpret_cast<Rec3*>(spos)[i] = reinterpret_cast<const Rec3*>(P)[j];
    else reinterpret_cast<Rec2*>(spos)[i] = reinterpret_cast<const Rec2*>(P)[j];
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
    unsigned* acc = (unsigned*)ws;
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, acc, 3, 0xffffffffu);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, acc + 3, 3, 0u);
    int blocks = (int)((n + 255) / 256);
    blocks = blocks > 512 ? 512 : blocks;
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(256), 0, s, P, (int)n, dim, acc);
    hipLaunchKernelGGL(k_decode_bounds, dim3(1), dim3(64), 0, s, acc, out_minmax, dim);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" size_t jt_hashgrid_workspace_size(int64_t n) {
    size_t temp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, (int)n, 0, 32);
    Arena a(nullptr, 0);
    a.take<unsigned>(n); a.take<unsigned>(n); a.take<int>(n); a.take<char>(temp);
    return a.used + 256;
}

extern "C" int jt_hashgrid_build(const jt_hashgrid* h, const float* P, void* ws, size_t ws_bytes, void* stream) {
    HGDev g;
    int rc = make_hg(h, g);
    if (rc) return rc;
    JT_REQUIRE(P && h->order && h->sorted_pos && h->cell_start && h->cell_end, "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    const int n = g.n;
    const long long nc = jt_hashgrid_num_cells(h);
    size_t temp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, n, 0, 32);
    Arena a(ws, ws_bytes);
    unsigned* keys = a.take<unsigned>(n);
    unsigned* skeys = a.take<unsigned>(n);
    int* idx = a.take<int>(n);
    char* tmp = a.take<char>(temp);
    if (!keys || !skeys || !idx || !tmp || !a.ok())
        return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", jt_hashgrid_workspace_size(n), ws_bytes);
    int end_bit = 1;
    while (end_bit < 32 && (1LL << end_bit) < nc) ++end_bit;
    dim3 b(256), gr((n + 255) / 256);
    hipLaunchKernelGGL(k_point_keys, gr, b, 0, s, g, P, keys, idx);
    JT_HIP(hipGetLastError());
    JT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, skeys, idx, h->order, n, 0, end_bit, s));
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nc + 255) / 256)), b, 0, s, (unsigned*)h->cell_start, (int)nc, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3((unsigned)((nc + 255) / 256)), b, 0, s, (unsigned*)h->cell_end, (int)nc, 0u);
    hipLaunchKernelGGL(k_cell_table, gr, b, 0, s, skeys, n, h->cell_start, h->cell_end);
    hipLaunchKernelGGL(k_gather_sorted, gr, b, 0, s, P, (const int*)h->order, n, g.dim, h->sorted_pos);
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
__global__ void __launch_bounds__(256)
k_sph_select(HGDev g, const float* __restrict__ Q, int m, float r, int cap, int K, int list_cap, float hmean,
             int kern, float* __restrict__ out_rho) {
    extern __shared__ unsigned sel_lds[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int qi = blockIdx.x * 4 + wv;
    if (qi >= m) return;                       // whole wave exits together (no block barrier below)
    if (__float_as_uint(out_rho[qi]) != JT_SPH_REDO) return;
    unsigned* dist = sel_lds + (size_t)wv * (list_cap + 256 + 64);
    unsigned* hist = dist + list_cap;
    int* toff = (int*)(hist + 256);
    int* tst = toff + 32;
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
