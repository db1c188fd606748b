// Device-side temporal blend + trilinear sampling on a structured grid, and
// the boundary epilogue.  Shared by the tracking, sampling and adjoint
// kernels.  All arithmetic is float32, one IEEE operation at a time, in the
// reference's evaluation order (the library is built with -ffp-contract=off).
#pragma once
#include "jt_common.h"

namespace jt {

// Snapshot storage layouts (element type x components per node).  A 16-bit
// node is four halves (x, y, 0, z): z is the high half of the second dword
// and the low half is zero, so for bfloat16 that dword is z's float32 value
// with no conversion.
enum StoreKind { ST_F32x3 = 0, ST_F32x4 = 1, ST_BF16x4 = 2, ST_F16x4 = 3 };

// Result of the scalar half of time_series.py:155-245 for one stage time.
enum BlendMode { BL_INTERP = 0, BL_COPY = 1, BL_LEFT = 2, BL_RIGHT = 3, BL_ZERO = 4, BL_NAN = 5 };
struct Blend {
    int i0, i1;   // snapshot indices
    float a;      // weight
    int mode;     // BlendMode
};

struct GridDev {
    const void* data;
    long long snap_stride;   // elements per snapshot (nodes * ncomp)
    const float* times;
    int T;
    int nx, ny, nz;
    int nzs;                 // allocated z extent (nz, or nz+1 with a z halo)
    float o[3], d[3];        // origin, spacing
    float inv[3];            // 1/spacing when that is exact (spacing a power of two), else 0
    float vhi[3];            // origin + f32(n-1)*d: inclusive validity bound
    float cmax[3];           // f32(max(n-1, 0)): clip bound of the grid coordinate
    int bmode;               // JT_BOUNDS_*
    int interp, extrap;      // JT_TIME_*, JT_EXTRAP_*
    const float* pair;       // consecutive-pair records (jt_grid_field::pair_data) or nullptr
    long long pair_stride;   // floats per pair (nx * ny * nz * 12)
};

// searchsorted(times, t, side='right') for non-decreasing times.
__host__ __device__ __forceinline__ int search_right(const float* times, int T, float t) {
    int lo = 0, hi = T;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (times[mid] <= t) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// time_series.py:155-245, scalar part (which snapshots and which weight).
// Also run on the host (float32 operations with the same values) where the
// host needs a step's snapshot pair without reading the device.
__host__ __device__ inline Blend time_blend(const float* times, int T, float t, int interp, int extrap) {
    Blend b;
    b.i0 = 0; b.i1 = 0; b.a = 0.0f; b.mode = BL_COPY;
    if (interp == JT_TIME_NEAREST) {                  // :171-174, first minimum
        int best = 0;
        float bd = fabsf(times[0] - t);
        for (int k = 1; k < T; ++k) {
            float dk = fabsf(times[k] - t);
            if (dk < bd) { bd = dk; best = k; }
        }
        b.i0 = b.i1 = best;
        return b;
    }
    int ir = search_right(times, T, t);               // :162-164
    int il = ir - 1 < 0 ? 0 : ir - 1;
    il = il > T - 1 ? T - 1 : il;
    float tmin = times[0], tmax = times[T - 1];
    bool left = t < tmin, right = t > tmax;
    if (left || right) {
        if (extrap == JT_EXTRAP_CONSTANT) {           // :217-219
            b.i0 = b.i1 = left ? 0 : T - 1;
            return b;
        }
        if (extrap == JT_EXTRAP_ZERO) { b.mode = BL_ZERO; return b; }
        if (extrap == JT_EXTRAP_NAN) { b.mode = BL_NAN; return b; }
        if (T <= 1) { b.i0 = b.i1 = left ? 0 : T - 1; return b; }
        if (left) {                                    // :190-199
            float t0 = times[0], t1 = times[1];
            float dd = t1 - t0;
            dd = dd < 1e-12f ? 1e-12f : dd;
            b.i0 = 0; b.i1 = 1; b.a = (t - t0) / dd; b.mode = BL_LEFT;
        } else {                                       // :201-210
            float t0 = times[T - 2], t1 = times[T - 1];
            float dd = t1 - t0;
            dd = dd < 1e-12f ? 1e-12f : dd;
            b.i0 = T - 2; b.i1 = T - 1; b.a = (t - t1) / dd; b.mode = BL_RIGHT;
        }
        return b;
    }
    int irc = ir > T - 1 ? T - 1 : ir;                 // :178-188 (ir >= 0)
    float t0 = times[il], t1 = times[irc];
    float dd = t1 - t0;
    dd = dd < 1e-12f ? 1e-12f : dd;
    b.i0 = il; b.i1 = irc; b.a = clipf((t - t0) / dd, 0.0f, 1.0f); b.mode = BL_INTERP;
    return b;
}

template <int ST>
__device__ __forceinline__ void load_node(const void* base, int node, float& a, float& b, float& c) {
    if constexpr (ST == ST_F32x3) {
        const float* p = (const float*)base + (size_t)node * 3;
        a = p[0]; b = p[1]; c = p[2];
    } else if constexpr (ST == ST_F32x4) {
        float4 v = ((const float4*)base)[node];
        a = v.x; b = v.y; c = v.z;
    } else if constexpr (ST == ST_BF16x4) {
        uint2 v = ((const uint2*)base)[node];
        a = bf16_to_f32((uint16_t)(v.x & 0xffffu));
        b = bf16_to_f32((uint16_t)(v.x >> 16));
        c = bf16_hi_to_f32(v.y);
    } else {
        uint2 v = ((const uint2*)base)[node];
        a = f16_to_f32((uint16_t)(v.x & 0xffffu));
        b = f16_to_f32((uint16_t)(v.x >> 16));
        c = f16_to_f32((uint16_t)(v.y >> 16));
    }
}

template <int ST> struct ElemBytes { static constexpr int v = 4; };
template <> struct ElemBytes<ST_BF16x4> { static constexpr int v = 2; };
template <> struct ElemBytes<ST_F16x4> { static constexpr int v = 2; };

// Blended value of one node (time_series.py:186-188, :199, :210).
template <int ST>
__device__ __forceinline__ void corner(const void* b0, const void* b1, const Blend& bl,
                                       float oma, int node, float v[3]) {
    if (bl.mode == BL_ZERO) { v[0] = v[1] = v[2] = 0.0f; return; }
    if (bl.mode == BL_NAN) { v[0] = v[1] = v[2] = __int_as_float(0x7fc00000); return; }
    float a0, a1, a2;
    load_node<ST>(b0, node, a0, a1, a2);
    if (bl.mode == BL_COPY) { v[0] = a0; v[1] = a1; v[2] = a2; return; }
    float c0, c1, c2;
    load_node<ST>(b1, node, c0, c1, c2);
    const float al = bl.a;
    if (bl.mode == BL_INTERP) {
        v[0] = oma * a0 + al * c0;
        v[1] = oma * a1 + al * c1;
        v[2] = oma * a2 + al * c2;
    } else if (bl.mode == BL_LEFT) {
        v[0] = a0 + al * (c0 - a0);
        v[1] = a1 + al * (c1 - a1);
        v[2] = a2 + al * (c2 - a2);
    } else {
        v[0] = c0 + al * (c0 - a0);
        v[1] = c1 + al * (c1 - a1);
        v[2] = c2 + al * (c2 - a2);
    }
}

// The six values of a raw 16-bit node pair (node, node+1), 4 components each.
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int ST>
__device__ __forceinline__ void unpack_pair16(const u32x4 t, float a[6]) {
    if constexpr (ST == ST_BF16x4) {
        a[0] = bf16_to_f32((uint16_t)(t.x & 0xffffu)); a[1] = bf16_to_f32((uint16_t)(t.x >> 16));
        a[2] = bf16_hi_to_f32(t.y);
        a[3] = bf16_to_f32((uint16_t)(t.z & 0xffffu)); a[4] = bf16_to_f32((uint16_t)(t.z >> 16));
        a[5] = bf16_hi_to_f32(t.w);
    } else {
        a[0] = f16_to_f32((uint16_t)(t.x & 0xffffu)); a[1] = f16_to_f32((uint16_t)(t.x >> 16));
        a[2] = f16_to_f32((uint16_t)(t.y >> 16));
        a[3] = f16_to_f32((uint16_t)(t.z & 0xffffu)); a[4] = f16_to_f32((uint16_t)(t.z >> 16));
        a[5] = f16_to_f32((uint16_t)(t.w >> 16));
    }
}

// Two z-adjacent nodes (node, node+1) in one contiguous load (z-halo layout).
struct alignas(4) F6 { float v[6]; };
struct alignas(4) F8 { float v[8]; };
struct alignas(4) U4 { uint32_t v[4]; };

template <int ST>
__device__ __forceinline__ void load_pair(const void* base, int node, float a[6]) {
    if constexpr (ST == ST_F32x3) {
        F6 t = *(const F6*)((const float*)base + (size_t)node * 3);
#pragma unroll
        for (int c = 0; c < 6; ++c) a[c] = t.v[c];
    } else if constexpr (ST == ST_F32x4) {
        F8 t = *(const F8*)((const float*)base + (size_t)node * 4);
        a[0] = t.v[0]; a[1] = t.v[1]; a[2] = t.v[2];
        a[3] = t.v[4]; a[4] = t.v[5]; a[5] = t.v[6];
    } else {
        U4 t = *(const U4*)((const char*)base + (size_t)node * 8);
        unpack_pair16<ST>(u32x4{t.v[0], t.v[1], t.v[2], t.v[3]}, a);
    }
}

// structured.py:176-220 for one axis.
__device__ __forceinline__ void axis_coord(const GridDev& g, int ax, float x, int n,
                                           int& i0, int& i1, float& w, bool& valid) {
    // true division (structured.py:176-178); when 1/d is exact (d a power of
    // two) the product is the identical correctly rounded quotient
    const float u = x - g.o[ax];
    float f = (g.inv[ax] != 0.0f) ? u * g.inv[ax] : u / g.d[ax];
    if (g.bmode == JT_BOUNDS_PERIODIC) {
        f = jmodf(f, (float)n);
    } else {
        valid = valid && (x >= g.o[ax]) && (x <= g.vhi[ax]);
        f = clipf(f, 0.0f, g.cmax[ax]);
    }
    float fl = floorf(f);
    int ii = (fl != fl) ? 0 : (int)fl;                 // NaN -> 0
    // periodic: (ii + 1) mod n; ii is in [0, n] here (jmodf can round up to n),
    // so one conditional subtract gives the remainder, and the clamp below
    // covers n = 1
    int jj = ii + 1;
    if (g.bmode == JT_BOUNDS_PERIODIC) jj = (jj >= n) ? jj - n : jj;
    else jj = min(max(jj, 0), n - 1);
    w = f - (float)ii;
    i0 = min(max(ii, 0), n - 1);                       // gather index clamp
    i1 = min(max(jj, 0), n - 1);
}

// float -> int32 by v_cvt_i32_f32, which returns 0 for NaN: the value of
// axis_coord's "(fl != fl) ? 0 : (int)fl" for every float below 2^31.
__device__ __forceinline__ int cvt_i32_nan0(float f) {
    int r;
    __asm__("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(f));
    return r;
}

// axis_coord for a periodic sampler whose spacing has an exact reciprocal
// (inv != 0) on an axis of n >= 2 nodes, with the same values for every
// input: f = jnp.mod(u * inv, n) lies in [0, n] (the last by rounding) or is
// NaN, so floor(f) needs no lower clamp and (floor(f) + 1) mod n no clamp at
// all; jnp.mod's three exact cases are selected without branches and only
// |f| >= 2n or NaN takes the general path.
__device__ __forceinline__ void axis_coord_pf(float x, float o, float inv, int n, float fn, int& i0, int& i1,
                                              float& w) {
    const float a = (x - o) * inv;
    float f;
    if (a > -fn && a < 2.0f * fn) {
        const float lo = a + fn, hi = a - fn;
        f = a < 0.0f ? lo : a;
        f = a >= fn ? hi : f;
    } else {
        f = jmodf(a, fn);
    }
    const int ii = cvt_i32_nan0(floorf(f));
    int jj = ii + 1;
    jj = (jj >= n) ? jj - n : jj;
    w = f - (float)ii;
    i0 = min(ii, n - 1);
    i1 = jj;
}

// The eight blended corner values around a point and its trilinear weights.
struct Corners {
    float v[8][3];      // corner (ci, cj, ck) at index ci*4 + cj*2 + ck
    float wx, wy, wz;
    bool valid;
};

// Blend n node values of the two snapshots with the same per-value arithmetic
// as corner(): (1-a) V0 + a V1 ('interp'), V0 + a (V1 - V0) / V1 + a (V1 - V0)
// (linear extrapolation left / right of the series).  COPY, ZERO and NAN are
// handled by the caller before V1 is loaded.
template <int N>
__device__ __forceinline__ void blend_n(const Blend& bl, const float (*A)[3], const float (*B)[3],
                                        float (*V)[3]) {
    const float al = bl.a, oma = 1.0f - bl.a;
    if (bl.mode == BL_INTERP) {
#pragma unroll
        for (int q = 0; q < N; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) V[q][c] = oma * A[q][c] + al * B[q][c];
    } else if (bl.mode == BL_LEFT) {
#pragma unroll
        for (int q = 0; q < N; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) V[q][c] = A[q][c] + al * (B[q][c] - A[q][c]);
    } else {
#pragma unroll
        for (int q = 0; q < N; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) V[q][c] = B[q][c] + al * (B[q][c] - A[q][c]);
    }
}

// Blended values of N arbitrary nodes of float32 (N, 3) snapshots, all loads
// issued before the blend-mode branch (per-value arithmetic of corner()).
template <int N>
__device__ __forceinline__ void blend_nodes_f32(const float* b0, const float* b1, const Blend& bl,
                                                const int* nodes, float (*V)[3]) {
    if (bl.mode == BL_ZERO || bl.mode == BL_NAN) {
        const float z = bl.mode == BL_ZERO ? 0.0f : __int_as_float(0x7fc00000);
#pragma unroll
        for (int q = 0; q < N; ++q) V[q][0] = V[q][1] = V[q][2] = z;
        return;
    }
    const float* bb = bl.mode == BL_COPY ? b0 : b1;
    float A[N][3], B[N][3];
#pragma unroll
    for (int q = 0; q < N; ++q) load_node<ST_F32x3>(b0, nodes[q], A[q][0], A[q][1], A[q][2]);
#pragma unroll
    for (int q = 0; q < N; ++q) load_node<ST_F32x3>(bb, nodes[q], B[q][0], B[q][1], B[q][2]);
    __asm__ volatile("" ::: "memory");                 // keep the loads ahead of the mode branch
    if (bl.mode == BL_COPY) {
#pragma unroll
        for (int q = 0; q < N; ++q) { V[q][0] = A[q][0]; V[q][1] = A[q][1]; V[q][2] = A[q][2]; }
        return;
    }
    blend_n<N>(bl, A, B, V);
}

// Bytes of one stored node.
template <int ST> struct NodeBytes { static constexpr unsigned v = (ST == ST_F32x3) ? 12u : (ST == ST_F32x4 ? 16u : 8u); };

// The eight stored corner values of one snapshot; corner (ci, cj, ck) at
// index ci*4 + cj*2 + ck.  Addresses are a uniform snapshot base plus 32-bit
// byte offsets (row[] of the (i, j) rows, k0b/k1b of the z nodes), so each
// load takes the scalar-base form with no 64-bit address arithmetic.
// ZH: the (k0, k0+1) corners are one contiguous pair.
template <int ST, bool ZH>
__device__ __forceinline__ void load_corners(const void* base, const unsigned row[4], unsigned k0b, unsigned k1b,
                                             float (*A)[3]) {
    const char* b = (const char*)base;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if constexpr (ZH) {
            float a[6];
            load_pair<ST>(b + (row[r] + k0b), 0, a);
            A[2 * r][0] = a[0]; A[2 * r][1] = a[1]; A[2 * r][2] = a[2];
            A[2 * r + 1][0] = a[3]; A[2 * r + 1][1] = a[4]; A[2 * r + 1][2] = a[5];
        } else {
            load_node<ST>(b + (row[r] + k0b), 0, A[2 * r][0], A[2 * r][1], A[2 * r][2]);
            load_node<ST>(b + (row[r] + k1b), 0, A[2 * r + 1][0], A[2 * r + 1][1], A[2 * r + 1][2]);
        }
    }
}

// The eight blended corners from snapshots b0/b1 (node arrays with y-row
// stride ny_ and z-row stride nzs_).  The blend mode is dispatched once, so
// all corner loads of both snapshots issue together.
// M24: the row offsets by 24-bit multiplies (full rate), valid when
// nx * ny and nzs * node bytes are below 2^24 (the periodic fast path checks
// it on the host).
template <int ST, bool ZH, bool M24 = false>
__device__ __forceinline__ void gather_from(const void* b0, const void* b1, const Blend& bl, int i0, int i1,
                                            int j0, int j1, int k0, int k1, int ny_, int nzs_, Corners& C) {
    if (bl.mode == BL_ZERO || bl.mode == BL_NAN) {
        const float z = bl.mode == BL_ZERO ? 0.0f : __int_as_float(0x7fc00000);
#pragma unroll
        for (int q = 0; q < 8; ++q) C.v[q][0] = C.v[q][1] = C.v[q][2] = z;
        return;
    }
    // 32-bit byte offsets (a snapshot is below 4 GiB, checked on the host)
    const unsigned NB = NodeBytes<ST>::v;
    const unsigned rs = (unsigned)nzs_ * NB;
    unsigned row[4];                                    // rows in corner order (ci, cj): (0,0) (0,1) (1,0) (1,1)
    if constexpr (M24) {
        const unsigned r0 = __umul24((unsigned)i0, (unsigned)ny_), r1 = __umul24((unsigned)i1, (unsigned)ny_);
        row[0] = __umul24(r0 + (unsigned)j0, rs);
        row[1] = __umul24(r0 + (unsigned)j1, rs);
        row[2] = __umul24(r1 + (unsigned)j0, rs);
        row[3] = __umul24(r1 + (unsigned)j1, rs);
    } else {
        const unsigned ps = (unsigned)ny_ * rs;
        const unsigned oi0 = (unsigned)i0 * ps, oi1 = (unsigned)i1 * ps;
        const unsigned oj0 = (unsigned)j0 * rs, oj1 = (unsigned)j1 * rs;
        row[0] = oi0 + oj0; row[1] = oi0 + oj1; row[2] = oi1 + oj0; row[3] = oi1 + oj1;
    }
    const unsigned k0b = (unsigned)k0 * NB, k1b = (unsigned)k1 * NB;
    // both snapshots' loads issue before any wait ('copy' reads V0 twice)
    float A[8][3], B[8][3];
    if constexpr (ZH && NodeBytes<ST>::v == 8u) {
        // 16-bit z-pairs: the eight raw 16 B loads are all held until the
        // asm below, so none waits for an earlier one's registers
        const char* pa = (const char*)b0;
        const char* pb = (const char*)(bl.mode == BL_COPY ? b0 : b1);
        u32x4 ra[4], rb[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const U4 ta = *(const U4*)(pa + (row[r] + k0b));
            const U4 tb = *(const U4*)(pb + (row[r] + k0b));
            ra[r] = u32x4{ta.v[0], ta.v[1], ta.v[2], ta.v[3]};
            rb[r] = u32x4{tb.v[0], tb.v[1], tb.v[2], tb.v[3]};
        }
        __asm__ volatile("" : "+v"(ra[0]), "+v"(ra[1]), "+v"(ra[2]), "+v"(ra[3]),
                              "+v"(rb[0]), "+v"(rb[1]), "+v"(rb[2]), "+v"(rb[3]) :: "memory");
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float a[6], b[6];
            unpack_pair16<ST>(ra[r], a);
            unpack_pair16<ST>(rb[r], b);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                A[2 * r][c] = a[c]; A[2 * r + 1][c] = a[3 + c];
                B[2 * r][c] = b[c]; B[2 * r + 1][c] = b[3 + c];
            }
        }
    } else {
        load_corners<ST, ZH>(b0, row, k0b, k1b, A);
        load_corners<ST, ZH>(bl.mode == BL_COPY ? b0 : b1, row, k0b, k1b, B);
        __asm__ volatile("" ::: "memory");             // keep the loads ahead of the mode branch
    }
    if (bl.mode == BL_COPY) {
#pragma unroll
        for (int q = 0; q < 8; ++q) { C.v[q][0] = A[q][0]; C.v[q][1] = A[q][1]; C.v[q][2] = A[q][2]; }
        return;
    }
    blend_n<8>(bl, A, B, C.v);
}

template <int ST>
__device__ __forceinline__ const void* snap_base(const GridDev& g, int snap) {
    return (const char*)g.data + (size_t)snap * (size_t)g.snap_stride * ElemBytes<ST>::v;
}

// Gather half of the trilinear sample (structured.py:176-232 after the blend).
// PF: the periodic fast path (periodic bounds, exact reciprocal spacing on
// every axis, every axis at least 2 nodes), chosen on the host.
template <int ST, bool ZH = false, bool PF = false>
__device__ __forceinline__ void gather_corners(const GridDev& g, const Blend& bl,
                                               float px, float py, float pz, Corners& C) {
    int i0, i1, j0, j1, k0, k1;
    C.valid = true;
    if constexpr (PF) {
        axis_coord_pf(px, g.o[0], g.inv[0], g.nx, (float)g.nx, i0, i1, C.wx);
        axis_coord_pf(py, g.o[1], g.inv[1], g.ny, (float)g.ny, j0, j1, C.wy);
        axis_coord_pf(pz, g.o[2], g.inv[2], g.nz, (float)g.nz, k0, k1, C.wz);
    } else {
        axis_coord(g, 0, px, g.nx, i0, i1, C.wx, C.valid);
        axis_coord(g, 1, py, g.ny, j0, j1, C.wy, C.valid);
        axis_coord(g, 2, pz, g.nz, k0, k1, C.wz, C.valid);
    }
    gather_from<ST, ZH, PF>(snap_base<ST>(g, bl.i0), snap_base<ST>(g, bl.i1), bl, i0, i1, j0, j1, k0, k1, g.ny,
                        g.nzs, C);
}

// Interpolation half: lerp along x, then y, then z (structured.py:234-247).
__device__ __forceinline__ void lerp_corners(const GridDev& g, const Corners& C, float out[3]) {
    const float wx = C.wx, wy = C.wy, wz = C.wz;
    const float ax = 1.0f - wx, ay = 1.0f - wy, az = 1.0f - wz;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float c00 = C.v[0][c] * ax + C.v[4][c] * wx;   // :234-237
        float c10 = C.v[2][c] * ax + C.v[6][c] * wx;
        float c01 = C.v[1][c] * ax + C.v[5][c] * wx;
        float c11 = C.v[3][c] * ax + C.v[7][c] * wx;
        float c0 = c00 * ay + c10 * wy;                 // :239-240
        float c1 = c01 * ay + c11 * wy;
        out[c] = c0 * az + c1 * wz;                     // :242
    }
    if (!C.valid) {                                     // :244-247
        if (g.bmode == JT_BOUNDS_ZERO) { out[0] = out[1] = out[2] = 0.0f; }
        else if (g.bmode == JT_BOUNDS_NAN) { out[0] = out[1] = out[2] = __int_as_float(0x7fc00000); }
    }
}

// Blended trilinear sample u(x, t) (structured.py:167-250 after the blend).
template <int ST, bool ZH = false, bool PF = false>
__device__ __forceinline__ void sample(const GridDev& g, const Blend& bl,
                                       float px, float py, float pz, float out[3]) {
    Corners C;
    gather_corners<ST, ZH, PF>(g, bl, px, py, pz, C);
    lerp_corners(g, C, out);
}

// ---- corner reuse across stages (z-halo layout, periodic fast path) ------
// Consecutive stage points of one particle mostly stay in one cell, and the
// stored snapshots do not change during a launch, so a lane keeps the raw
// corner records of its last sample in registers and reloads them only when
// its cell (i0, i1, j0, j1, k0) or the snapshot pair changes.  The reloads
// are exec-masked loads: lanes that reuse issue no memory requests, which is
// the vector-memory address/data work the tracking kernel is bound by.  The
// values used are the same stored values, so results are unchanged.
// 16-bit records are unpacked inside the reload, so a stage where no lane
// reloads skips the unpack too (measured faster than caching the raw records,
// which take 16 fewer registers).
template <int ST>
struct CornerCache {
    float A[8][3], B[8][3];          // the two snapshots' corner values
    int i0, i1, j0, j1, k0;
    const char* pa;
    const char* pb;
};

template <int ST>
__device__ __forceinline__ void cache_reset(CornerCache<ST>& cc) {
    cc.i0 = -1; cc.i1 = -1; cc.j0 = -1; cc.j1 = -1; cc.k0 = -1;
    cc.pa = nullptr; cc.pb = nullptr;
}

// gather_corners<ST, true, true> through a CornerCache.  IO: the caller
// knows the blend is 'interp' (a per-step uniform test), so no mode branches.
template <int ST, bool IO = false>
__device__ __forceinline__ void gather_cached(const GridDev& g, const Blend& bl, float px, float py, float pz,
                                              Corners& C, CornerCache<ST>& cc) {
    C.valid = true;
    int i0, i1, j0, j1, k0, k1;
    axis_coord_pf(px, g.o[0], g.inv[0], g.nx, (float)g.nx, i0, i1, C.wx);
    axis_coord_pf(py, g.o[1], g.inv[1], g.ny, (float)g.ny, j0, j1, C.wy);
    axis_coord_pf(pz, g.o[2], g.inv[2], g.nz, (float)g.nz, k0, k1, C.wz);
    if (!IO && (bl.mode == BL_ZERO || bl.mode == BL_NAN)) {
        const float z = bl.mode == BL_ZERO ? 0.0f : __int_as_float(0x7fc00000);
#pragma unroll
        for (int q = 0; q < 8; ++q) C.v[q][0] = C.v[q][1] = C.v[q][2] = z;
        return;
    }
    const char* pa = (const char*)snap_base<ST>(g, bl.i0);
    const char* pb = (!IO && bl.mode == BL_COPY) ? pa : (const char*)snap_base<ST>(g, bl.i1);
    const bool same = pa == cc.pa && pb == cc.pb && i0 == cc.i0 && i1 == cc.i1 && j0 == cc.j0 && j1 == cc.j1 &&
                      k0 == cc.k0;
    if (!same) {
        const unsigned NB = NodeBytes<ST>::v;
        const unsigned rs = (unsigned)g.nzs * NB;
        const unsigned r0 = (unsigned)__umul24((unsigned)i0, (unsigned)g.ny);
        const unsigned r1 = (unsigned)__umul24((unsigned)i1, (unsigned)g.ny);
        const unsigned row[4] = {(unsigned)__umul24(r0 + (unsigned)j0, rs), (unsigned)__umul24(r0 + (unsigned)j1, rs),
                                 (unsigned)__umul24(r1 + (unsigned)j0, rs), (unsigned)__umul24(r1 + (unsigned)j1, rs)};
        const unsigned k0b = (unsigned)k0 * NB;
        if constexpr (NodeBytes<ST>::v == 8u) {
            u32x4 ra[4], rb[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const U4 ta = *(const U4*)(pa + (row[r] + k0b));
                const U4 tb = *(const U4*)(pb + (row[r] + k0b));
                ra[r] = u32x4{ta.v[0], ta.v[1], ta.v[2], ta.v[3]};
                rb[r] = u32x4{tb.v[0], tb.v[1], tb.v[2], tb.v[3]};
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                float a[6], b[6];
                unpack_pair16<ST>(ra[r], a);
                unpack_pair16<ST>(rb[r], b);
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    cc.A[2 * r][c] = a[c]; cc.A[2 * r + 1][c] = a[3 + c];
                    cc.B[2 * r][c] = b[c]; cc.B[2 * r + 1][c] = b[3 + c];
                }
            }
        } else {
            load_corners<ST, true>(pa, row, k0b, 0u, cc.A);
            load_corners<ST, true>(pb, row, k0b, 0u, cc.B);
        }
        cc.i0 = i0; cc.i1 = i1; cc.j0 = j0; cc.j1 = j1; cc.k0 = k0;
    }
    cc.pa = pa;
    cc.pb = pb;
    const float (*A)[3] = cc.A;
    const float (*B)[3] = cc.B;
    if constexpr (IO) {
        const float al = bl.a, oma = 1.0f - bl.a;      // blend_n's 'interp' arithmetic
#pragma unroll
        for (int q = 0; q < 8; ++q)
#pragma unroll
            for (int c = 0; c < 3; ++c) C.v[q][c] = oma * A[q][c] + al * B[q][c];
    } else if (bl.mode == BL_COPY) {
#pragma unroll
        for (int q = 0; q < 8; ++q) { C.v[q][0] = A[q][0]; C.v[q][1] = A[q][1]; C.v[q][2] = A[q][2]; }
    } else {
        blend_n<8>(bl, A, B, C.v);
    }
}

// gather_cached<ST_F32x4, true> from the pair records of snapshots
// (i0, i0 + 1): one row of a cell is 12 consecutive floats (48 B, three
// 16-byte loads) holding both snapshots' z-pairs, against four 16-byte loads
// of the padded per-snapshot layout.  Same values, same arithmetic.
__device__ __forceinline__ void gather_cached_pair(const GridDev& g, const Blend& bl, float px, float py, float pz,
                                                   Corners& C, CornerCache<ST_F32x4>& cc) {
    C.valid = true;
    int i0, i1, j0, j1, k0, k1;
    axis_coord_pf(px, g.o[0], g.inv[0], g.nx, (float)g.nx, i0, i1, C.wx);
    axis_coord_pf(py, g.o[1], g.inv[1], g.ny, (float)g.ny, j0, j1, C.wy);
    axis_coord_pf(pz, g.o[2], g.inv[2], g.nz, (float)g.nz, k0, k1, C.wz);
    const char* pp = (const char*)(g.pair + (size_t)bl.i0 * (size_t)g.pair_stride);
    const bool same = pp == cc.pa && pp == cc.pb && i0 == cc.i0 && i1 == cc.i1 && j0 == cc.j0 && j1 == cc.j1 &&
                      k0 == cc.k0;
    if (!same) {
        const unsigned rs = (unsigned)g.nz * 48u;
        const unsigned r0 = (unsigned)__umul24((unsigned)i0, (unsigned)g.ny);
        const unsigned r1 = (unsigned)__umul24((unsigned)i1, (unsigned)g.ny);
        const unsigned kb = (unsigned)k0 * 48u;
        const unsigned row[4] = {(unsigned)__umul24(r0 + (unsigned)j0, rs) + kb,
                                 (unsigned)__umul24(r0 + (unsigned)j1, rs) + kb,
                                 (unsigned)__umul24(r1 + (unsigned)j0, rs) + kb,
                                 (unsigned)__umul24(r1 + (unsigned)j1, rs) + kb};
        float4 f[4][3];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float4* q = (const float4*)(pp + row[r]);
            f[r][0] = q[0]; f[r][1] = q[1]; f[r][2] = q[2];
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cc.A[2 * r][0] = f[r][0].x; cc.A[2 * r][1] = f[r][0].y; cc.A[2 * r][2] = f[r][0].z;
            cc.A[2 * r + 1][0] = f[r][0].w; cc.A[2 * r + 1][1] = f[r][1].x; cc.A[2 * r + 1][2] = f[r][1].y;
            cc.B[2 * r][0] = f[r][1].z; cc.B[2 * r][1] = f[r][1].w; cc.B[2 * r][2] = f[r][2].x;
            cc.B[2 * r + 1][0] = f[r][2].y; cc.B[2 * r + 1][1] = f[r][2].z; cc.B[2 * r + 1][2] = f[r][2].w;
        }
        cc.i0 = i0; cc.i1 = i1; cc.j0 = j0; cc.j1 = j1; cc.k0 = k0;
    }
    cc.pa = pp;
    cc.pb = pp;
    const float al = bl.a, oma = 1.0f - bl.a;      // blend_n's 'interp' arithmetic
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) C.v[q][c] = oma * cc.A[q][c] + al * cc.B[q][c];
}

// sample<ST, true, true> through a CornerCache.
template <int ST, bool IO = false>
__device__ __forceinline__ void sample_cached(const GridDev& g, const Blend& bl, float px, float py, float pz,
                                              float out[3], CornerCache<ST>& cc) {
    Corners C;
    gather_cached<ST, IO>(g, bl, px, py, pz, C, cc);
    lerp_corners(g, C, out);
}

// Whether the periodic fast path applies to a grid (its row offsets use
// 24-bit multiplies: nx * ny and nzs * node bytes below 2^24).
inline bool periodic_fast(const GridDev& g, int node_bytes) {
    return g.bmode == JT_BOUNDS_PERIODIC && g.inv[0] != 0.0f && g.inv[1] != 0.0f && g.inv[2] != 0.0f &&
           g.nx >= 2 && g.ny >= 2 && g.nz >= 2 && (long long)g.nx * g.ny < (1LL << 24) &&
           (long long)g.nzs * node_bytes < (1LL << 24);
}

// Waves per SIMD the tracking kernels' register budget is sized for (128
// VGPRs at 4).  The per-lane corner cache of the fast path holds 48 floats
// across stages; at 4 waves the RK4 kernels spill (a private segment of
// 112 B float32 / 108 B 16-bit per lane) and still run faster than the
// spill-free 3-wave build (9.8 against 10.4 ms and 7.7 against 8.9 ms per 25
// steps of 2^24 particles); 5 waves is far slower.
#ifndef JT_WAVES_F32
#define JT_WAVES_F32 4
#endif
#ifndef JT_WAVES_16
#define JT_WAVES_16 4
#endif
template <int ST> struct TrackWaves {
    static constexpr int v = (ST == ST_F32x3 || ST == ST_F32x4) ? JT_WAVES_F32 : JT_WAVES_16;
};

// ---- boundary epilogue (boundary.py) ----------------------------------------
struct BcDev {
    int kind[3];
    float lo[3], hi[3];
    // periodic / reflective axes as one formula (apply_bc_axes): y = jnp.mod(x
    // - lo, per), then y <= rw ? y : rw2 - y, then lo + y.  Periodic: per =
    // width (1 if empty), rw = +inf; reflective: per = 2 * width (2 if
    // empty), rw = width, rw2 = 2 * width.
    float per[3], rw[3], rw2[3];
    int absorb_rows;
    float outside;
    int inlet, ia, ipos;
    float backstop;
    const float* coords[2];
    int nc[2];
};

__device__ __forceinline__ float bc_axis(int kind, float x, float lo, float hi, float outside) {
    if (kind == JT_BC_PERIODIC) {                       // :117-123
        float w = hi - lo;
        w = w < 0.0f ? 0.0f : w;
        float ws = w > 0.0f ? w : 1.0f;
        return lo + jmodf(x - lo, ws);
    } else if (kind == JT_BC_REFLECTIVE) {              // :165-172
        float w = hi - lo;
        w = w < 0.0f ? 0.0f : w;
        float y = jmodf(x - lo, 2.0f * (w > 0.0f ? w : 1.0f));
        float yr = (y <= w) ? y : 2.0f * w - y;
        return lo + yr;
    } else if (kind == JT_BC_CLAMP) {                   // :143
        return clipf(x, lo, hi);
    } else if (kind == JT_BC_ABSORB) {                  // per-axis, as mixed_boundary applies it
        return (x < lo || x > hi) ? outside : x;
    }
    return x;
}

__device__ __forceinline__ float snap_nearest(const float* c, int nc, float v) {
    int best = 0;
    float bd = fabsf(c[0] - v);
    for (int k = 1; k < nc; ++k) {
        float dk = fabsf(c[k] - v);
        if (dk < bd) { bd = dk; best = k; }
    }
    return c[best];
}

__device__ __forceinline__ void apply_bc(const BcDev& b, float p[3]) {
    if (b.inlet != JT_INLET_OFF) {                      // boundary.py:1515-1599
        const int a = b.ia;
        const int o0 = (a == 0) ? 1 : 0;
        const int o1 = (a == 2) ? 1 : 2;
        const bool exited = b.ipos ? (p[a] > b.hi[a]) : (p[a] < b.lo[a]);
        if (exited) {
            float g0 = 0.0f, g1 = 0.0f;
            if (b.inlet == JT_INLET_GRID) {             // :1301-1317
                g0 = snap_nearest(b.coords[0], b.nc[0], p[o0]);
                g1 = snap_nearest(b.coords[1], b.nc[1], p[o1]);
            }
            p[a] = b.ipos ? b.lo[a] : b.hi[a];
            p[o0] = g0;
            p[o1] = g1;
        }
        if (b.ipos) { if (p[a] < b.lo[a]) p[a] = b.backstop; }   // :1580-1587
        else        { if (p[a] > b.hi[a]) p[a] = b.backstop; }
#pragma unroll
        for (int i = 0; i < 3; ++i) {                   // :1590-1597
            if (i == a) continue;
            float v = p[i];
            if (v < b.lo[i]) p[i] = 2.0f * b.lo[i] - v;
            else if (v > b.hi[i]) p[i] = 2.0f * b.hi[i] - v;
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) p[i] = bc_axis(b.kind[i], p[i], b.lo[i], b.hi[i], b.outside);
    if (b.absorb_rows) {                                // :190-210
        bool out = false;
#pragma unroll
        for (int i = 0; i < 3; ++i) out = out || (p[i] < b.lo[i]) || (p[i] > b.hi[i]);
        if (out) { p[0] = p[1] = p[2] = b.outside; }
    }
}

// jnp.mod(a, m) for m > 0 with jmodf's values: its three exact cases by
// selects, the general path only for |a| beyond them or NaN.
__device__ __forceinline__ float jmod_pos(float a, float m) {
    if (a > -m && a < 2.0f * m) {
        const float lo = a + m, hi = a - m;
        float f = a < 0.0f ? lo : a;
        return a >= m ? hi : f;
    }
    return jmodf(a, m);
}

// apply_bc for boundaries of per-axis none / periodic / reflective /
// clamping kinds only (no inlet, no absorbing rows), with the same values.
__device__ __forceinline__ void apply_bc_axes(const BcDev& b, float p[3]) {
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const int k = b.kind[i];
        if (k == JT_BC_PERIODIC || k == JT_BC_REFLECTIVE) {
            float y = jmod_pos(p[i] - b.lo[i], b.per[i]);
            y = (y <= b.rw[i]) ? y : b.rw2[i] - y;
            p[i] = b.lo[i] + y;
        } else if (k == JT_BC_CLAMP) {
            p[i] = clipf(p[i], b.lo[i], b.hi[i]);
        }
    }
}

// Whether apply_bc_axes covers a boundary.
inline bool bc_axes_only(const BcDev& b) {
    if (b.inlet != JT_INLET_OFF || b.absorb_rows) return false;
    for (int i = 0; i < 3; ++i)
        if (b.kind[i] == JT_BC_ABSORB) return false;
    return true;
}

// Stage-time blend table: 3 entries per step (stage origin, midpoint, end).
// rk4.py:88-92 (f32 t + 0.5*dt, t + dt); rk2.py:165 (float64 midpoint sum).
static __global__ void k_blend_table(GridDev g, const float* __restrict__ step_times, int n_steps,
                              float dt, int integrator, Blend* __restrict__ tab) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_steps) return;
    const float t = step_times[s];
    float t1 = t, t2 = t;
    if (integrator == JT_RK4) {
        const float dth = 0.5f * dt;
        t1 = t + dth;
        t2 = t + dt;
    } else if (integrator == JT_RK2) {
        const float dth = dt * 0.5f;
        t1 = (float)((double)t + (double)dth);
    }
    tab[3 * s + 0] = time_blend(g.times, g.T, t, g.interp, g.extrap);
    tab[3 * s + 1] = time_blend(g.times, g.T, t1, g.interp, g.extrap);
    tab[3 * s + 2] = time_blend(g.times, g.T, t2, g.interp, g.extrap);
}

// One integrator step (rk4.py:84-106, rk2.py:157-177, euler.py:150-157) for
// any sampler f(blend, q, out); tab holds the step's three stage blends.
template <int INTEG, class F>
__device__ __forceinline__ void rk_step(F&& f, const Blend* __restrict__ tab, int s, float p[3], float dt,
                                        float dth, float dt6) {
    const Blend b0 = tab[3 * s];
    if constexpr (INTEG == JT_RK4) {
        // acc accumulates ((k1 + 2 k2) + 2 k3) + k4 in that order as each
        // stage completes, so only one k is live at a time
        const Blend b1 = tab[3 * s + 1];
        const Blend b2 = tab[3 * s + 2];
        float k[3], q[3], acc[3];
        f(b0, p, k);
#pragma unroll
        for (int c = 0; c < 3; ++c) { acc[c] = k[c]; q[c] = p[c] + dth * k[c]; }
        f(b1, q, k);
#pragma unroll
        for (int c = 0; c < 3; ++c) { acc[c] = acc[c] + 2.0f * k[c]; q[c] = p[c] + dth * k[c]; }
        __asm__ volatile("" : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]));   // sum now, not at the end
        f(b1, q, k);
#pragma unroll
        for (int c = 0; c < 3; ++c) { acc[c] = acc[c] + 2.0f * k[c]; q[c] = p[c] + dt * k[c]; }
        __asm__ volatile("" : "+v"(acc[0]), "+v"(acc[1]), "+v"(acc[2]));
        f(b2, q, k);
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = p[c] + dt6 * (acc[c] + k[c]);   // rk4.py:44
    } else if constexpr (INTEG == JT_RK2) {
        const Blend b1 = tab[3 * s + 1];
        float v1[3], v2[3], q[3];
        f(b0, p, v1);
#pragma unroll
        for (int c = 0; c < 3; ++c) q[c] = p[c] + dth * v1[c];
        f(b1, q, v2);
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = p[c] + dt * v2[c];
    } else {
        float v[3];
        f(b0, p, v);
#pragma unroll
        for (int c = 0; c < 3; ++c) p[c] = p[c] + dt * v[c];
    }
}

// One step of the fast path (z-halo layout, periodic sampler with exact
// reciprocal spacing, per-axis boundary kinds) through the per-lane corner
// cache, then the boundary.  A step whose three stage blends are all
// 'interp' (a uniform test) runs a copy with no blend-mode branches; the
// rare other steps sample without the cache and leave it empty, so the cache
// registers are not live across that copy.
template <int INTEG, int ST>
__device__ __forceinline__ void fast_step(const GridDev& g, const BcDev& bc, const Blend* __restrict__ tab, int s,
                                          float p[3], float dt, float dth, float dt6, CornerCache<ST>& cc) {
    const Blend* t = tab + 3 * s;
    if (t[0].mode == BL_INTERP && t[1].mode == BL_INTERP && t[2].mode == BL_INTERP) {
        auto fi = [&](const Blend& b, const float q[3], float out[3]) {
            sample_cached<ST, true>(g, b, q[0], q[1], q[2], out, cc);
        };
        rk_step<INTEG>(fi, tab, s, p, dt, dth, dt6);
    } else {
        auto f = [&](const Blend& b, const float q[3], float out[3]) {
            sample<ST, true, true>(g, b, q[0], q[1], q[2], out);
        };
        rk_step<INTEG>(f, tab, s, p, dt, dth, dt6);
        cache_reset(cc);
    }
    apply_bc_axes(bc, p);                               // tracker.py:238-239
}

// fast_step over the consecutive-pair records (float32 fields that carry
// them): a step whose three stage blends are 'interp' between snapshots
// (i0, i0 + 1) gathers through gather_cached_pair; any other step samples the
// per-snapshot layout without the cache and empties it.
template <int INTEG>
__device__ __forceinline__ void fast_step_pair(const GridDev& g, const BcDev& bc, const Blend* __restrict__ tab, int s,
                                               float p[3], float dt, float dth, float dt6,
                                               CornerCache<ST_F32x4>& cc) {
    const Blend* t = tab + 3 * s;
    if (t[0].mode == BL_INTERP && t[1].mode == BL_INTERP && t[2].mode == BL_INTERP && t[0].i1 == t[0].i0 + 1 &&
        t[1].i1 == t[1].i0 + 1 && t[2].i1 == t[2].i0 + 1) {
        auto fp = [&](const Blend& b, const float q[3], float out[3]) {
            Corners C;
            gather_cached_pair(g, b, q[0], q[1], q[2], C, cc);
            lerp_corners(g, C, out);
        };
        rk_step<INTEG>(fp, tab, s, p, dt, dth, dt6);
    } else {
        auto f = [&](const Blend& b, const float q[3], float out[3]) {
            sample<ST_F32x4, true, true>(g, b, q[0], q[1], q[2], out);
        };
        rk_step<INTEG>(f, tab, s, p, dt, dth, dt6);
        cache_reset(cc);
    }
    apply_bc_axes(bc, p);                               // tracker.py:238-239
}

// Host-side translation of the public descriptors.
int make_grid_dev(const jt_grid_field* f, GridDev& g, int& store_kind);
int make_bc_dev(const jt_boundary* b, BcDev& d);

}  // namespace jt
