// Fused multi-step particle tracking through a time-blended structured grid.
//
// One thread owns one particle and keeps its position in registers across all
// stages and all steps of a launch: the reference's lax.scan over
// rk4_step -> field_fn -> boundary_fn (jaxtrace/tracking/tracker.py:228-246)
// materialises k1..k4 and the blended node arrays in memory every stage; here
// each stage is 8 corner gathers from each of the two bracketing snapshots
// (4 paired loads with the z-halo layout).
//
// Locality: optionally (sort_every > 0) particles are re-ordered every k steps
// by the Morton code of their grid cell, so the 64 lanes of a wavefront gather
// from one small block of cells; trajectories are written back in the
// caller's particle order through the permutation.  Blocks are mapped to
// particles XCD-contiguously: the hardware deals workgroups round-robin over
// the 8 XCDs, so block b takes chunk (b % 8) * (grid / 8) + b / 8 and each
// XCD's L2 sees one contiguous (spatially compact, once sorted) particle range.
#include "grid_sampler.h"
#include "cell_sort.h"
#include <algorithm>
#include <cmath>
#include <mutex>

namespace jt {

int make_grid_dev(const jt_grid_field* f, GridDev& g, int& st) {
    JT_REQUIRE(f != nullptr, "field descriptor is NULL");
    JT_REQUIRE(f->data != nullptr && f->times != nullptr, "field data/times pointer is NULL");
    JT_REQUIRE(f->n_snap >= 1, "n_snap must be >= 1 (got %d)", f->n_snap);
    JT_REQUIRE(f->nx >= 1 && f->ny >= 1 && f->nz >= 1, "grid dims must be >= 1");
    JT_REQUIRE(f->ncomp == 3 || f->ncomp == 4, "ncomp must be 3 or 4 (got %d)", f->ncomp);
    JT_REQUIRE(f->bounds_mode >= JT_BOUNDS_CLAMP && f->bounds_mode <= JT_BOUNDS_PERIODIC,
               "bad bounds_mode %d", f->bounds_mode);
    JT_REQUIRE(f->time_interp == JT_TIME_LINEAR || f->time_interp == JT_TIME_NEAREST,
               "bad time_interp %d", f->time_interp);
    JT_REQUIRE(f->time_extrap >= JT_EXTRAP_CONSTANT && f->time_extrap <= JT_EXTRAP_ZERO,
               "bad time_extrap %d", f->time_extrap);
    JT_REQUIRE(f->z_halo == 0 || f->z_halo == 1, "z_halo must be 0 or 1");
    long long nodes = (long long)f->nx * f->ny * (f->nz + f->z_halo);
    JT_REQUIRE(nodes * f->ncomp < (1LL << 31), "grid too large for 32-bit node offsets");
    {   // the gather addresses one snapshot with 32-bit byte offsets
        const long long eb = (f->dtype == JT_DTYPE_F32) ? 4 : 2;
        if (nodes * f->ncomp * eb >= (1LL << 32))
            return fail(JT_EUNSUPPORTED, "one snapshot is %lld bytes; at most 4 GiB per snapshot is supported",
                        nodes * f->ncomp * eb);
    }
    if (f->dtype == JT_DTYPE_F32) st = (f->ncomp == 3) ? ST_F32x3 : ST_F32x4;
    else if (f->dtype == JT_DTYPE_BF16) st = ST_BF16x4;
    else if (f->dtype == JT_DTYPE_F16) st = ST_F16x4;
    else return fail(JT_EINVAL, "bad dtype %d", f->dtype);
    if (f->dtype != JT_DTYPE_F32 && f->ncomp != 4)
        return fail(JT_EUNSUPPORTED, "16-bit snapshots must be stored with ncomp=4");
    g.data = f->data;
    g.snap_stride = nodes * f->ncomp;
    g.pair = nullptr;
    g.pair_stride = 0;
    if (f->pair_data != nullptr) {
        JT_REQUIRE(f->dtype == JT_DTYPE_F32 && f->ncomp == 4 && f->z_halo == 1,
                   "pair_data needs a float32 field with 4 components and a z halo");
        const long long ps = (long long)f->nx * f->ny * f->nz * 12;
        JT_REQUIRE(ps * 4 < (1LL << 32), "pair records of one snapshot pair exceed 4 GiB");
        g.pair = f->pair_data;
        g.pair_stride = ps;
    }
    g.times = f->times;
    g.T = f->n_snap;
    g.nx = f->nx; g.ny = f->ny; g.nz = f->nz;
    g.nzs = f->nz + f->z_halo;
    const int n[3] = {f->nx, f->ny, f->nz};
    for (int a = 0; a < 3; ++a) {
        g.o[a] = f->origin[a];
        g.d[a] = f->spacing[a];
        {   // power-of-two spacing: its reciprocal is exact
            int e;
            float m = frexpf(f->spacing[a], &e);
            g.inv[a] = (m == 0.5f && std::isfinite(f->spacing[a])) ? 1.0f / f->spacing[a] : 0.0f;
        }
        // structured.py:188-190: x_min + (N-1)*dx, int32*f32 -> f32
        float nm1 = (float)(n[a] - 1);
        float span = nm1 * f->spacing[a];
        g.vhi[a] = f->origin[a] + span;
        g.cmax[a] = (float)(n[a] - 1 > 0 ? n[a] - 1 : 0);
    }
    g.bmode = f->bounds_mode;
    g.interp = f->time_interp;
    g.extrap = f->time_extrap;
    return JT_OK;
}

int make_bc_dev(const jt_boundary* b, BcDev& d) {
    if (b == nullptr) {
        for (int i = 0; i < 3; ++i) {
            d.kind[i] = JT_BC_NONE; d.lo[i] = 0; d.hi[i] = 0;
            d.per[i] = 1.0f; d.rw[i] = INFINITY; d.rw2[i] = 0.0f;
        }
        d.absorb_rows = 0; d.outside = 0; d.inlet = JT_INLET_OFF; d.ia = 0; d.ipos = 1;
        d.backstop = 0; d.coords[0] = d.coords[1] = nullptr; d.nc[0] = d.nc[1] = 0;
        return JT_OK;
    }
    for (int i = 0; i < 3; ++i) {
        JT_REQUIRE(b->axis_kind[i] >= JT_BC_NONE && b->axis_kind[i] <= JT_BC_ABSORB,
                   "bad boundary kind %d on axis %d", b->axis_kind[i], i);
        d.kind[i] = b->axis_kind[i];
        d.lo[i] = b->lo[i];
        d.hi[i] = b->hi[i];
        // the constants of apply_bc_axes, in bc_axis's float32 operations
        float w = b->hi[i] - b->lo[i];
        w = w < 0.0f ? 0.0f : w;
        const float ws = w > 0.0f ? w : 1.0f;
        if (b->axis_kind[i] == JT_BC_REFLECTIVE) {
            d.per[i] = 2.0f * ws;
            d.rw[i] = w;
            d.rw2[i] = 2.0f * w;
        } else {
            d.per[i] = ws;
            d.rw[i] = INFINITY;
            d.rw2[i] = 0.0f;
        }
    }
    d.absorb_rows = b->absorb_rows;
    d.outside = b->outside_value;
    JT_REQUIRE(b->inlet >= JT_INLET_OFF && b->inlet <= JT_INLET_HOST, "bad inlet mode %d", b->inlet);
    d.inlet = b->inlet;
    d.ia = b->inlet_axis;
    d.ipos = b->inlet_positive ? 1 : 0;
    d.backstop = b->inlet_backstop;
    if (b->inlet != JT_INLET_OFF)
        JT_REQUIRE(b->inlet_axis >= 0 && b->inlet_axis <= 2, "bad inlet axis %d", b->inlet_axis);
    for (int j = 0; j < 2; ++j) {
        d.coords[j] = b->inlet_coords[j];
        d.nc[j] = b->inlet_ncoords[j];
        if (b->inlet == JT_INLET_GRID)
            JT_REQUIRE(d.coords[j] != nullptr && d.nc[j] >= 1, "inlet grid coords missing");
    }
    return JT_OK;
}

// ---- kernels -------------------------------------------------------------

// Where one tracking launch reads and writes its particles.  perm == nullptr:
// in place (xin == xout), thread i owns slot i, whose caller index is idin[i]
// (i without ids).  Otherwise thread i takes particle perm[i] of xin/idin (the
// cell order just sorted) and leaves it in slot i of xout/idout, so the
// re-ordering costs no pass of its own.
struct Reorder {
    const float* xin;
    float* xout;
    const int* idin;
    int* idout;
    const int* perm;
};

#ifndef JT_BLOCK_F32
#define JT_BLOCK_F32 256
#endif
#ifndef JT_BLOCK_16
#define JT_BLOCK_16 256
#endif
template <int ST> struct TrackBlock {
    static constexpr int v = (ST == ST_F32x3 || ST == ST_F32x4) ? JT_BLOCK_F32 : JT_BLOCK_16;
};

// PF: the fast path (fast_step: periodic sampler index math, per-lane corner
// reuse, per-axis boundary epilogue), chosen on the host where it applies.
// PR: the fast path over the field's consecutive-pair records
// (fast_step_pair), chosen on the host when a float32 field carries them.
template <int INTEG, int ST, bool ZH, bool PF, bool PR = false>
__global__ void __launch_bounds__(TrackBlock<ST>::v, TrackWaves<ST>::v)
k_track(GridDev g, BcDev bc, Reorder ro, int n,
        const Blend* __restrict__ tab, int s0, int s1, float dt, float dth, float dt6,
        int rec_every, float* __restrict__ traj) {
    const int per = gridDim.x >> 3;                     // gridDim.x is a multiple of 8
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    const int i = lb * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int src = ro.perm ? ro.perm[i] : i;
    float p[3] = {ro.xin[3 * (size_t)src], ro.xin[3 * (size_t)src + 1], ro.xin[3 * (size_t)src + 2]};
    const int id = ro.idin ? ro.idin[src] : src;
    if (ro.idout) ro.idout[i] = id;
    const size_t dst = (size_t)id;
    auto f = [&](const Blend& b, const float q[3], float out[3]) { sample<ST, ZH, PF>(g, b, q[0], q[1], q[2], out); };
    CornerCache<ST> cc;
    cache_reset(cc);
    for (int s = s0; s < s1; ++s) {
        if constexpr (PR) {
            fast_step_pair<INTEG>(g, bc, tab, s, p, dt, dth, dt6, cc);
        } else if constexpr (PF) {
            fast_step<INTEG, ST>(g, bc, tab, s, p, dt, dth, dt6, cc);
        } else {
            rk_step<INTEG>(f, tab, s, p, dt, dth, dt6);     // rk4.py:84-106 / rk2.py / euler.py
            apply_bc(bc, p);                                // tracker.py:238-239
        }
        if (traj != nullptr && (s + 1) % rec_every == 0) {
            const size_t r = (size_t)((s + 1) / rec_every);
            float* o = traj + (r * (size_t)n + dst) * 3;
            o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        }
    }
    ro.xout[3 * (size_t)i] = p[0];
    ro.xout[3 * (size_t)i + 1] = p[1];
    ro.xout[3 * (size_t)i + 2] = p[2];
}

// Consecutive-pair records from a float32 z-halo field with 4 components:
// thread e takes record (p, i, j, k) and writes its 12 floats with three
// 16-byte stores: the xyz of V_p(k), V_p(k+1), V_p+1(k), V_p+1(k+1), where
// k+1 = nz is the halo plane.
__global__ void __launch_bounds__(256)
k_pair_records(const float4* __restrict__ snaps, unsigned n_rec, unsigned nz, unsigned nxy, unsigned nzs,
               float4* __restrict__ out) {
    const unsigned e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_rec) return;
    const unsigned k = e % nz;
    const unsigned r = e / nz;                          // p * nxy + ij
    const unsigned p = r / nxy, ij = r - p * nxy;
    const size_t snap = (size_t)nxy * nzs;
    const size_t a0 = (size_t)p * snap + (size_t)ij * nzs + k;
    const float4 a = snaps[a0], b = snaps[a0 + 1], c = snaps[a0 + snap], d = snaps[a0 + snap + 1];
    float4* o = out + 3 * (size_t)e;
    o[0] = make_float4(a.x, a.y, a.z, b.x);
    o[1] = make_float4(b.y, b.z, c.x, c.y);
    o[2] = make_float4(c.z, d.x, d.y, d.z);
}

template <int ST, bool ZH>
__global__ void __launch_bounds__(256)
k_sample(GridDev g, const float* __restrict__ x, int n, float t, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Blend b = time_blend(g.times, g.T, t, g.interp, g.extrap);
    float v[3];
    sample<ST, ZH>(g, b, x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2], v);
    out[3 * (size_t)i] = v[0];
    out[3 * (size_t)i + 1] = v[1];
    out[3 * (size_t)i + 2] = v[2];
}

template <bool AXES>
__global__ void __launch_bounds__(256) k_boundary(BcDev bc, float* __restrict__ x, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p[3] = {x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2]};
    if constexpr (AXES) apply_bc_axes(bc, p);
    else apply_bc(bc, p);
    x[3 * (size_t)i] = p[0];
    x[3 * (size_t)i + 1] = p[1];
    x[3 * (size_t)i + 2] = p[2];
}

// ---- continuous inlet with host-drawn replacements -------------------------
__device__ __forceinline__ bool inlet_exited(const BcDev& b, const float* p) {
    return b.ipos ? (p[b.ia] > b.hi[b.ia]) : (p[b.ia] < b.lo[b.ia]);   // boundary.py:1526-1537
}

__global__ void __launch_bounds__(256)
k_inlet_flags(BcDev bc, const float* __restrict__ x, int n, int* __restrict__ flags) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = inlet_exited(bc, x + 3 * (size_t)i) ? 1 : 0;
}

__global__ void k_inlet_count(const int* __restrict__ flags, const int* __restrict__ rank, int n,
                              int* __restrict__ count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) count[0] = rank[n - 1] + flags[n - 1];
}

__global__ void __launch_bounds__(256)
k_inlet_place(BcDev b, float* __restrict__ x, int n, const int* __restrict__ rank, const float* __restrict__ draws,
              long long n_exited) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p[3] = {x[3 * (size_t)i], x[3 * (size_t)i + 1], x[3 * (size_t)i + 2]};
    const int a = b.ia;
    const int o0 = (a == 0) ? 1 : 0;
    const int o1 = (a == 2) ? 1 : 2;
    if (inlet_exited(b, p)) {                           // :1552-1571
        const long long r = rank[i];
        p[a] = b.ipos ? b.lo[a] : b.hi[a];
        p[o0] = draws[r];
        p[o1] = draws[n_exited + r];
    }
    if (b.ipos) { if (p[a] < b.lo[a]) p[a] = b.backstop; }   // :1576-1585
    else        { if (p[a] > b.hi[a]) p[a] = b.backstop; }
#pragma unroll
    for (int k = 0; k < 3; ++k) {                       // :1588-1597
        if (k == a) continue;
        float v = p[k];
        if (v < b.lo[k]) p[k] = 2.0f * b.lo[k] - v;
        else if (v > b.hi[k]) p[k] = 2.0f * b.hi[k] - v;
    }
    x[3 * (size_t)i] = p[0];
    x[3 * (size_t)i + 1] = p[1];
    x[3 * (size_t)i + 2] = p[2];
}

// ---- particle re-ordering by grid cell ------------------------------------
__device__ __forceinline__ unsigned cell_bits(float x, float o, float d, int n, int shift) {
    float f = (x - o) / d;
    int c = (f != f) ? 0 : (int)fminf(fmaxf(floorf(f), 0.0f), (float)(n - 1));
    return (unsigned)c >> shift;
}

// Bit layout of a cell's sort key: key bit p is bit bit[p] of axis axis[p]
// (0 x, 1 y, 2 z) of the cell index >> shift.
struct KeyBits {
    unsigned char axis[30], bit[30];
    int nbits, shift;
};

__global__ void __launch_bounds__(256)
k_cell_keys(GridDev g, const float* __restrict__ xs, int n, KeyBits kb, unsigned* __restrict__ keys,
            int* __restrict__ perm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned cx = cell_bits(xs[3 * (size_t)i], g.o[0], g.d[0], g.nx, kb.shift);
    const unsigned cy = cell_bits(xs[3 * (size_t)i + 1], g.o[1], g.d[1], g.ny, kb.shift);
    const unsigned cz = cell_bits(xs[3 * (size_t)i + 2], g.o[2], g.d[2], g.nz, kb.shift);
    unsigned key = 0;
#pragma unroll
    for (int p = 0; p < 30; ++p) {                      // constant offsets: the layout sits in SGPRs
        if (p < kb.nbits) {
            const unsigned c = kb.axis[p] == 0 ? cx : (kb.axis[p] == 1 ? cy : cz);
            key |= ((c >> kb.bit[p]) & 1u) << p;
        }
    }
    keys[i] = key;
    if (perm) perm[i] = i;
}

// Z-order (Morton) keys over the grid's cells with each axis's own bit count
// (at most 10 bits per axis: coarser cells beyond 1024 nodes), preceded by
// an optional low block of lz z-bits, ly y-bits and lx x-bits.  Bits follow
// z, y, x from the least significant end, so 64 consecutive keys are a
// 4 x 4 x 4 cell cube with no low block, or a 2 x 2 x 16 column with
// (lz, ly, lx) = (4, 1, 1).
static KeyBits make_key_bits(const GridDev& g, int lz, int ly, int lx) {
    KeyBits kb{};
    const int n[3] = {g.nx, g.ny, g.nz};
    const int maxn = std::max(n[0], std::max(n[1], n[2]));
    kb.shift = 0;
    while ((maxn - 1) >> kb.shift >= 1024) ++kb.shift;
    int nb[3], used[3] = {0, 0, 0};
    for (int a = 0; a < 3; ++a) {
        nb[a] = 0;
        while (((n[a] - 1) >> kb.shift) >= (1 << nb[a])) ++nb[a];
    }
    const int low[3] = {lx, ly, lz};
    int p = 0;
    for (int a = 2; a >= 0; --a)
        for (int k = 0; k < low[a] && used[a] < nb[a]; ++k) { kb.axis[p] = (unsigned char)a; kb.bit[p] = (unsigned char)used[a]++; ++p; }
    while (used[0] < nb[0] || used[1] < nb[1] || used[2] < nb[2])
        for (int a = 2; a >= 0; --a)
            if (used[a] < nb[a]) { kb.axis[p] = (unsigned char)a; kb.bit[p] = (unsigned char)used[a]++; ++p; }
    if (p == 0) { kb.axis[0] = 0; kb.bit[0] = 0; p = 1; }   // one cell: all keys 0
    kb.nbits = p;
    return kb;
}

#ifndef TRACK_BUCKET_BITS_MAX
#define TRACK_BUCKET_BITS_MAX 22
#endif
#ifndef TRACK_BUCKET_SORT_MAX_N
#define TRACK_BUCKET_SORT_MAX_N (1 << 22)
#endif

#ifndef JT_KEY_LOW
#define JT_KEY_LOW 3, 2, 2
#endif

// One block of TrackBlock threads per particle range; the grid is padded to a
// multiple of 8 for the kernel's XCD mapping.
template <int INTEG, int ST, bool ZH, bool PF>
static void launch_track(hipStream_t s, const GridDev& g, const BcDev& bc, const Reorder& ro, int n,
                         const Blend* tab, int s0, int s1, float dt, float dth, float dt6, int rec, float* traj) {
    const unsigned B = TrackBlock<ST>::v;
    const unsigned nb = ((unsigned)n + B - 1) / B;
    if constexpr (PF && ST == ST_F32x4) {
        if (g.pair != nullptr) {                        // the fast path over consecutive-pair records
            hipLaunchKernelGGL((k_track<INTEG, ST, ZH, PF, true>), dim3((nb + 7) / 8 * 8), dim3(B), 0, s, g, bc, ro, n,
                               tab, s0, s1, dt, dth, dt6, rec, traj);
            return;
        }
    }
    hipLaunchKernelGGL((k_track<INTEG, ST, ZH, PF>), dim3((nb + 7) / 8 * 8), dim3(B), 0, s, g, bc, ro, n, tab, s0, s1,
                       dt, dth, dt6, rec, traj);
}

template <int INTEG, bool ZH, bool PF>
static void launch_track_st(int st, hipStream_t s, const GridDev& g,
                            const BcDev& bc, const Reorder& ro, int n, const Blend* tab, int s0, int s1,
                            float dt, float dth, float dt6, int rec, float* traj) {
    switch (st) {
    case ST_F32x3: launch_track<INTEG, ST_F32x3, ZH, PF>(s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj); break;
    case ST_F32x4: launch_track<INTEG, ST_F32x4, ZH, PF>(s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj); break;
    case ST_BF16x4: launch_track<INTEG, ST_BF16x4, ZH, PF>(s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj); break;
    default: launch_track<INTEG, ST_F16x4, ZH, PF>(s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj); break;
    }
}

// The z-halo layout with the fast paths (periodic sampler with exact
// reciprocal spacing, per-axis boundary kinds), the z-halo layout, or the
// node-per-load layout.
template <int INTEG>
static void launch_track_any(int st, hipStream_t s, const GridDev& g, const BcDev& bc, const Reorder& ro, int n,
                             const Blend* tab, int s0, int s1, float dt, float dth, float dt6, int rec,
                             float* traj) {
    if (g.nzs == g.nz) launch_track_st<INTEG, false, false>(st, s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj);
    else if (periodic_fast(g, st == ST_F32x3 ? 12 : (st == ST_F32x4 ? 16 : 8)) && bc_axes_only(bc))
        launch_track_st<INTEG, true, true>(st, s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj);
    else launch_track_st<INTEG, true, false>(st, s, g, bc, ro, n, tab, s0, s1, dt, dth, dt6, rec, traj);
}

template <bool ZH>
static void launch_sample(int st, dim3 grid, dim3 block, hipStream_t s, const GridDev& g, const float* x, int n,
                          float t, float* out) {
    switch (st) {
    case ST_F32x3: hipLaunchKernelGGL((k_sample<ST_F32x3, ZH>), grid, block, 0, s, g, x, n, t, out); break;
    case ST_F32x4: hipLaunchKernelGGL((k_sample<ST_F32x4, ZH>), grid, block, 0, s, g, x, n, t, out); break;
    case ST_BF16x4: hipLaunchKernelGGL((k_sample<ST_BF16x4, ZH>), grid, block, 0, s, g, x, n, t, out); break;
    default: hipLaunchKernelGGL((k_sample<ST_F16x4, ZH>), grid, block, 0, s, g, x, n, t, out); break;
    }
}


}  // namespace jt

using namespace jt;

extern "C" int jt_grid_sample(const jt_grid_field* field, const float* x, int64_t n, float t,
                              float* out, void* stream) {
    GridDev g;
    int st;
    int rc = make_grid_dev(field, g, st);
    if (rc) return rc;
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    if (n == 0) return JT_OK;
    JT_REQUIRE(x != nullptr && out != nullptr, "NULL x/out");
    hipStream_t s = (hipStream_t)stream;
    dim3 block(256), grid((unsigned)((n + 255) / 256));
    if (g.nzs != g.nz) launch_sample<true>(st, grid, block, s, g, x, (int)n, t, out);
    else launch_sample<false>(st, grid, block, s, g, x, (int)n, t, out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_grid_pair_records(const jt_grid_field* field, float* out, void* stream) {
    JT_REQUIRE(field != nullptr && out != nullptr, "NULL pointer");
    jt_grid_field f = *field;
    f.pair_data = nullptr;
    GridDev g;
    int st;
    int rc = make_grid_dev(&f, g, st);
    if (rc) return rc;
    JT_REQUIRE(st == ST_F32x4 && g.nzs == g.nz + 1, "pair records need a float32 field with 4 components and a z halo");
    JT_REQUIRE(g.T >= 2, "pair records need at least two snapshots");
    const long long n_rec = (long long)(g.T - 1) * g.nx * g.ny * g.nz;
    JT_REQUIRE(n_rec < (1LL << 31), "too many pair records (%lld)", n_rec);
    hipLaunchKernelGGL(k_pair_records, dim3((unsigned)((n_rec + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       (const float4*)g.data, (unsigned)n_rec, (unsigned)g.nz, (unsigned)(g.nx * g.ny),
                       (unsigned)g.nzs, (float4*)out);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" size_t jt_inlet_exits_workspace_size(int64_t n) {
    Arena a(nullptr, 0);
    const int nn = (int)(n > 0 ? n : 1);
    a.take<int>(nn);
    size_t temp = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (const int*)nullptr, (int*)nullptr, nn);
    a.take<char>(temp);
    return a.used + 256;
}

extern "C" int jt_inlet_exits(const jt_boundary* bc, const float* x, int64_t n, int32_t* rank, int32_t* n_exited,
                              void* ws, size_t ws_bytes, void* stream) {
    BcDev d;
    JT_REQUIRE(bc != nullptr, "boundary descriptor is NULL");
    int rc = make_bc_dev(bc, d);
    if (rc) return rc;
    JT_REQUIRE(d.inlet == JT_INLET_HOST, "jt_inlet_exits needs a JT_INLET_HOST boundary");
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(n_exited != nullptr, "NULL n_exited");
    hipStream_t s = (hipStream_t)stream;
    if (n == 0) { JT_HIP(hipMemsetAsync(n_exited, 0, sizeof(int32_t), s)); return JT_OK; }
    JT_REQUIRE(x && rank, "NULL x/rank");
    Arena arena(ws, ws_bytes);
    int* flags = arena.take<int>(n);
    size_t temp = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, temp, (const int*)nullptr, (int*)nullptr, (int)n);
    char* tmp = arena.take<char>(temp);
    if (!flags || !tmp || !arena.ok())
        return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu", jt_inlet_exits_workspace_size(n), ws_bytes);
    const dim3 blk(256), grd((unsigned)((n + 255) / 256));
    hipLaunchKernelGGL(k_inlet_flags, grd, blk, 0, s, d, x, (int)n, flags);
    JT_HIP(hipGetLastError());
    JT_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, temp, flags, rank, (int)n, s));
    hipLaunchKernelGGL(k_inlet_count, dim3(1), dim3(1), 0, s, flags, rank, (int)n, n_exited);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_inlet_place(const jt_boundary* bc, float* x, int64_t n, const int32_t* rank, const float* draws,
                              int64_t n_exited, void* stream) {
    BcDev d;
    JT_REQUIRE(bc != nullptr, "boundary descriptor is NULL");
    int rc = make_bc_dev(bc, d);
    if (rc) return rc;
    JT_REQUIRE(d.inlet == JT_INLET_HOST, "jt_inlet_place needs a JT_INLET_HOST boundary");
    JT_REQUIRE(n >= 0 && n < (1LL << 31) && n_exited >= 0 && n_exited <= n, "n / n_exited out of range");
    if (n == 0) return JT_OK;
    JT_REQUIRE(x && rank && (n_exited == 0 || draws), "NULL x/rank/draws");
    hipLaunchKernelGGL(k_inlet_place, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d, x,
                       (int)n, rank, draws, (long long)n_exited);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

extern "C" int jt_apply_boundary(const jt_boundary* bc, float* x, int64_t n, void* stream) {
    BcDev d;
    JT_REQUIRE(bc != nullptr, "boundary descriptor is NULL");
    int rc = make_bc_dev(bc, d);
    if (rc) return rc;
    if (d.inlet == JT_INLET_HOST)
        return fail(JT_EUNSUPPORTED, "a host-drawn inlet boundary is applied with jt_inlet_exits/jt_inlet_place");
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    if (n == 0) return JT_OK;
    JT_REQUIRE(x != nullptr, "NULL x");
    // the same fused epilogue the tracking kernels run on this boundary
    if (bc_axes_only(d))
        hipLaunchKernelGGL(k_boundary<true>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, d, x, (int)n);
    else
        hipLaunchKernelGGL(k_boundary<false>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, d, x, (int)n);
    JT_HIP(hipGetLastError());
    return JT_OK;
}

static int default_steps_per_launch(int64_t n) {
    // Keep one launch to a fraction of a second at the largest sizes.
    if (n >= (1 << 22)) return 50;
    return 1000;
}

extern "C" size_t jt_track_grid_workspace_size(int64_t n, int32_t n_steps, const jt_track_opts* opts) {
    Arena a(nullptr, 0);
    a.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    if (opts && opts->sort_every > 0 && n > 0) {
        a.take<unsigned>(n); a.take<unsigned>(n);      // keys, sorted keys (radix sort)
        a.take<int>(n); a.take<int>(n);                // identity / sorted order
        a.take<int>(n); a.take<int>(n);                // particle ids (double buffer)
        a.take<float>((size_t)3 * n); a.take<float>((size_t)3 * n);   // sorted state (double buffer)
        const int nb = 1 << TRACK_BUCKET_BITS_MAX;
        a.take<int>(nb); a.take<int>(nb);              // bucket counts, offsets (counting sort)
        const size_t rt = sort_temp_bytes((int)n), bt = bucket_scan_bytes(nb);
        a.take<char>(rt > bt ? rt : bt);
    }
    return a.used + 256;
}

// The cell-sort buffers of one tracking call (all NULL: no sorting).
struct SortBufs {
    unsigned *keys = nullptr, *skeys = nullptr;
    int *perm = nullptr, *sperm = nullptr, *ids = nullptr, *ids2 = nullptr, *bcount = nullptr, *boffs = nullptr;
    float *xa = nullptr, *xb = nullptr;
    char* tmp = nullptr;
    size_t temp = 0;
};

// Checks shared by the tracking entry points; fills g/st/d.
static int track_grid_args(const jt_grid_field* field, const jt_boundary* bc, int32_t integrator, const float* x,
                           int64_t n, const float* step_times, int32_t n_steps, int32_t record_every,
                           const jt_track_opts* opts, GridDev& g, int& st, BcDev& d) {
    int rc = make_grid_dev(field, g, st);
    if (rc) return rc;
    rc = make_bc_dev(bc, d);
    if (rc) return rc;
    JT_REQUIRE(integrator >= JT_EULER && integrator <= JT_RK4, "bad integrator %d", integrator);
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(n_steps >= 0, "n_steps must be >= 0");
    JT_REQUIRE(record_every >= 1, "record_every must be >= 1");
    if (d.inlet == JT_INLET_HOST)
        return fail(JT_EUNSUPPORTED, "a host-drawn inlet boundary is not fused into tracking; step with an open "
                                     "boundary and apply it with jt_inlet_exits/jt_inlet_place");
    JT_REQUIRE(n == 0 || x != nullptr, "NULL x");
    JT_REQUIRE(n_steps == 0 || step_times != nullptr, "NULL step_times");
    JT_REQUIRE(opts == nullptr || (opts->sort_every >= 0 && opts->steps_per_launch >= 0), "bad opts");
    return JT_OK;
}

// The tracking run itself: tab holds room for 3 * n_steps blends; sb holds
// the cell-sort buffers when opts->sort_every > 0.
static int track_grid_run(const GridDev& g, int st, const BcDev& d, int32_t integrator, float* x, int64_t n,
                          const float* step_times, float dt, int32_t n_steps, int32_t record_every, float* traj,
                          const jt_track_opts* opts, Blend* tab, const SortBufs& sb, hipStream_t s) {
    if (traj != nullptr && n > 0)
        JT_HIP(hipMemcpyAsync(traj, x, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (n == 0 || n_steps == 0) return JT_OK;
    const int sort_every = opts ? opts->sort_every : 0;
    // a counting sort by bucket below TRACK_BUCKET_SORT_MAX_N particles (2^20:
    // 12.5 against 14.0 ms for 500 steps), a radix sort above (2^24: 0.57
    // against 1.1 ms per sort, and better ordered waves)
    const bool bucketed = n < TRACK_BUCKET_SORT_MAX_N;
    hipLaunchKernelGGL(k_blend_table, dim3((n_steps + 127) / 128), dim3(128), 0, s, g, step_times,
                       n_steps, dt, integrator, tab);
    JT_HIP(hipGetLastError());

    int spl = (opts && opts->steps_per_launch > 0) ? opts->steps_per_launch : default_steps_per_launch(n);
    const float dth = 0.5f * dt;       // rk4.py:90 (rk2.py:169 is dt*0.5, the same value)
    const float dt6 = dt / 6.0f;       // rk4.py:44
    const unsigned nb = (unsigned)((n + 255) / 256);
    dim3 eb(256), eg(nb);
    const KeyBits kb = make_key_bits(g, JT_KEY_LOW);
    // buckets of 64 cells (the low 6 key bits: a wave's 2 x 4 x 8-cell block;
    // the order within a wave does not matter to its gathers), at most
    // 2^TRACK_BUCKET_BITS_MAX of them
    int bshift = kb.nbits > 6 ? 6 : kb.nbits;
    if (kb.nbits - bshift > TRACK_BUCKET_BITS_MAX) bshift = kb.nbits - TRACK_BUCKET_BITS_MAX;
    float* cur = x;
    const int* cur_ids = nullptr;
    const int n_ev = (opts && opts->launch_events) ? opts->n_launch_events : 0;
    int launch = 0;
    for (int s0 = 0; s0 < n_steps;) {
        int s1 = s0 + spl < n_steps ? s0 + spl : n_steps;
        Reorder ro{cur, cur, cur_ids, nullptr, nullptr};
        if (sort_every > 0) {
            int next_sort = (s0 / sort_every + 1) * sort_every;
            if (s1 > next_sort) s1 = next_sort;
            if (s0 % sort_every == 0) {
                hipLaunchKernelGGL(k_cell_keys, eg, eb, 0, s, g, cur, (int)n, kb, sb.keys,
                                   bucketed ? nullptr : sb.perm);
                if (bucketed)
                    JT_HIP(bucket_sort(s, sb.keys, (int)n, bshift, 1 << (kb.nbits - bshift), sb.bcount, sb.boffs,
                                       sb.tmp, sb.temp, sb.sperm));
                else {
                    size_t temp = sb.temp;
                    JT_HIP(hipcub::DeviceRadixSort::SortPairs(sb.tmp, temp, sb.keys, sb.skeys, sb.perm, sb.sperm,
                                                              (int)n, 0, kb.nbits, s));
                }
                float* nx = (cur == sb.xa) ? sb.xb : sb.xa;
                int* nid = (cur_ids == sb.ids) ? sb.ids2 : sb.ids;
                ro = Reorder{cur, nx, cur_ids, nid, sb.sperm};   // the launch below applies the permutation
                cur = nx;
                cur_ids = nid;
            }
        }
        if (2 * launch + 1 < n_ev) JT_HIP(hipEventRecord((hipEvent_t)opts->launch_events[2 * launch], s));
        switch (integrator) {
        case JT_RK4:
            launch_track_any<JT_RK4>(st, s, g, d, ro, (int)n, tab, s0, s1, dt, dth, dt6, record_every, traj);
            break;
        case JT_RK2:
            launch_track_any<JT_RK2>(st, s, g, d, ro, (int)n, tab, s0, s1, dt, dth, dt6, record_every, traj);
            break;
        default:
            launch_track_any<JT_EULER>(st, s, g, d, ro, (int)n, tab, s0, s1, dt, dth, dt6, record_every, traj);
            break;
        }
        JT_HIP(hipGetLastError());
        if (2 * launch + 1 < n_ev) JT_HIP(hipEventRecord((hipEvent_t)opts->launch_events[2 * launch + 1], s));
        ++launch;
        s0 = s1;
    }
    if (cur != x) {
        hipLaunchKernelGGL(k_unpermute, eg, eb, 0, s, cur, cur_ids, (int)n, x);
        JT_HIP(hipGetLastError());
    }
    return JT_OK;
}

extern "C" int jt_track_grid(const jt_grid_field* field, const jt_boundary* bc, int32_t integrator,
                             float* x, int64_t n, const float* step_times, float dt, int32_t n_steps,
                             int32_t record_every, float* traj, const jt_track_opts* opts,
                             void* ws, size_t ws_bytes, void* stream) {
    GridDev g;
    int st;
    BcDev d;
    int rc = track_grid_args(field, bc, integrator, x, n, step_times, n_steps, record_every, opts, g, st, d);
    if (rc) return rc;
    if (n == 0 || n_steps == 0) return track_grid_run(g, st, d, integrator, x, n, step_times, dt, n_steps,
                                                      record_every, traj, opts, nullptr, SortBufs{},
                                                      (hipStream_t)stream);
    const int sort_every = opts ? opts->sort_every : 0;
    const size_t need = jt_track_grid_workspace_size(n, n_steps, opts);
    Arena arena(ws, ws_bytes);
    Blend* tab = arena.take<Blend>((size_t)3 * n_steps);
    SortBufs sb;
    if (sort_every > 0) {
        const bool bucketed = n < TRACK_BUCKET_SORT_MAX_N;
        sb.keys = arena.take<unsigned>(n); sb.skeys = arena.take<unsigned>(n);
        sb.perm = arena.take<int>(n); sb.sperm = arena.take<int>(n);
        sb.ids = arena.take<int>(n); sb.ids2 = arena.take<int>(n);
        sb.xa = arena.take<float>((size_t)3 * n); sb.xb = arena.take<float>((size_t)3 * n);
        sb.bcount = arena.take<int>(1 << TRACK_BUCKET_BITS_MAX); sb.boffs = arena.take<int>(1 << TRACK_BUCKET_BITS_MAX);
        sb.temp = bucketed ? bucket_scan_bytes(1 << TRACK_BUCKET_BITS_MAX) : sort_temp_bytes((int)n);
        sb.tmp = arena.take<char>(sb.temp);
    }
    if (tab == nullptr || !arena.ok() || (sort_every > 0 && sb.tmp == nullptr))
        return fail(JT_ENOMEM, "workspace too small: need %zu bytes, got %zu", need, ws_bytes);
    return track_grid_run(g, st, d, integrator, x, n, step_times, dt, n_steps, record_every, traj, opts, tab, sb,
                          (hipStream_t)stream);
}

// ---- allocation-free tracking -------------------------------------------------
// jt_track_grid_noalloc's only scratch, the stage-time blend table, lives in
// this file-scope array: JT_NOALLOC_MAX_STEPS steps x 3 blends x 16 B =
// 3 MiB, held on every GPU the library runs on until the process ends.  Host
// code never passes the array's name to a kernel (on the host it is a
// stand-in, not a device address): its device address is looked up once per
// device with hipGetSymbolAddress and kept in noalloc_ptr[].
#define JT_NOALLOC_MAX_STEPS (1 << 16)
#define JT_NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) Blend scratch[3 * JT_NOALLOC_MAX_STEPS];

static std::mutex noalloc_mu;
static void* noalloc_ptr[JT_NOALLOC_MAX_DEVICES];

// The current device's address of `scratch` (looked up on its first use).
static int noalloc_scratch(Blend** out) {
    int dev = 0;
    JT_HIP(hipGetDevice(&dev));
    JT_REQUIRE(dev >= 0 && dev < JT_NOALLOC_MAX_DEVICES, "device %d beyond the supported %d", dev,
               JT_NOALLOC_MAX_DEVICES);
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_ptr[dev] == nullptr) {
        void* p = nullptr;
        JT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        JT_REQUIRE(p != nullptr, "hipGetSymbolAddress returned NULL for the tracking scratch");
        noalloc_ptr[dev] = p;
    }
    *out = (Blend*)noalloc_ptr[dev];
    return JT_OK;
}

extern "C" int jt_track_grid_noalloc(const jt_grid_field* field, const jt_boundary* bc, int32_t integrator,
                                     float* x, int64_t n, const float* step_times, float dt, int32_t n_steps,
                                     int32_t record_every, float* traj, const jt_track_opts* opts, void* stream) {
    GridDev g;
    int st;
    BcDev d;
    int rc = track_grid_args(field, bc, integrator, x, n, step_times, n_steps, record_every, opts, g, st, d);
    if (rc) return rc;
    if (n_steps > JT_NOALLOC_MAX_STEPS)
        return fail(JT_EINVAL, "jt_track_grid_noalloc takes at most %d steps per call (got %d): split the run "
                               "into calls of at most that many steps", JT_NOALLOC_MAX_STEPS, n_steps);
    if (opts != nullptr && opts->sort_every > 0)
        return fail(JT_EUNSUPPORTED, "jt_track_grid_noalloc does not sort particles by cell (the sort needs "
                                     "O(n) buffers): pass sort_every = 0, or use jt_track_grid with a workspace");
    Blend* tab = nullptr;
    rc = noalloc_scratch(&tab);
    if (rc) return rc;
    return track_grid_run(g, st, d, integrator, x, n, step_times, dt, n_steps, record_every, traj, opts, tab,
                          SortBufs{}, (hipStream_t)stream);
}
