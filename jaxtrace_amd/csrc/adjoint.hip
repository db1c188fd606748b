// Adjoint (reverse-mode) RK4 tracking through a time-blended structured grid.
//
// The reference has no differentiation; this is the build-only gradient of
// the forward path of track_grid.hip: given dL/dx_T it returns dL/dx_0 and
// dL/dU for the snapshots U, with positions checkpointed every K steps by the
// forward run (its trajectory recorded every K steps) and recomputed in
// between.  Per step the forward stages are recomputed with the forward
// kernel's exact arithmetic, then reversed:
//   x' = BC(x + dt/6 (k1 + 2k2 + 2k3 + k4)),  k_i = u(x_i, t_i)
// using the trilinear Jacobian (corner weights' derivatives times the blended
// corner values) and adding  coef_snap * W_corner * abar  into dL/dU.
//
// The dL/dU scatter is reduced in three levels, because atomics are the cost
// that dominates (measured on the whole chip with random addresses: about
// 1.8 T lane-ops/s for LDS ds_add_f64 but only 0.2 T/s for ds_add_f32 and
// 5.7 T/s for ds_add_u32, and about 21 G/s for scattered global_atomic_add_f32):
//   1. registers: a thread sums its contributions for one cell (the 8 corner
//      nodes) and one snapshot pair while its stage points stay in that cell,
//      and flushes the 48 sums only when the cell changes;
//   2. LDS: a block of Hilbert-sorted particles flushes, in float64, into a
//      box of nodes (two snapshot slots) around its particles' start cells;
//   3. HBM: the box is added to dL/dU with coalesced float atomics at the end
//      of the launch; flushes outside the box go straight to HBM.
#include "grid_sampler.h"
#include "cell_sort.h"
#include <algorithm>
#include <vector>

namespace jt {

// d(blended value)/d(V[i0]) and d/d(V[i1]) for each blend mode.
__device__ __forceinline__ void blend_coefs(const Blend& b, float& c0, float& c1) {
    switch (b.mode) {
    case BL_INTERP: c0 = 1.0f - b.a; c1 = b.a; break;
    case BL_COPY:   c0 = 1.0f; c1 = 0.0f; break;
    case BL_LEFT:   c0 = 1.0f - b.a; c1 = b.a; break;
    case BL_RIGHT:  c0 = -b.a; c1 = 1.0f + b.a; break;
    default:        c0 = 0.0f; c1 = 0.0f; break;
    }
}

// d BC(x)/dx per axis (periodic 1, reflective +-1, clamp 1 inside / 0 outside).
__device__ __forceinline__ float bc_deriv(int kind, float x, float lo, float hi) {
    if (kind == JT_BC_REFLECTIVE) {
        float w = hi - lo;
        w = w < 0.0f ? 0.0f : w;
        float y = jmodf(x - lo, 2.0f * (w > 0.0f ? w : 1.0f));
        return (y <= w) ? 1.0f : -1.0f;
    }
    if (kind == JT_BC_CLAMP) return (x >= lo && x <= hi) ? 1.0f : 0.0f;
    return 1.0f;
}

// ---- LDS box -----------------------------------------------------------------
#ifndef TILE_NODES
#define TILE_NODES 3320            // 2 slots x 3 float64 per node + a spare: the CU's 160 KiB, one block per CU
#endif
#ifndef ADJ_BLOCK
#define ADJ_BLOCK 512              // 8 waves: two per SIMD, as the registers allow
#endif
#ifndef ADJ_RESORT
#define ADJ_RESORT 16              // longest sub-segment: 16 measured fastest with the middle-step box (8: 5 % slower, 24: 3 %)
#endif
#ifndef ADJ_MARGIN
#define ADJ_MARGIN 1               // the widest box margin tried, in cells each way
#endif
#ifndef ADJ_REC_WAVES
#define ADJ_REC_WAVES 4
#endif
#ifndef ADJ_MINW
#define ADJ_MINW 1
#endif

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

struct Tile {
    int m[3];           // cell the offsets are taken from
    int b[3], e[3];     // box origin (wrapped node offsets from m) and extents
    int smin;           // snapshot of slot 0 (slot 1 is smin + 1)
    int vol;            // e0*e1*e2 (0: no box)
};

// v in (-n, n) wrapped to [-(n/2), n - n/2): a node offset from the box's
// origin cell that is the same node whichever way round a periodic axis it
// is counted.
__device__ __forceinline__ int wrap_off(int v, int n) {
    const int h = n >> 1;
    v = v < -h ? v + n : v;
    return v >= n - h ? v - n : v;
}

struct Stencil {
    int I[2], J[2], K[2];
    float wx[2], wy[2], wz[2];   // corner weight of (ci, cj, ck) = wx[ci] * wy[cj] * wz[ck]
};

// gather_corners<ST_F32x4, true, true> of an 'interp' blend between snapshots
// (i0, i0 + 1) from the consecutive-pair records: a cell row of both
// snapshots is 12 consecutive floats (three 16-byte loads), the blend is
// blend_n's 'interp' arithmetic.
__device__ __forceinline__ void gather_pair(const GridDev& g, const Blend& bl, float px, float py, float pz,
                                            Corners& C) {
    C.valid = true;
    int i0, i1, j0, j1, k0, k1;
    axis_coord_pf(px, g.o[0], g.inv[0], g.nx, (float)g.nx, i0, i1, C.wx);
    axis_coord_pf(py, g.o[1], g.inv[1], g.ny, (float)g.ny, j0, j1, C.wy);
    axis_coord_pf(pz, g.o[2], g.inv[2], g.nz, (float)g.nz, k0, k1, C.wz);
    const char* pp = (const char*)(g.pair + (size_t)bl.i0 * (size_t)g.pair_stride);
    const unsigned rs = (unsigned)g.nz * 48u;
    const unsigned r0 = (unsigned)__umul24((unsigned)i0, (unsigned)g.ny);
    const unsigned r1 = (unsigned)__umul24((unsigned)i1, (unsigned)g.ny);
    const unsigned kb = (unsigned)k0 * 48u;
    const unsigned row[4] = {(unsigned)__umul24(r0 + (unsigned)j0, rs) + kb, (unsigned)__umul24(r0 + (unsigned)j1, rs) + kb,
                             (unsigned)__umul24(r1 + (unsigned)j0, rs) + kb, (unsigned)__umul24(r1 + (unsigned)j1, rs) + kb};
    float4 f[4][3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const float4* qq = (const float4*)(pp + row[r]);
        f[r][0] = qq[0]; f[r][1] = qq[1]; f[r][2] = qq[2];
    }
    float A[8][3], B[8][3];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        A[2 * r][0] = f[r][0].x; A[2 * r][1] = f[r][0].y; A[2 * r][2] = f[r][0].z;
        A[2 * r + 1][0] = f[r][0].w; A[2 * r + 1][1] = f[r][1].x; A[2 * r + 1][2] = f[r][1].y;
        B[2 * r][0] = f[r][1].z; B[2 * r][1] = f[r][1].w; B[2 * r][2] = f[r][2].x;
        B[2 * r + 1][0] = f[r][2].y; B[2 * r + 1][1] = f[r][2].z; B[2 * r + 1][2] = f[r][2].w;
    }
    const float al = bl.a, oma = 1.0f - bl.a;
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
        for (int c = 0; c < 3; ++c) C.v[q][c] = oma * A[q][c] + al * B[q][c];
}

// u(q) with the forward sampler's exact arithmetic, plus its Jacobian
// J[c*3 + ax] = du_c/dx_ax from the same eight blended corner values: the
// trilinear weights' derivatives are +-1/d per axis (0 where a non-periodic
// axis is clipped).  depends = false where u is the constant 0 of the 'zero'
// bounds mode.
// PF: the periodic fast path's index math.  PR: a float32 field with
// consecutive-pair records: an 'interp' blend between snapshots (s, s + 1)
// gathers each cell row of both snapshots as three 16-byte loads from the
// records (gather_pair) instead of four; the values are the same.
template <int ST, bool ZH, bool PF = false, bool PR = false>
__device__ __forceinline__ void sample_jac(const GridDev& g, const Blend& bl, const float q[3], float u[3],
                                           float J[9], bool& depends) {
    Corners C;
    if (PR && bl.mode == BL_INTERP && bl.i1 == bl.i0 + 1) gather_pair(g, bl, q[0], q[1], q[2], C);
    else gather_corners<ST, ZH, PF>(g, bl, q[0], q[1], q[2], C);
    lerp_corners(g, C, u);
    depends = !(!C.valid && g.bmode == JT_BOUNDS_ZERO);
    float gd[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float raw = (q[ax] - g.o[ax]) / g.d[ax];
        gd[ax] = 1.0f / g.d[ax];
        if (g.bmode != JT_BOUNDS_PERIODIC && !(raw >= 0.0f && raw <= g.cmax[ax])) gd[ax] = 0.0f;
        if (!depends) gd[ax] = 0.0f;
    }
    const float wx = C.wx, wy = C.wy, wz = C.wz;
    const float ax_ = 1.0f - wx, ay = 1.0f - wy, az = 1.0f - wz;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float Z[2][2], dZ[2][2];
#pragma unroll
        for (int ci = 0; ci < 2; ++ci)
#pragma unroll
            for (int cj = 0; cj < 2; ++cj) {
                const float z0 = C.v[ci * 4 + cj * 2][c], z1 = C.v[ci * 4 + cj * 2 + 1][c];
                Z[ci][cj] = fmaf(z1, wz, z0 * az);
                dZ[ci][cj] = z1 - z0;
            }
        float Y[2], dYy[2], dYz[2];
#pragma unroll
        for (int ci = 0; ci < 2; ++ci) {
            Y[ci] = fmaf(Z[ci][1], wy, Z[ci][0] * ay);
            dYy[ci] = Z[ci][1] - Z[ci][0];
            dYz[ci] = fmaf(dZ[ci][1], wy, dZ[ci][0] * ay);
        }
        J[c * 3 + 0] = (Y[1] - Y[0]) * gd[0];
        J[c * 3 + 1] = fmaf(dYy[1], wx, dYy[0] * ax_) * gd[1];
        J[c * 3 + 2] = fmaf(dYz[1], wx, dYz[0] * ax_) * gd[2];
    }
}

// l = J^T a
__device__ __forceinline__ void jac_t(const float J[9], const float a[3], float l[3]) {
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) l[ax] = fmaf(J[6 + ax], a[2], fmaf(J[3 + ax], a[1], J[ax] * a[0]));
}

// Level-1 accumulator: sums of c_h * W_q * a over the stage points of one
// run (same stencil, same snapshot pair) of one thread.
struct RunAcc {
    int I[2], J[2], K[2];
    int snap[2];
    bool live;
    float v[2][8][3];
};

__device__ __forceinline__ void run_reset(RunAcc& r, const Stencil& st, const Blend& bl) {
    r.I[0] = st.I[0]; r.I[1] = st.I[1];
    r.J[0] = st.J[0]; r.J[1] = st.J[1];
    r.K[0] = st.K[0]; r.K[1] = st.K[1];
    r.snap[0] = bl.i0; r.snap[1] = bl.i1;
    r.live = true;
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int q = 0; q < 8; ++q) r.v[h][q][0] = r.v[h][q][1] = r.v[h][q][2] = 0.0f;
}

__device__ __forceinline__ void run_add(RunAcc& r, const Stencil& st, const Blend& bl, const float a[3]) {
    float c[2];
    blend_coefs(bl, c[0], c[1]);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const float W = st.wx[q >> 2] * st.wy[(q >> 1) & 1] * st.wz[q & 1];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float f = c[h] * W;
            r.v[h][q][0] = fmaf(f, a[0], r.v[h][q][0]);
            r.v[h][q][1] = fmaf(f, a[1], r.v[h][q][1]);
            r.v[h][q][2] = fmaf(f, a[2], r.v[h][q][2]);
        }
    }
}

// Add a run's sums into the LDS box (ds_add_f64) where its node and snapshot
// are covered, else into dL/dU (global_atomic_add_f32).  The LDS adds are
// unconditional (a missing corner adds into a spare node past the box, never
// read back) so the wave takes no branch per corner; only lanes with a miss
// run the second, global loop.  The two atomics carry different scopes so the
// compiler keeps them apart instead of merging them into one flat atomic.
__device__ __forceinline__ void run_flush(const GridDev& g, const RunAcc& r, const Tile& tl, double* tile,
                                          float* __restrict__ gU) {
    const int tv = tl.vol * 3;
    const int spare = 2 * TILE_NODES * 3;
    unsigned miss = 0;
    const int oi[2] = {wrap_off(r.I[0] - tl.m[0], g.nx) - tl.b[0], wrap_off(r.I[1] - tl.m[0], g.nx) - tl.b[0]};
    const int oj[2] = {wrap_off(r.J[0] - tl.m[1], g.ny) - tl.b[1], wrap_off(r.J[1] - tl.m[1], g.ny) - tl.b[1]};
    const int ok_[2] = {wrap_off(r.K[0] - tl.m[2], g.nz) - tl.b[2], wrap_off(r.K[1] - tl.m[2], g.nz) - tl.b[2]};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int li = oi[q >> 2], lj = oj[(q >> 1) & 1], lk = ok_[q & 1];
        const bool in = (unsigned)li < (unsigned)tl.e[0] && (unsigned)lj < (unsigned)tl.e[1] &&
                        (unsigned)lk < (unsigned)tl.e[2];
        const int loc = ((li * tl.e[1] + lj) * tl.e[2] + lk) * 3;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int slot = r.snap[h] - tl.smin;
            const bool ok = in && (slot == 0 || slot == 1);
            const int e = ok ? slot * tv + loc : spare;
            const float* v = r.v[h][q];
            __hip_atomic_fetch_add(&tile[e], (double)v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(&tile[e + 1], (double)v[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            __hip_atomic_fetch_add(&tile[e + 2], (double)v[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (!ok && (v[0] != 0.0f || v[1] != 0.0f || v[2] != 0.0f)) miss |= 1u << (q * 2 + h);
        }
    }
    if (miss) {
        const size_t gsnap = (size_t)g.nx * g.ny * g.nz * 3;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int I = r.I[q >> 2], J = r.J[(q >> 1) & 1], K = r.K[q & 1];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (!((miss >> (q * 2 + h)) & 1u)) continue;
                const float* v = r.v[h][q];
                float* gp = gU + (size_t)r.snap[h] * gsnap + ((size_t)(I * g.ny + J) * g.nz + K) * 3;
                __hip_atomic_fetch_add(gp, v[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gp + 1, v[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(gp + 2, v[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Forward recompute of one checkpoint segment [sa, sb): thread t takes
// particle perm[t] (cell-sorted) and stores the state before each step at
// seg[(s - sa) * n + t] (thread order: coalesced stores).
// PF: the tracking kernel's fast path (periodic sampler with exact reciprocal
// spacing, z-halo layout, per-axis boundary kinds) with its per-lane corner
// reuse (PR: over the consecutive-pair records): the same values as the
// forward run.
template <int ST, bool ZH, bool PF = false, bool PR = false>
__global__ void __launch_bounds__(256, ADJ_REC_WAVES)
k_adj_recompute(GridDev g, BcDev bc, const float* __restrict__ x0s, int n, const int* __restrict__ perm,
                const Blend* __restrict__ tab, int sa, int sb, float dt, float dth, float dt6,
                float* __restrict__ seg) {
    const int per = gridDim.x >> 3;                    // gridDim.x is a multiple of 8: consecutive
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);   // tiles share an XCD's L2
    const int t = lb * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int pid = perm[t];
    float x[3] = {x0s[3 * (size_t)pid], x0s[3 * (size_t)pid + 1], x0s[3 * (size_t)pid + 2]};
    auto f = [&](const Blend& b, const float q[3], float out[3]) { sample<ST, ZH>(g, b, q[0], q[1], q[2], out); };
    CornerCache<ST> cc;
    cache_reset(cc);
#pragma unroll 1
    for (int s = sa; s < sb; ++s) {
        float* o = seg + ((size_t)(s - sa) * n + t) * 3;
        o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
        if constexpr (PF) {
            if constexpr (PR) fast_step_pair<JT_RK4>(g, bc, tab, s, x, dt, dth, dt6, cc);
            else fast_step<JT_RK4, ST>(g, bc, tab, s, x, dt, dth, dt6, cc);
        } else {
            rk_step<JT_RK4>(f, tab, s, x, dt, dth, dt6);
            apply_bc(bc, x);
        }
    }
}

// Reverse steps s_hi-1 .. s_lo of the segment starting at sa.  Thread t takes
// segment slot slot = order[t] (slots sorted by their cell at the middle step
// s_m = (s_lo + s_hi) / 2, so a block's particles stay in a small region
// around their middle-step cells while the kernel runs), particle
// pid = seg_pid[slot]; its adjoint lives in lam_io[pid] between launches.
template <int ST, bool ZH, bool PF = false, bool PR = false>
__global__ void __launch_bounds__(ADJ_BLOCK, ADJ_MINW)
k_adj_reverse(GridDev g, BcDev bc, int n, const int* __restrict__ order, const int* __restrict__ seg_pid,
              const Blend* __restrict__ tab, int sa, int s_lo, int s_hi, float dt, float dth, float dt6,
              const float* __restrict__ seg, float* __restrict__ lam_io, float* __restrict__ gU,
              float* __restrict__ gx0) {
    __shared__ double tile[2 * TILE_NODES * 3 + 3];   // + the spare node
    __shared__ int sred[(ADJ_BLOCK / 64) * 6 + 3];
    const int per = gridDim.x >> 3;                    // gridDim.x is a multiple of 8: consecutive
    const int lb = (blockIdx.x & 7) * per + (blockIdx.x >> 3);   // tiles share an XCD's L2
    if (lb * ADJ_BLOCK >= n) return;                   // whole block past the end (block-uniform)
    const int t = lb * ADJ_BLOCK + threadIdx.x;
    const bool live = t < n;                           // every other thread stays: the block synchronises
    const int s_m = (s_lo + s_hi) >> 1;               // the box's step (a measured 5 % faster than s_lo)
    const int slot = live ? order[t] : order[0];
    const int nn[3] = {g.nx, g.ny, g.nz};
    // box: offsets (wrapped to half a grid each way, so a block that meets a
    // periodic face stays compact) from the step-s_m cell m of the block's
    // middle particle, spanning the block's step-s_m cells, widened by the
    // stencil (one node above) and ADJ_MARGIN cells each way.  (Spanning every
    // step's cells of the sub-segment instead measured slower: the larger
    // boxes outgrow the LDS and are clamped.)
    auto cell_of = [&](const float* x, int c[3]) {
        for (int d = 0; d < 3; ++d) {
            float f = floorf((x[d] - g.o[d]) / g.d[d]);
            c[d] = (f != f) ? 0 : (int)fminf(fmaxf(f, 0.0f), (float)(nn[d] - 1));
        }
    };
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    constexpr int NW = ADJ_BLOCK / 64;
    const int mid = (min(n - 1, lb * ADJ_BLOCK + ADJ_BLOCK / 2) & (ADJ_BLOCK - 1));
    if ((int)threadIdx.x == mid) {
        int c[3];
        cell_of(seg + ((size_t)(s_m - sa) * n + slot) * 3, c);
        for (int d = 0; d < 3; ++d) sred[NW * 6 + d] = c[d];
    }
    for (int e = threadIdx.x; e < 2 * TILE_NODES * 3 + 3; e += blockDim.x) tile[e] = 0.0;
    __syncthreads();
    Tile tl;
    for (int d = 0; d < 3; ++d) tl.m[d] = sred[NW * 6 + d];
    {
        int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-0x7fffffff, -0x7fffffff, -0x7fffffff};
        if (live) {
            int c[3];
            cell_of(seg + ((size_t)(s_m - sa) * n + slot) * 3, c);
            for (int d = 0; d < 3; ++d) lo[d] = hi[d] = wrap_off(c[d] - tl.m[d], nn[d]);
        }
        for (int d = 0; d < 3; ++d) {
            const int l = wave_min_i(lo[d]), h = wave_max_i(hi[d]);
            if (lane == 0) { sred[wv * 6 + d] = l; sred[wv * 6 + 3 + d] = h; }
        }
    }
    __syncthreads();
    {
        long long vol = 1;
        int ext[3], l3[3], h3[3];
        for (int d = 0; d < 3; ++d) {
            int l = sred[d], h = sred[3 + d];
            for (int w = 1; w < NW; ++w) { l = min(l, sred[w * 6 + d]); h = max(h, sred[w * 6 + 3 + d]); }
            l3[d] = l;
            h3[d] = h + 1;
        }
        // the widest margin (ADJ_MARGIN down to 0 cells each way) whose box
        // still fits: particles drift out of a tight box during the sub-segment
        for (int mg = ADJ_MARGIN; mg >= 0; --mg) {
            vol = 1;
            for (int d = 0; d < 3; ++d) {
                tl.b[d] = l3[d] - mg;
                ext[d] = h3[d] + mg - tl.b[d] + 1;
                vol *= (ext[d] > 0 ? ext[d] : 0);
            }
            if (vol <= TILE_NODES) break;
        }
        if (vol > TILE_NODES) {                        // a cube of the largest fitting side around the
            int cap = 1;                               // block's middle particle (offset 0)
            while ((cap + 1) * (cap + 1) * (cap + 1) <= TILE_NODES) ++cap;
            vol = 1;
            for (int d = 0; d < 3; ++d) {
                if (ext[d] > cap) {
                    int b0 = -(cap / 2);
                    b0 = max(b0, tl.b[d]);
                    b0 = min(b0, tl.b[d] + ext[d] - cap);
                    tl.b[d] = b0;
                    ext[d] = cap;
                }
                vol *= ext[d];
            }
        }
        for (int d = 0; d < 3; ++d) tl.e[d] = ext[d];
        tl.vol = vol > 0 ? (int)vol : 0;
        int smin = tab[3 * s_lo].i0;                   // snapshot slots: the sub-segment's lowest index
        for (int st = s_lo; st < s_hi; ++st)
            for (int k = 0; k < 3; ++k) smin = min(smin, min(tab[3 * st + k].i0, tab[3 * st + k].i1));
        tl.smin = smin;
        if (tl.vol == 0) { tl.e[0] = tl.e[1] = tl.e[2] = 0; }
    }
    const int pid = live ? seg_pid[slot] : 0;
    float lam[3] = {0.0f, 0.0f, 0.0f};
    if (live) { lam[0] = lam_io[3 * (size_t)pid]; lam[1] = lam_io[3 * (size_t)pid + 1]; lam[2] = lam_io[3 * (size_t)pid + 2]; }
    RunAcc run;
    run.live = false;
    if (live) {
        for (int s = s_hi - 1; s >= s_lo; --s) {
            const float* xi = seg + ((size_t)(s - sa) * n + slot) * 3;
            const float p[3] = {xi[0], xi[1], xi[2]};
            const Blend bl0 = tab[3 * s], bl1 = tab[3 * s + 1], bl2 = tab[3 * s + 2];
            // forward stages with the forward kernel's arithmetic (rk_step),
            // keeping the Jacobians of stages 1-3 (Ja, Jb, Jc) and the stage
            // points x2, x3, x4; stage 4's Jacobian is used at once
            float q[3] = {p[0], p[1], p[2]}, acc[3], x2[3], x3[3], x4[3];
            float Ja[9], Jb[9], Jc[9];
            bool da = true, db = true, dc = true, d4 = true;
            float L[3], A3[3], l[3];
#pragma unroll 1
            for (int r = 0; r < 4; ++r) {
                const Blend b = r == 0 ? bl0 : (r == 3 ? bl2 : bl1);
                float k[3], Jn[9];
                bool dn;
                sample_jac<ST, ZH, PF, PR>(g, b, q, k, Jn, dn);
                if (r == 3) {
                    float xp[3];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        xp[c] = p[c] + dt6 * (acc[c] + k[c]);
                        L[c] = lam[c] * bc_deriv(bc.kind[c], xp[c], bc.lo[c], bc.hi[c]);
                        A3[c] = dt6 * L[c];
                    }
                    jac_t(Jn, A3, l);
                    d4 = dn;
                } else {
                    const float h = r == 2 ? dt : dth;
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        acc[c] = r == 0 ? k[c] : acc[c] + 2.0f * k[c];
                        q[c] = p[c] + h * k[c];
                        x2[c] = x3[c]; x3[c] = x4[c]; x4[c] = q[c];
                    }
#pragma unroll
                    for (int e = 0; e < 9; ++e) { Ja[e] = Jb[e]; Jb[e] = Jc[e]; Jc[e] = Jn[e]; }
                    da = db; db = dc; dc = dn;
                }
            }
            // reverse: a_r = w_r L + h_r l_(r+1), l_r = J_r^T a_r, lam = L + sum l_r
            float A[3][3];
#pragma unroll
            for (int c = 0; c < 3; ++c) { lam[c] = L[c] + l[c]; A[2][c] = fmaf(dt, l[c], 2.0f * dt6 * L[c]); }
            jac_t(Jc, A[2], l);
#pragma unroll
            for (int c = 0; c < 3; ++c) { lam[c] += l[c]; A[1][c] = fmaf(dth, l[c], 2.0f * dt6 * L[c]); }
            jac_t(Jb, A[1], l);
#pragma unroll
            for (int c = 0; c < 3; ++c) { lam[c] += l[c]; A[0][c] = fmaf(dth, l[c], dt6 * L[c]); }
            jac_t(Ja, A[0], l);
#pragma unroll
            for (int c = 0; c < 3; ++c) lam[c] += l[c];
            // dL/dU: stage points 4, 3, 2, 1 through the run accumulator
            float cur[3] = {x4[0], x4[1], x4[2]}, av[3] = {A3[0], A3[1], A3[2]};
            float y2[3] = {x3[0], x3[1], x3[2]}, y1[3] = {x2[0], x2[1], x2[2]};
            float b2[3] = {A[2][0], A[2][1], A[2][2]}, b1[3] = {A[1][0], A[1][1], A[1][2]};
            bool dcur = d4, d2 = dc, d1 = db;
#pragma unroll 1
            for (int r = 3; r >= 0; --r) {
                if (dcur) {
                    const Blend b = r == 0 ? bl0 : (r == 3 ? bl2 : bl1);
                    Stencil stc;
                    bool valid = true;
                    float w;
                    if constexpr (PF) {
                        axis_coord_pf(cur[0], g.o[0], g.inv[0], g.nx, (float)g.nx, stc.I[0], stc.I[1], w);
                        stc.wx[0] = 1.0f - w; stc.wx[1] = w;
                        axis_coord_pf(cur[1], g.o[1], g.inv[1], g.ny, (float)g.ny, stc.J[0], stc.J[1], w);
                        stc.wy[0] = 1.0f - w; stc.wy[1] = w;
                        axis_coord_pf(cur[2], g.o[2], g.inv[2], g.nz, (float)g.nz, stc.K[0], stc.K[1], w);
                        stc.wz[0] = 1.0f - w; stc.wz[1] = w;
                    } else {
                        axis_coord(g, 0, cur[0], g.nx, stc.I[0], stc.I[1], w, valid); stc.wx[0] = 1.0f - w; stc.wx[1] = w;
                        axis_coord(g, 1, cur[1], g.ny, stc.J[0], stc.J[1], w, valid); stc.wy[0] = 1.0f - w; stc.wy[1] = w;
                        axis_coord(g, 2, cur[2], g.nz, stc.K[0], stc.K[1], w, valid); stc.wz[0] = 1.0f - w; stc.wz[1] = w;
                    }
                    const bool same = run.live && stc.I[0] == run.I[0] && stc.J[0] == run.J[0] &&
                                      stc.K[0] == run.K[0] && stc.I[1] == run.I[1] && stc.J[1] == run.J[1] &&
                                      stc.K[1] == run.K[1] && b.i0 == run.snap[0] && b.i1 == run.snap[1];
                    if (!same) {
                        if (run.live) run_flush(g, run, tl, tile, gU);
                        run_reset(run, stc, b);
                    }
                    run_add(run, stc, b, av);
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    cur[c] = y2[c]; y2[c] = y1[c]; y1[c] = p[c];
                    av[c] = b2[c]; b2[c] = b1[c]; b1[c] = A[0][c];
                }
                dcur = d2; d2 = d1; d1 = da;
            }
        }
        if (run.live) run_flush(g, run, tl, tile, gU);
        lam_io[3 * (size_t)pid] = lam[0];
        lam_io[3 * (size_t)pid + 1] = lam[1];
        lam_io[3 * (size_t)pid + 2] = lam[2];
        if (s_lo == 0 && gx0 != nullptr) {
            gx0[3 * (size_t)pid] = lam[0];
            gx0[3 * (size_t)pid + 1] = lam[1];
            gx0[3 * (size_t)pid + 2] = lam[2];
        }
    }
    __syncthreads();
    const int tv = tl.vol * 3;
    const size_t gsnap = (size_t)g.nx * g.ny * g.nz * 3;
    for (int e = threadIdx.x; e < 2 * tv; e += blockDim.x) {        // flush the box
        const float v = (float)tile[e];
        if (v == 0.0f) continue;
        const int sl = e >= tv;
        const int r = e - sl * tv;
        const int node = r / 3, comp = r - node * 3;
        auto wrap_node = [](int v, int n) { v = v < 0 ? v + n : v; return v >= n ? v - n : v; };
        const int k = wrap_node(tl.m[2] + node % tl.e[2] + tl.b[2], g.nz);
        const int t2 = node / tl.e[2];
        const int j = wrap_node(tl.m[1] + t2 % tl.e[1] + tl.b[1], g.ny);
        const int i = wrap_node(tl.m[0] + t2 / tl.e[1] + tl.b[0], g.nx);
        atomicAdd(gU + (size_t)(tl.smin + sl) * gsnap + ((size_t)(i * g.ny + j) * g.nz + k) * 3 + comp, v);
    }
}

#define JT_ADJ_DISPATCH(KERNEL, GRID, BLOCK, ...)                                                        \
    do {                                                                                                  \
        if (zh) {                                                                                         \
            switch (st) {                                                                                 \
            case ST_F32x3: hipLaunchKernelGGL((KERNEL<ST_F32x3, true>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
            case ST_F32x4: hipLaunchKernelGGL((KERNEL<ST_F32x4, true>), GRID, BLOCK, 0, s, __VA_ARGS__); break;   \
            case ST_BF16x4: hipLaunchKernelGGL((KERNEL<ST_BF16x4, true>), GRID, BLOCK, 0, s, __VA_ARGS__); break; \
            default: hipLaunchKernelGGL((KERNEL<ST_F16x4, true>), GRID, BLOCK, 0, s, __VA_ARGS__); break;         \
            }                                                                                             \
        } else {                                                                                          \
            switch (st) {                                                                                 \
            case ST_F32x3: hipLaunchKernelGGL((KERNEL<ST_F32x3, false>), GRID, BLOCK, 0, s, __VA_ARGS__); break;  \
            case ST_F32x4: hipLaunchKernelGGL((KERNEL<ST_F32x4, false>), GRID, BLOCK, 0, s, __VA_ARGS__); break;  \
            case ST_BF16x4: hipLaunchKernelGGL((KERNEL<ST_BF16x4, false>), GRID, BLOCK, 0, s, __VA_ARGS__); break;\
            default: hipLaunchKernelGGL((KERNEL<ST_F16x4, false>), GRID, BLOCK, 0, s, __VA_ARGS__); break;        \
            }                                                                                             \
        }                                                                                                 \
    } while (0)

// ---- bucket ordering for the backward --------------------------------------
// The reverse kernel needs a block of consecutive threads to start in a
// compact region (its LDS box), not a total order, so particles are ordered
// by a counting sort over super-cells (2^ADJ_BUCKET_SHIFT cells a side) taken
// along the Hilbert curve: three small kernels and a scan instead of a radix
// sort.  The order within a super-cell is that of the atomic slot grabs.
#ifndef ADJ_BUCKET_SHIFT
#define ADJ_BUCKET_SHIFT 1
#endif

// Hilbert index of each particle's super-cell.
__global__ void __launch_bounds__(256)
k_bucket_keys(GridDev g, const float* __restrict__ xs, int n, int shift, int cbits, unsigned* __restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned c[3];
    const int nn[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
        float f = (xs[3 * (size_t)i + a] - g.o[a]) / g.d[a];
        int ci = (f != f) ? 0 : (int)fminf(fmaxf(floorf(f), 0.0f), (float)(nn[a] - 1));
        c[a] = (unsigned)ci >> shift;
    }
    keys[i] = hilbert3(c[0], c[1], c[2], cbits);
}

// Bits per axis of the super-cell lattice: the grid's Hilbert bits less
// ADJ_BUCKET_SHIFT, and at most log2(n)/3 (about one super-cell per
// particle at most).
static int bucket_bits(int bits, int64_t n) {
    int cb = bits - ADJ_BUCKET_SHIFT;
    int lg = 0;
    while ((2LL << lg) <= n) ++lg;
    cb = cb < lg / 3 ? cb : lg / 3;
    return cb < 1 ? 1 : cb;
}

// out[j] = the index i of the j-th particle of xs in super-cell order.
static int bucket_order(hipStream_t s, const GridDev& g, const float* xs, int n, int shift, int bits, int cb,
                        unsigned* keys, int* counts, int* offs, char* tmp, size_t temp, int* out) {
    const int nb = 1 << (3 * cb);
    hipLaunchKernelGGL(k_bucket_keys, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, xs, n,
                       shift + (bits - cb), cb, keys);
    JT_HIP(hipGetLastError());
    JT_HIP(bucket_sort(s, keys, n, 0, nb, counts, offs, tmp, temp, out));
    return JT_OK;
}

}  // namespace jt

using namespace jt;

extern "C" size_t jt_track_grid_adjoint_workspace_size(int64_t n, int32_t n_steps, int32_t ckpt_every) {
    Arena a(nullptr, 0);
    const size_t nn = (size_t)(n > 0 ? n : 1);
    a.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    a.take<float>(3 * nn);                                          // adjoint state
    a.take<float>(3 * nn * (ckpt_every > 0 ? ckpt_every : 1));      // segment states
    a.take<unsigned>(nn); a.take<int>(nn); a.take<int>(nn);       // bucket keys, the two orders
    const int nb = 1 << (3 * bucket_bits(10, (int64_t)nn));         // the largest bucket lattice for n
    a.take<int>(nb); a.take<int>(nb);                               // counts, offsets
    a.take<char>(bucket_scan_bytes(nb));
    return a.used + 256;
}

extern "C" int jt_track_grid_adjoint(const jt_grid_field* field, const jt_boundary* bc, const float* ckpt,
                                     int64_t n, const float* step_times, const float* host_step_times,
                                     const float* host_field_times, float dt, int32_t n_steps,
                                     int32_t ckpt_every, const float* grad_xT, float* grad_x0, float* grad_U,
                                     void* ws, size_t ws_bytes, void* stream) {
    GridDev g;
    int st;
    int rc = make_grid_dev(field, g, st);
    if (rc) return rc;
    BcDev d;
    rc = make_bc_dev(bc, d);
    if (rc) return rc;
    if (d.inlet != JT_INLET_OFF || d.absorb_rows)
        return fail(JT_EUNSUPPORTED, "the adjoint supports periodic, reflective, clamping and open boundaries");
    for (int a = 0; a < 3; ++a)
        if (d.kind[a] == JT_BC_ABSORB)
            return fail(JT_EUNSUPPORTED, "the adjoint does not differentiate absorbing boundaries");
    if (g.bmode == JT_BOUNDS_NAN) return fail(JT_EUNSUPPORTED, "the adjoint needs clamp, zero or periodic bounds");
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(n_steps >= 0 && ckpt_every >= 1, "bad n_steps / ckpt_every");
    if (n == 0) return JT_OK;
    JT_REQUIRE(ckpt && grad_xT && grad_U && (n_steps == 0 || (step_times && host_step_times && host_field_times)),
               "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    Arena arena(ws, ws_bytes);
    Blend* tab = arena.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    float* lam = arena.take<float>((size_t)3 * n);
    float* seg = arena.take<float>((size_t)3 * n * ckpt_every);
    unsigned* keys = arena.take<unsigned>(n);
    int* perm = arena.take<int>(n);                   // segment slot -> particle
    int* order = arena.take<int>(n);                  // reverse thread -> segment slot
    const int nb_max = 1 << (3 * bucket_bits(10, n));
    int* counts = arena.take<int>(nb_max);
    int* offs = arena.take<int>(nb_max);
    const size_t temp = bucket_scan_bytes(nb_max);
    char* tmp = arena.take<char>(temp);
    if (!tab || !lam || !seg || !order || !tmp || !arena.ok())
        return fail(JT_ENOMEM, "workspace too small: need %zu, got %zu",
                    jt_track_grid_adjoint_workspace_size(n, n_steps, ckpt_every), ws_bytes);
    JT_HIP(hipMemcpyAsync(lam, grad_xT, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (n_steps == 0) {
        if (grad_x0) JT_HIP(hipMemcpyAsync(grad_x0, grad_xT, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
        return JT_OK;
    }
    hipLaunchKernelGGL(k_blend_table, dim3((n_steps + 127) / 128), dim3(128), 0, s, g, step_times, n_steps, dt,
                       JT_RK4, tab);
    JT_HIP(hipGetLastError());
    const float dth = 0.5f * dt, dt6 = dt / 6.0f;
    const int n_seg = (n_steps + ckpt_every - 1) / ckpt_every;
    const int R = ADJ_RESORT;                          // longest sub-segment (re-sort interval)
    int maxn = g.nx > g.ny ? g.nx : g.ny;
    maxn = maxn > g.nz ? maxn : g.nz;
    int shift = 0;
    while ((maxn - 1) >> shift >= 1024) ++shift;
    int bits = 1;                                      // bits per axis of the shifted cell index
    while (((maxn - 1) >> shift) >> bits) ++bits;
    const int cb = bucket_bits(bits, n);
    // the stage blends on the host, from the caller's host copies of the
    // step times and snapshot times (k_blend_table's float32 operations, so
    // the same snapshot pairs), to cut sub-segments where the pair changes
    // (the LDS box holds two snapshot slots) without reading the device
    std::vector<Blend> htab((size_t)3 * n_steps);
    for (int k = 0; k < n_steps; ++k) {
        const float t = host_step_times[k];
        const float t1 = t + dth, t2 = t + dt;
        htab[3 * (size_t)k + 0] = time_blend(host_field_times, g.T, t, g.interp, g.extrap);
        htab[3 * (size_t)k + 1] = time_blend(host_field_times, g.T, t1, g.interp, g.extrap);
        htab[3 * (size_t)k + 2] = time_blend(host_field_times, g.T, t2, g.interp, g.extrap);
    }
    dim3 b256(256), g256((unsigned)((n + 255) / 256));
    dim3 g256x((unsigned)((((n + 255) / 256) + 7) / 8 * 8));   // multiple of 8 for the XCD remap
    dim3 badj(ADJ_BLOCK), gadj((unsigned)((((n + ADJ_BLOCK - 1) / ADJ_BLOCK) + 7) / 8 * 8));
    const bool zh = g.nzs != g.nz;
    const bool pf = zh && periodic_fast(g, st == ST_F32x3 ? 12 : (st == ST_F32x4 ? 16 : 8)) && bc_axes_only(d);
    const bool pr = pf && st == ST_F32x4 && g.pair != nullptr;
    for (int c = n_seg - 1; c >= 0; --c) {
        const int sa = c * ckpt_every;
        const int sb = sa + ckpt_every < n_steps ? sa + ckpt_every : n_steps;
        const float* xc = ckpt + (size_t)c * n * 3;
        rc = bucket_order(s, g, xc, (int)n, shift, bits, cb, keys, counts, offs, tmp, temp, perm);
        if (rc) return rc;
        if (pf) {
            switch (st) {
            case ST_F32x3: hipLaunchKernelGGL((k_adj_recompute<ST_F32x3, true, true>), g256x, b256, 0, s, g, d, xc,
                                              (int)n, (const int*)perm, tab, sa, sb, dt, dth, dt6, seg); break;
            case ST_F32x4:
                if (pr) hipLaunchKernelGGL((k_adj_recompute<ST_F32x4, true, true, true>), g256x, b256, 0, s, g, d, xc,
                                           (int)n, (const int*)perm, tab, sa, sb, dt, dth, dt6, seg);
                else hipLaunchKernelGGL((k_adj_recompute<ST_F32x4, true, true>), g256x, b256, 0, s, g, d, xc,
                                        (int)n, (const int*)perm, tab, sa, sb, dt, dth, dt6, seg);
                break;
            case ST_BF16x4: hipLaunchKernelGGL((k_adj_recompute<ST_BF16x4, true, true>), g256x, b256, 0, s, g, d, xc,
                                               (int)n, (const int*)perm, tab, sa, sb, dt, dth, dt6, seg); break;
            default: hipLaunchKernelGGL((k_adj_recompute<ST_F16x4, true, true>), g256x, b256, 0, s, g, d, xc,
                                        (int)n, (const int*)perm, tab, sa, sb, dt, dth, dt6, seg); break;
            }
        } else {
            JT_ADJ_DISPATCH(k_adj_recompute, g256x, b256, g, d, xc, (int)n, (const int*)perm, tab, sa, sb, dt, dth,
                            dt6, seg);
        }
        JT_HIP(hipGetLastError());
        // sub-segments of at most R steps, cut where a third snapshot would enter
        int s_hi = sb;
        while (s_hi > sa) {
            int s_lo = s_hi - 1;
            int u0 = htab[3 * (size_t)s_lo].i0, u1 = u0;
            auto fits = [&](int st) {
                int a0 = u0, a1 = u1;
                for (int k = 0; k < 3; ++k) {
                    const Blend& b = htab[3 * (size_t)st + k];
                    a0 = std::min(a0, std::min(b.i0, b.i1));
                    a1 = std::max(a1, std::max(b.i0, b.i1));
                }
                if (a1 - a0 > 1) return false;
                u0 = a0; u1 = a1;
                return true;
            };
            fits(s_lo);
            while (s_lo > sa && s_hi - s_lo < R && fits(s_lo - 1)) --s_lo;
            // order the segment slots by their cell at the sub-segment's middle step
            rc = bucket_order(s, g, seg + (size_t)(((s_lo + s_hi) >> 1) - sa) * n * 3, (int)n, shift, bits, cb, keys, counts, offs,
                              tmp, temp, order);
            if (rc) return rc;
            if (pf) {
                switch (st) {
                case ST_F32x3: hipLaunchKernelGGL((k_adj_reverse<ST_F32x3, true, true>), gadj, badj, 0, s, g, d, (int)n,
                                                  (const int*)order, (const int*)perm, tab, sa, s_lo, s_hi, dt, dth, dt6,
                                                  (const float*)seg, lam, grad_U, grad_x0); break;
                case ST_F32x4:
                    if (pr) hipLaunchKernelGGL((k_adj_reverse<ST_F32x4, true, true, true>), gadj, badj, 0, s, g, d,
                                               (int)n, (const int*)order, (const int*)perm, tab, sa, s_lo, s_hi, dt,
                                               dth, dt6, (const float*)seg, lam, grad_U, grad_x0);
                    else hipLaunchKernelGGL((k_adj_reverse<ST_F32x4, true, true>), gadj, badj, 0, s, g, d, (int)n,
                                            (const int*)order, (const int*)perm, tab, sa, s_lo, s_hi, dt, dth, dt6,
                                            (const float*)seg, lam, grad_U, grad_x0);
                    break;
                case ST_BF16x4: hipLaunchKernelGGL((k_adj_reverse<ST_BF16x4, true, true>), gadj, badj, 0, s, g, d,
                                                   (int)n, (const int*)order, (const int*)perm, tab, sa, s_lo, s_hi, dt,
                                                   dth, dt6, (const float*)seg, lam, grad_U, grad_x0); break;
                default: hipLaunchKernelGGL((k_adj_reverse<ST_F16x4, true, true>), gadj, badj, 0, s, g, d, (int)n,
                                            (const int*)order, (const int*)perm, tab, sa, s_lo, s_hi, dt, dth, dt6,
                                            (const float*)seg, lam, grad_U, grad_x0); break;
                }
            } else {
                JT_ADJ_DISPATCH(k_adj_reverse, gadj, badj, g, d, (int)n, (const int*)order, (const int*)perm, tab, sa,
                                s_lo, s_hi, dt, dth, dt6, (const float*)seg, lam, grad_U, grad_x0);
            }
            JT_HIP(hipGetLastError());
            s_hi = s_lo;
        }
    }
    return JT_OK;
}
