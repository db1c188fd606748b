// Adjoint (reverse-mode) RK4 tracking through a time-blended structured grid.
//
// The reference has no differentiation; this is the build-only gradient of
// the forward path of track_grid.hip: given dL/dx_T it returns dL/dx_0 and
// dL/dU for the snapshots U, with positions checkpointed every K steps by the
// forward run (its trajectory recorded every K steps) and recomputed in
// between.  One thread owns one particle; its adjoint state stays in registers
// across all segments of a launch.  Per step the forward stages are recomputed
// with the forward kernel's exact arithmetic, then reversed:
//   x' = BC(x + dt/6 (k1 + 2k2 + 2k3 + k4)),  k_i = u(x_i, t_i)
// using the trilinear Jacobian (corner weights' derivatives times the blended
// corner values) and scattering  coef_snap * W_corner * abar  into dL/dU with
// hardware float atomics (global_atomic_add_f32).
#include "grid_sampler.h"
#include <hipcub/hipcub.hpp>

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

// Per-axis weights (1-w, w) and their x-derivatives.
__device__ __forceinline__ void axis_w(const GridDev& g, int ax, float x, int n, int& i0, int& i1,
                                       float w[2], float dw[2], bool& valid) {
    float raw = (x - g.o[ax]) / g.d[ax];
    float ww;
    axis_coord(g, ax, x, n, i0, i1, ww, valid);
    float gd = 1.0f / g.d[ax];
    if (g.bmode != JT_BOUNDS_PERIODIC && !(raw >= 0.0f && raw <= g.cmax[ax])) gd = 0.0f;   // clipped
    w[0] = 1.0f - ww; w[1] = ww;
    dw[0] = -gd; dw[1] = gd;
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

// ---- block-level accumulation of the dL/dU scatter ------------------------
// The threads of a block hold cell-sorted particles, so over a few steps the
// corners they scatter into stay in a small box of nodes.  A reverse launch
// sums its whole sub-segment's contributions for the box (two snapshot
// slots) in LDS with ds_add_f32 and flushes each touched entry to HBM once
// with a float atomic; corners outside the box (or in another snapshot) go
// straight to HBM.
#define TILE_NODES 3072
#define ADJ_BLOCK 256

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

struct Tile {
    int b[3], e[3];     // box origin (node indices) and extents
    int smin;           // first snapshot slot
    int vol;            // e0*e1*e2 (0: no box)
};

struct Stencil {
    int I[2], J[2], K[2];
    float W[8];
};

// J^T a and the stencil of u at q (no scatter).
template <int ST, bool ZH>
__device__ __forceinline__ void vjp_local(const GridDev& g, const Blend& bl, const float q[3], const float a[3],
                                          float lam[3], Stencil& st, bool& depends) {
    float wx[2], wy[2], wz[2], dx[2], dy[2], dz[2];
    bool valid = true;
    axis_w(g, 0, q[0], g.nx, st.I[0], st.I[1], wx, dx, valid);
    axis_w(g, 1, q[1], g.ny, st.J[0], st.J[1], wy, dy, valid);
    axis_w(g, 2, q[2], g.nz, st.K[0], st.K[1], wz, dz, valid);
    lam[0] = lam[1] = lam[2] = 0.0f;
    depends = !(!valid && g.bmode == JT_BOUNDS_ZERO);
    const char* base = (const char*)g.data;
    const size_t sb = (size_t)g.snap_stride * ElemBytes<ST>::v;
    const void* b0 = base + (size_t)bl.i0 * sb;
    const void* b1 = base + (size_t)bl.i1 * sb;
    const float oma = 1.0f - bl.a;
#pragma unroll
    for (int ci = 0; ci < 2; ++ci)
#pragma unroll
        for (int cj = 0; cj < 2; ++cj)
#pragma unroll
            for (int ck = 0; ck < 2; ++ck) {
                const int c = ci * 4 + cj * 2 + ck;
                st.W[c] = depends ? wx[ci] * wy[cj] * wz[ck] : 0.0f;
                if (!depends) continue;
                float v[3];
                corner<ST>(b0, b1, bl, oma, (st.I[ci] * g.ny + st.J[cj]) * g.nzs + st.K[ck], v);
                const float s = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
                lam[0] += s * dx[ci] * wy[cj] * wz[ck];
                lam[1] += s * wx[ci] * dy[cj] * wz[ck];
                lam[2] += s * wx[ci] * wy[cj] * dz[ck];
            }
}

// Add coef_snap * W_c * a for the 8 corners: into the LDS tile when the
// corner and snapshot are covered, else straight to HBM.
__device__ __forceinline__ void scatter(const GridDev& g, const Blend& bl, const Stencil& st, const float a[3],
                                        const Tile& tl, float* tile, float* __restrict__ gU) {
    float c[2];
    blend_coefs(bl, c[0], c[1]);
    const int snap[2] = {bl.i0, bl.i1};
    const size_t gsnap = (size_t)g.nx * g.ny * g.nz * 3;
    const int tv = tl.vol * 3;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        const int ci = q >> 2, cj = (q >> 1) & 1, ck = q & 1;
        const int I = st.I[ci], J = st.J[cj], K = st.K[ck];
        const int li = I - tl.b[0], lj = J - tl.b[1], lk = K - tl.b[2];
        const bool in = tl.vol > 0 && (unsigned)li < (unsigned)tl.e[0] && (unsigned)lj < (unsigned)tl.e[1] &&
                        (unsigned)lk < (unsigned)tl.e[2];
        const int loc = (li * tl.e[1] + lj) * tl.e[2] + lk;
        const size_t node = ((size_t)(I * g.ny + J) * g.nz + K) * 3;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (c[h] == 0.0f) continue;
            const float f = c[h] * st.W[q];
            const int slot = snap[h] - tl.smin;
            if (in && (slot == 0 || slot == 1)) {
                float* tp = tile + slot * tv + loc * 3;
                atomicAdd(tp, f * a[0]);
                atomicAdd(tp + 1, f * a[1]);
                atomicAdd(tp + 2, f * a[2]);
            } else {
                float* gp = gU + (size_t)snap[h] * gsnap + node;
                atomicAdd(gp, f * a[0]);
                atomicAdd(gp + 1, f * a[1]);
                atomicAdd(gp + 2, f * a[2]);
            }
        }
    }
}

// Forward recompute of one checkpoint segment [sa, sb): thread t takes
// particle perm[t] (cell-sorted) and stores the state before each step at
// seg[(s - sa) * n + pid] (particle order, so any later ordering can read it).
template <int ST, bool ZH>
__global__ void __launch_bounds__(256)
k_adj_recompute(GridDev g, BcDev bc, const float* __restrict__ x0s, int n, const int* __restrict__ perm,
                const Blend* __restrict__ tab, int sa, int sb, float dt, float dth, float dt6,
                float* __restrict__ seg) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int pid = perm[t];
    float x[3] = {x0s[3 * (size_t)pid], x0s[3 * (size_t)pid + 1], x0s[3 * (size_t)pid + 2]};
    auto f = [&](const Blend& b, const float q[3], float out[3]) { sample<ST, ZH>(g, b, q[0], q[1], q[2], out); };
    for (int s = sa; s < sb; ++s) {
        float* o = seg + ((size_t)(s - sa) * n + pid) * 3;
        o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
        rk_step<JT_RK4>(f, tab, s, x, dt, dth, dt6);
        apply_bc(bc, x);
    }
}

// Reverse steps s_hi-1 .. s_lo of the segment starting at sa.  Thread t takes
// particle perm[t], sorted by its cell at step s_lo, so the block's corners
// stay in a small box; its adjoint lives in lam_io[pid] between launches.
template <int ST, bool ZH>
__global__ void __launch_bounds__(ADJ_BLOCK)
k_adj_reverse(GridDev g, BcDev bc, int n, const int* __restrict__ perm, const Blend* __restrict__ tab,
              int sa, int s_lo, int s_hi, float dt, float dth, float dt6, const float* __restrict__ seg,
              float* __restrict__ lam_io, float* __restrict__ gU, float* __restrict__ gx0) {
    __shared__ float tile[2 * TILE_NODES * 3];
    __shared__ int sred[(ADJ_BLOCK / 64) * 6];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < n;                           // every thread stays: the block synchronises
    const int pid = live ? perm[t] : perm[0];
    // box: cells of this thread's states over the sub-segment, widened by the
    // stencil reach (one cell below, two above), reduced over the block
    int lo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, hi[3] = {-1, -1, -1};
    if (live) {
        const int nn[3] = {g.nx, g.ny, g.nz};
        for (int s = s_lo; s < s_hi; ++s) {
            const float* xi = seg + ((size_t)(s - sa) * n + pid) * 3;
            for (int d = 0; d < 3; ++d) {
                float f = floorf((xi[d] - g.o[d]) / g.d[d]);
                int c = (f != f) ? 0 : (int)fminf(fmaxf(f, -4.0f), (float)(nn[d] + 4));
                lo[d] = min(lo[d], c - 1);
                hi[d] = max(hi[d], c + 2);
            }
        }
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int d = 0; d < 3; ++d) {
        int l = wave_min_i(lo[d]), h = wave_max_i(hi[d]);
        if (lane == 0) { sred[wv * 6 + d] = l; sred[wv * 6 + 3 + d] = h; }
    }
    for (int e = threadIdx.x; e < 2 * TILE_NODES * 3; e += blockDim.x) tile[e] = 0.0f;
    __syncthreads();
    Tile tl;
    {
        const int nn[3] = {g.nx, g.ny, g.nz};
        long long vol = 1;
        int ext[3];
        for (int d = 0; d < 3; ++d) {
            int l = sred[d], h = sred[3 + d];
            for (int w = 1; w < ADJ_BLOCK / 64; ++w) { l = min(l, sred[w * 6 + d]); h = max(h, sred[w * 6 + 3 + d]); }
            l = max(l, 0);
            h = min(h, nn[d] - 1);
            tl.b[d] = l;
            ext[d] = h - l + 1;
            vol *= (ext[d] > 0 ? ext[d] : 0);
        }
        if (vol > TILE_NODES) {                        // keep a central sub-box that fits
            int cap = 1;
            while ((cap + 1) * (cap + 1) * (cap + 1) <= TILE_NODES) ++cap;
            vol = 1;
            for (int d = 0; d < 3; ++d) {
                if (ext[d] > cap) { tl.b[d] += (ext[d] - cap) / 2; ext[d] = cap; }
                vol *= ext[d];
            }
        }
        for (int d = 0; d < 3; ++d) tl.e[d] = ext[d];
        tl.vol = vol > 0 ? (int)vol : 0;
        int smin = tab[3 * s_lo].i0;                   // snapshot slots: the sub-segment's lowest index
        for (int s = s_lo; s < s_hi; ++s)
            for (int k = 0; k < 3; ++k) smin = min(smin, min(tab[3 * s + k].i0, tab[3 * s + k].i1));
        tl.smin = smin;
    }
    float lam[3] = {lam_io[3 * (size_t)pid], lam_io[3 * (size_t)pid + 1], lam_io[3 * (size_t)pid + 2]};
    auto f = [&](const Blend& b, const float q[3], float out[3]) { sample<ST, ZH>(g, b, q[0], q[1], q[2], out); };
    bool needs_xp = false;
    for (int a = 0; a < 3; ++a) needs_xp = needs_xp || bc.kind[a] == JT_BC_REFLECTIVE || bc.kind[a] == JT_BC_CLAMP;
    if (live) {
        for (int s = s_hi - 1; s >= s_lo; --s) {
            const float* xi = seg + ((size_t)(s - sa) * n + pid) * 3;
            float p[3] = {xi[0], xi[1], xi[2]};
            const Blend bl0 = tab[3 * s], bl1 = tab[3 * s + 1], bl2 = tab[3 * s + 2];
            float k1[3], k2[3], k3[3], k4[3], x2[3], x3[3], x4[3], xp[3];
            f(bl0, p, k1);
            for (int a = 0; a < 3; ++a) x2[a] = p[a] + dth * k1[a];
            f(bl1, x2, k2);
            for (int a = 0; a < 3; ++a) x3[a] = p[a] + dth * k2[a];
            f(bl1, x3, k3);
            for (int a = 0; a < 3; ++a) x4[a] = p[a] + dt * k3[a];
            if (needs_xp) {                            // the boundary's derivative depends on x'
                f(bl2, x4, k4);
                for (int a = 0; a < 3; ++a) xp[a] = p[a] + dt6 * (((k1[a] + 2.0f * k2[a]) + 2.0f * k3[a]) + k4[a]);
            } else {
                xp[0] = xp[1] = xp[2] = 0.0f;
            }
            float a1[3], a2[3], a3[3], a4[3], l[3];
            for (int a = 0; a < 3; ++a) {
                lam[a] *= bc_deriv(bc.kind[a], xp[a], bc.lo[a], bc.hi[a]);
                a4[a] = dt6 * lam[a];
                a3[a] = 2.0f * dt6 * lam[a];
                a2[a] = 2.0f * dt6 * lam[a];
                a1[a] = dt6 * lam[a];
            }
            Stencil stc;
            bool dep;
            vjp_local<ST, ZH>(g, bl2, x4, a4, l, stc, dep);
            if (dep) scatter(g, bl2, stc, a4, tl, tile, gU);
            for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a3[a] += dt * l[a]; }
            vjp_local<ST, ZH>(g, bl1, x3, a3, l, stc, dep);
            if (dep) scatter(g, bl1, stc, a3, tl, tile, gU);
            for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a2[a] += dth * l[a]; }
            vjp_local<ST, ZH>(g, bl1, x2, a2, l, stc, dep);
            if (dep) scatter(g, bl1, stc, a2, tl, tile, gU);
            for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a1[a] += dth * l[a]; }
            vjp_local<ST, ZH>(g, bl0, p, a1, l, stc, dep);
            if (dep) scatter(g, bl0, stc, a1, tl, tile, gU);
            for (int a = 0; a < 3; ++a) lam[a] += l[a];
        }
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
        const float v = tile[e];
        if (v == 0.0f) continue;
        const int slot = e >= tv;
        const int r = e - slot * tv;
        const int node = r / 3, comp = r - node * 3;
        const int k = node % tl.e[2] + tl.b[2];
        const int t2 = node / tl.e[2];
        const int j = t2 % tl.e[1] + tl.b[1];
        const int i = t2 / tl.e[1] + tl.b[0];
        atomicAdd(gU + (size_t)(tl.smin + slot) * gsnap + ((size_t)(i * g.ny + j) * g.nz + k) * 3 + comp, v);
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

// Morton key of each particle's cell (10 bits per axis) for the backward sort.
__device__ __forceinline__ unsigned adj_spread3(unsigned v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ void __launch_bounds__(256)
k_adj_keys(GridDev g, const float* __restrict__ xs, int n, int shift, unsigned* __restrict__ keys,
           int* __restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned c[3];
    const int nn[3] = {g.nx, g.ny, g.nz};
    for (int a = 0; a < 3; ++a) {
        float f = (xs[3 * (size_t)i + a] - g.o[a]) / g.d[a];
        int ci = (f != f) ? 0 : (int)fminf(fmaxf(floorf(f), 0.0f), (float)(nn[a] - 1));
        c[a] = (unsigned)ci >> shift;
    }
    keys[i] = (adj_spread3(c[0]) << 2) | (adj_spread3(c[1]) << 1) | adj_spread3(c[2]);
    idx[i] = i;
}

}  // namespace jt

using namespace jt;

extern "C" size_t jt_track_grid_adjoint_workspace_size(int64_t n, int32_t n_steps, int32_t ckpt_every) {
    Arena a(nullptr, 0);
    const size_t nn = (size_t)(n > 0 ? n : 1);
    a.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    a.take<float>(3 * nn);                                          // adjoint state
    a.take<float>(3 * nn * (ckpt_every > 0 ? ckpt_every : 1));      // segment states
    a.take<unsigned>(nn); a.take<unsigned>(nn); a.take<int>(nn); a.take<int>(nn);   // cell sort
    size_t temp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, (int)nn, 0, 30);
    a.take<char>(temp);
    return a.used + 256;
}

extern "C" int jt_track_grid_adjoint(const jt_grid_field* field, const jt_boundary* bc, const float* ckpt,
                                     int64_t n, const float* step_times, float dt, int32_t n_steps,
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
    JT_REQUIRE(ckpt && grad_xT && grad_U && (n_steps == 0 || step_times), "NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    Arena arena(ws, ws_bytes);
    Blend* tab = arena.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    float* lam = arena.take<float>((size_t)3 * n);
    float* seg = arena.take<float>((size_t)3 * n * ckpt_every);
    unsigned* keys = arena.take<unsigned>(n);
    unsigned* skeys = arena.take<unsigned>(n);
    int* idx = arena.take<int>(n);
    int* perm = arena.take<int>(n);
    size_t temp = 0;
    hipcub::DeviceRadixSort::SortPairs(nullptr, temp, (const unsigned*)nullptr, (unsigned*)nullptr,
                                       (const int*)nullptr, (int*)nullptr, (int)n, 0, 30);
    char* tmp = arena.take<char>(temp);
    if (!tab || !lam || !seg || !tmp || !arena.ok())
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
    const int R = 8;                                   // re-sort interval inside a segment
    int maxn = g.nx > g.ny ? g.nx : g.ny;
    maxn = maxn > g.nz ? maxn : g.nz;
    int shift = 0;
    while ((maxn - 1) >> shift >= 1024) ++shift;
    dim3 b256(256), g256((unsigned)((n + 255) / 256));
    dim3 badj(ADJ_BLOCK), gadj((unsigned)((n + ADJ_BLOCK - 1) / ADJ_BLOCK));
    const bool zh = g.nzs != g.nz;
    for (int c = n_seg - 1; c >= 0; --c) {
        const int sa = c * ckpt_every;
        const int sb = sa + ckpt_every < n_steps ? sa + ckpt_every : n_steps;
        const float* xc = ckpt + (size_t)c * n * 3;
        hipLaunchKernelGGL(k_adj_keys, g256, b256, 0, s, g, xc, (int)n, shift, keys, idx);
        JT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, skeys, idx, perm, (int)n, 0, 30, s));
        JT_ADJ_DISPATCH(k_adj_recompute, g256, b256, g, d, xc, (int)n, (const int*)perm, tab, sa, sb, dt, dth, dt6,
                        seg);
        JT_HIP(hipGetLastError());
        const int nsub = (sb - sa + R - 1) / R;
        for (int j = nsub - 1; j >= 0; --j) {
            const int s_lo = sa + j * R;
            const int s_hi = s_lo + R < sb ? s_lo + R : sb;
            // order particles by their cell at the start of this sub-segment
            hipLaunchKernelGGL(k_adj_keys, g256, b256, 0, s, g, seg + (size_t)(s_lo - sa) * n * 3, (int)n, shift,
                               keys, idx);
            JT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, skeys, idx, perm, (int)n, 0, 30, s));
            JT_ADJ_DISPATCH(k_adj_reverse, gadj, badj, g, d, (int)n, (const int*)perm, tab, sa, s_lo, s_hi, dt, dth,
                            dt6, (const float*)seg, lam, grad_U, grad_x0);
            JT_HIP(hipGetLastError());
        }
    }
    return JT_OK;
}
