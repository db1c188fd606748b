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

// ---- block-level pre-reduction of the dL/dU scatter ----------------------
// The threads of a block hold cell-sorted particles, so the corners they
// scatter into cluster in a small box of nodes.  Per stage the block reduces
// that box; if it holds at most TILE_NODES nodes (and does not wrap around a
// periodic edge) contributions are summed in LDS with ds_add_f32 and each
// touched node is flushed to HBM with one float atomic per component;
// otherwise every contribution goes straight to HBM.
#define TILE_NODES 2048

__device__ __forceinline__ int wave_min_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
    return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
    return v;
}

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

// Block-cooperative scatter of coef_snap * W_c * a into dL/dU.  Every thread
// of the block must call it (active=false for threads without a particle).
__device__ inline void block_scatter(const GridDev& g, const Blend& bl, const Stencil& st, const float a[3],
                                     bool active, float* __restrict__ gU, float* tile, int* sred) {
    float c0, c1;
    blend_coefs(bl, c0, c1);
    const size_t gsnap = (size_t)g.nx * g.ny * g.nz * 3;
    float* g0 = gU + (size_t)bl.i0 * gsnap;
    float* g1 = gU + (size_t)bl.i1 * gsnap;
    // box of the block's corners
    int lo[3], hi[3];
    const int* ax[3] = {st.I, st.J, st.K};
    for (int d = 0; d < 3; ++d) {
        int l = min(ax[d][0], ax[d][1]), h = max(ax[d][0], ax[d][1]);
        if (ax[d][1] < ax[d][0]) { l = 0; h = 1 << 20; }          // periodic wrap: no tile
        lo[d] = active ? l : 0x7fffffff;
        hi[d] = active ? h : -1;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int d = 0; d < 3; ++d) {
        int l = wave_min_i(lo[d]), h = wave_max_i(hi[d]);
        if (lane == 0) { sred[wv * 6 + d] = l; sred[wv * 6 + 3 + d] = h; }
    }
    __syncthreads();
    int bx[3], ex[3];
    long long vol = 1;
    for (int d = 0; d < 3; ++d) {
        int l = sred[d], h = sred[3 + d];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w) { l = min(l, sred[w * 6 + d]); h = max(h, sred[w * 6 + 3 + d]); }
        bx[d] = l;
        ex[d] = h - l + 1;
        vol *= (ex[d] > 0 ? ex[d] : 0);
    }
    __syncthreads();                                   // sred is reused by the next call
    if (vol <= 0) return;                              // no active thread in the block
    if (vol > TILE_NODES) {                            // direct atomics
        if (!active) return;
        for (int c = 0; c < 8; ++c) {
            const int ci = c >> 2, cj = (c >> 1) & 1, ck = c & 1;
            const size_t node = ((size_t)(st.I[ci] * g.ny + st.J[cj]) * g.nz + st.K[ck]) * 3;
            if (c0 != 0.0f) { const float f = c0 * st.W[c]; for (int k = 0; k < 3; ++k) atomicAdd(g0 + node + k, f * a[k]); }
            if (c1 != 0.0f) { const float f = c1 * st.W[c]; for (int k = 0; k < 3; ++k) atomicAdd(g1 + node + k, f * a[k]); }
        }
        return;
    }
    const int tv = (int)vol * 3;
    for (int e = threadIdx.x; e < 2 * tv; e += blockDim.x) tile[e] = 0.0f;
    __syncthreads();
    if (active) {
        for (int c = 0; c < 8; ++c) {
            const int ci = c >> 2, cj = (c >> 1) & 1, ck = c & 1;
            const int li = ((st.I[ci] - bx[0]) * ex[1] + (st.J[cj] - bx[1])) * ex[2] + (st.K[ck] - bx[2]);
            const float f0 = c0 * st.W[c], f1 = c1 * st.W[c];
            for (int k = 0; k < 3; ++k) {
                atomicAdd(&tile[li * 3 + k], f0 * a[k]);
                atomicAdd(&tile[tv + li * 3 + k], f1 * a[k]);
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * tv; e += blockDim.x) {
        const float v = tile[e];
        if (v == 0.0f) continue;
        const int snap = e >= tv;
        const int r = e - snap * tv;
        const int node = r / 3, comp = r - node * 3;
        const int k = node % ex[2] + bx[2];
        const int t2 = node / ex[2];
        const int j = t2 % ex[1] + bx[1];
        const int i = t2 / ex[1] + bx[0];
        float* gs = snap ? g1 : g0;
        atomicAdd(gs + ((size_t)(i * g.ny + j) * g.nz + k) * 3 + comp, v);
    }
    __syncthreads();                                   // tile is reused by the next call
}

template <int ST, bool ZH>
__global__ void __launch_bounds__(256)
k_adjoint(GridDev g, BcDev bc, const float* __restrict__ ckpt, int n, const int* __restrict__ perm,
          const Blend* __restrict__ tab, int n_steps, int K, int c, float dt, float dth, float dt6,
          float* __restrict__ lam_io, float* __restrict__ seg, float* __restrict__ gU, float* __restrict__ gx0) {
    __shared__ float tile[2 * TILE_NODES * 3];
    __shared__ int sred[4 * 6];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = t < n;                           // every thread stays: block_scatter synchronises
    const int pid = live ? (perm ? perm[t] : t) : 0;
    float lam[3] = {lam_io[3 * (size_t)pid], lam_io[3 * (size_t)pid + 1], lam_io[3 * (size_t)pid + 2]};
    auto f = [&](const Blend& b, const float q[3], float out[3]) { sample<ST, ZH>(g, b, q[0], q[1], q[2], out); };
    const int sa = c * K;
    const int sbnd = min(sa + K, n_steps);
    const int slot = live ? t : 0;
    float x[3] = {ckpt[((size_t)c * n + pid) * 3], ckpt[((size_t)c * n + pid) * 3 + 1],
                  ckpt[((size_t)c * n + pid) * 3 + 2]};
    for (int s = sa; s < sbnd; ++s) {                   // recompute the segment
        if (live) {
            float* o = seg + ((size_t)(s - sa) * n + slot) * 3;
            o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
        }
        rk_step<JT_RK4>(f, tab, s, x, dt, dth, dt6);
        apply_bc(bc, x);
    }
    for (int s = sbnd - 1; s >= sa; --s) {              // reverse it
        float p[3];
        if (live) {
            const float* xi = seg + ((size_t)(s - sa) * n + slot) * 3;
            p[0] = xi[0]; p[1] = xi[1]; p[2] = xi[2];
        } else {
            p[0] = x[0]; p[1] = x[1]; p[2] = x[2];
        }
        const Blend bl0 = tab[3 * s], bl1 = tab[3 * s + 1], bl2 = tab[3 * s + 2];
        float k1[3], k2[3], k3[3], k4[3], x2[3], x3[3], x4[3], xp[3];
        f(bl0, p, k1);
        for (int a = 0; a < 3; ++a) x2[a] = p[a] + dth * k1[a];
        f(bl1, x2, k2);
        for (int a = 0; a < 3; ++a) x3[a] = p[a] + dth * k2[a];
        f(bl1, x3, k3);
        for (int a = 0; a < 3; ++a) x4[a] = p[a] + dt * k3[a];
        f(bl2, x4, k4);
        for (int a = 0; a < 3; ++a) xp[a] = p[a] + dt6 * (((k1[a] + 2.0f * k2[a]) + 2.0f * k3[a]) + k4[a]);
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
        block_scatter(g, bl2, stc, a4, live && dep, gU, tile, sred);
        for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a3[a] += dt * l[a]; }
        vjp_local<ST, ZH>(g, bl1, x3, a3, l, stc, dep);
        block_scatter(g, bl1, stc, a3, live && dep, gU, tile, sred);
        for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a2[a] += dth * l[a]; }
        vjp_local<ST, ZH>(g, bl1, x2, a2, l, stc, dep);
        block_scatter(g, bl1, stc, a2, live && dep, gU, tile, sred);
        for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a1[a] += dth * l[a]; }
        vjp_local<ST, ZH>(g, bl0, p, a1, l, stc, dep);
        block_scatter(g, bl0, stc, a1, live && dep, gU, tile, sred);
        for (int a = 0; a < 3; ++a) lam[a] += l[a];
    }
    if (!live) return;
    lam_io[3 * (size_t)pid] = lam[0];
    lam_io[3 * (size_t)pid + 1] = lam[1];
    lam_io[3 * (size_t)pid + 2] = lam[2];
    if (c == 0 && gx0 != nullptr) {
        gx0[3 * (size_t)pid] = lam[0];
        gx0[3 * (size_t)pid + 1] = lam[1];
        gx0[3 * (size_t)pid + 2] = lam[2];
    }
}

template <bool ZH>
static void launch_adj(int st, dim3 gr, dim3 bl, hipStream_t s, const GridDev& g, const BcDev& bc,
                       const float* ckpt, int n, const int* perm, const Blend* tab, int n_steps, int K, int c,
                       float dt, float dth, float dt6, float* lam, float* seg, float* gU, float* gx0) {
    switch (st) {
    case ST_F32x3: hipLaunchKernelGGL((k_adjoint<ST_F32x3, ZH>), gr, bl, 0, s, g, bc, ckpt, n, perm, tab, n_steps, K, c, dt, dth, dt6, lam, seg, gU, gx0); break;
    case ST_F32x4: hipLaunchKernelGGL((k_adjoint<ST_F32x4, ZH>), gr, bl, 0, s, g, bc, ckpt, n, perm, tab, n_steps, K, c, dt, dth, dt6, lam, seg, gU, gx0); break;
    case ST_BF16x4: hipLaunchKernelGGL((k_adjoint<ST_BF16x4, ZH>), gr, bl, 0, s, g, bc, ckpt, n, perm, tab, n_steps, K, c, dt, dth, dt6, lam, seg, gU, gx0); break;
    default: hipLaunchKernelGGL((k_adjoint<ST_F16x4, ZH>), gr, bl, 0, s, g, bc, ckpt, n, perm, tab, n_steps, K, c, dt, dth, dt6, lam, seg, gU, gx0); break;
    }
}

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
    int maxn = g.nx > g.ny ? g.nx : g.ny;
    maxn = maxn > g.nz ? maxn : g.nz;
    int shift = 0;
    while ((maxn - 1) >> shift >= 1024) ++shift;
    dim3 block(256), grid((unsigned)((n + 255) / 256));
    const bool zh = g.nzs != g.nz;
    for (int c = n_seg - 1; c >= 0; --c) {
        // order this segment's particles by the cell of their checkpoint
        hipLaunchKernelGGL(k_adj_keys, grid, block, 0, s, g, ckpt + (size_t)c * n * 3, (int)n, shift, keys, idx);
        JT_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, temp, keys, skeys, idx, perm, (int)n, 0, 30, s));
        if (zh) launch_adj<true>(st, grid, block, s, g, d, ckpt, (int)n, perm, tab, n_steps, ckpt_every, c, dt, dth, dt6, lam, seg, grad_U, grad_x0);
        else launch_adj<false>(st, grid, block, s, g, d, ckpt, (int)n, perm, tab, n_steps, ckpt_every, c, dt, dth, dt6, lam, seg, grad_U, grad_x0);
        JT_HIP(hipGetLastError());
    }
    return JT_OK;
}
