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

// Vector-Jacobian product of u(q) for adjoint a: returns J^T a in lam and adds
// the snapshot gradient into gU (layout (T, nx, ny, nz, 3), f32).
template <int ST, bool ZH>
__device__ __forceinline__ void vjp(const GridDev& g, const Blend& bl, const float q[3], const float a[3],
                                    float lam[3], float* __restrict__ gU) {
    int I[2], J[2], K[2];
    float wx[2], wy[2], wz[2], dx[2], dy[2], dz[2];
    bool valid = true;
    axis_w(g, 0, q[0], g.nx, I[0], I[1], wx, dx, valid);
    axis_w(g, 1, q[1], g.ny, J[0], J[1], wy, dy, valid);
    axis_w(g, 2, q[2], g.nz, K[0], K[1], wz, dz, valid);
    lam[0] = lam[1] = lam[2] = 0.0f;
    if (!valid && g.bmode == JT_BOUNDS_ZERO) return;   // output is 0 there: no dependence
    const char* base = (const char*)g.data;
    const size_t sb = (size_t)g.snap_stride * ElemBytes<ST>::v;
    const void* b0 = base + (size_t)bl.i0 * sb;
    const void* b1 = base + (size_t)bl.i1 * sb;
    const float oma = 1.0f - bl.a;
    float c0, c1;
    blend_coefs(bl, c0, c1);
    const size_t gsnap = (size_t)g.nx * g.ny * g.nz * 3;
    float* g0 = gU + (size_t)bl.i0 * gsnap;
    float* g1 = gU + (size_t)bl.i1 * gsnap;
#pragma unroll
    for (int ci = 0; ci < 2; ++ci)
#pragma unroll
        for (int cj = 0; cj < 2; ++cj)
#pragma unroll
            for (int ck = 0; ck < 2; ++ck) {
                float v[3];
                corner<ST>(b0, b1, bl, oma, (I[ci] * g.ny + J[cj]) * g.nzs + K[ck], v);
                const float s = a[0] * v[0] + a[1] * v[1] + a[2] * v[2];
                const float W = wx[ci] * wy[cj] * wz[ck];
                lam[0] += s * dx[ci] * wy[cj] * wz[ck];
                lam[1] += s * wx[ci] * dy[cj] * wz[ck];
                lam[2] += s * wx[ci] * wy[cj] * dz[ck];
                const size_t node = ((size_t)(I[ci] * g.ny + J[cj]) * g.nz + K[ck]) * 3;
                if (c0 != 0.0f) {
                    const float f = c0 * W;
                    atomicAdd(g0 + node, f * a[0]);
                    atomicAdd(g0 + node + 1, f * a[1]);
                    atomicAdd(g0 + node + 2, f * a[2]);
                }
                if (c1 != 0.0f) {
                    const float f = c1 * W;
                    atomicAdd(g1 + node, f * a[0]);
                    atomicAdd(g1 + node + 1, f * a[1]);
                    atomicAdd(g1 + node + 2, f * a[2]);
                }
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

template <int ST, bool ZH>
__global__ void __launch_bounds__(256)
k_adjoint(GridDev g, BcDev bc, const float* __restrict__ ckpt, int n, const Blend* __restrict__ tab,
          int n_steps, int K, int c_hi, int c_lo, float dt, float dth, float dt6, float* __restrict__ lam_io,
          float* __restrict__ seg, float* __restrict__ gU, float* __restrict__ gx0) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float lam[3] = {lam_io[3 * (size_t)i], lam_io[3 * (size_t)i + 1], lam_io[3 * (size_t)i + 2]};
    auto f = [&](const Blend& b, const float q[3], float out[3]) { sample<ST, ZH>(g, b, q[0], q[1], q[2], out); };
    for (int c = c_hi; c >= c_lo; --c) {
        const int sa = c * K;
        const int sbnd = min(sa + K, n_steps);
        float x[3] = {ckpt[((size_t)c * n + i) * 3], ckpt[((size_t)c * n + i) * 3 + 1],
                      ckpt[((size_t)c * n + i) * 3 + 2]};
        for (int s = sa; s < sbnd; ++s) {                   // recompute the segment
            float* o = seg + ((size_t)(s - sa) * n + i) * 3;
            o[0] = x[0]; o[1] = x[1]; o[2] = x[2];
            rk_step<JT_RK4>(f, tab, s, x, dt, dth, dt6);
            apply_bc(bc, x);
        }
        for (int s = sbnd - 1; s >= sa; --s) {              // reverse it
            const float* xi = seg + ((size_t)(s - sa) * n + i) * 3;
            float p[3] = {xi[0], xi[1], xi[2]};
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
            vjp<ST, ZH>(g, bl2, x4, a4, l, gU);
            for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a3[a] += dt * l[a]; }
            vjp<ST, ZH>(g, bl1, x3, a3, l, gU);
            for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a2[a] += dth * l[a]; }
            vjp<ST, ZH>(g, bl1, x2, a2, l, gU);
            for (int a = 0; a < 3; ++a) { lam[a] += l[a]; a1[a] += dth * l[a]; }
            vjp<ST, ZH>(g, bl0, p, a1, l, gU);
            for (int a = 0; a < 3; ++a) lam[a] += l[a];
        }
    }
    lam_io[3 * (size_t)i] = lam[0];
    lam_io[3 * (size_t)i + 1] = lam[1];
    lam_io[3 * (size_t)i + 2] = lam[2];
    if (c_lo == 0 && gx0 != nullptr) {
        gx0[3 * (size_t)i] = lam[0];
        gx0[3 * (size_t)i + 1] = lam[1];
        gx0[3 * (size_t)i + 2] = lam[2];
    }
}

template <bool ZH>
static void launch_adj(int st, dim3 gr, dim3 bl, hipStream_t s, const GridDev& g, const BcDev& bc,
                       const float* ckpt, int n, const Blend* tab, int n_steps, int K, int c_hi, int c_lo,
                       float dt, float dth, float dt6, float* lam, float* seg, float* gU, float* gx0) {
    switch (st) {
    case ST_F32x3: hipLaunchKernelGGL((k_adjoint<ST_F32x3, ZH>), gr, bl, 0, s, g, bc, ckpt, n, tab, n_steps, K, c_hi, c_lo, dt, dth, dt6, lam, seg, gU, gx0); break;
    case ST_F32x4: hipLaunchKernelGGL((k_adjoint<ST_F32x4, ZH>), gr, bl, 0, s, g, bc, ckpt, n, tab, n_steps, K, c_hi, c_lo, dt, dth, dt6, lam, seg, gU, gx0); break;
    case ST_BF16x4: hipLaunchKernelGGL((k_adjoint<ST_BF16x4, ZH>), gr, bl, 0, s, g, bc, ckpt, n, tab, n_steps, K, c_hi, c_lo, dt, dth, dt6, lam, seg, gU, gx0); break;
    default: hipLaunchKernelGGL((k_adjoint<ST_F16x4, ZH>), gr, bl, 0, s, g, bc, ckpt, n, tab, n_steps, K, c_hi, c_lo, dt, dth, dt6, lam, seg, gU, gx0); break;
    }
}

}  // namespace jt

using namespace jt;

extern "C" size_t jt_track_grid_adjoint_workspace_size(int64_t n, int32_t n_steps, int32_t ckpt_every) {
    Arena a(nullptr, 0);
    a.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    a.take<float>((size_t)3 * (n > 0 ? n : 1));                                   // adjoint state
    a.take<float>((size_t)3 * (n > 0 ? n : 1) * (ckpt_every > 0 ? ckpt_every : 1)); // segment states
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
    if (!tab || !lam || !seg || !arena.ok())
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
    // a few segments per launch keeps each launch short at the largest sizes
    const int seg_per_launch = n >= (1 << 20) ? 2 : n_seg;
    dim3 block(256), grid((unsigned)((n + 255) / 256));
    const bool zh = g.nzs != g.nz;
    for (int c_hi = n_seg - 1; c_hi >= 0; c_hi -= seg_per_launch) {
        int c_lo = c_hi - seg_per_launch + 1;
        c_lo = c_lo < 0 ? 0 : c_lo;
        if (zh) launch_adj<true>(st, grid, block, s, g, d, ckpt, (int)n, tab, n_steps, ckpt_every, c_hi, c_lo, dt, dth, dt6, lam, seg, grad_U, grad_x0);
        else launch_adj<false>(st, grid, block, s, g, d, ckpt, (int)n, tab, n_steps, ckpt_every, c_hi, c_lo, dt, dth, dt6, lam, seg, grad_U, grad_x0);
        JT_HIP(hipGetLastError());
    }
    return JT_OK;
}
