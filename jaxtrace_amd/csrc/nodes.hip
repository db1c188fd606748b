// Nearest-node sampling of a TimeSeriesField (jaxtrace/fields/time_series.py:
// 249-259, 391-398): u(q, t) = blend_t(V)[argmin_n |q - p_n|^2], ties to the
// lowest node index, with |.|^2 = (dx*dx + dy*dy) + dz*dz in float32.
//
// Exhaustive, as the reference: each thread owns one query and scans all
// nodes, which stream through LDS in tiles of 256 shared by the block's 256
// queries.  The blend touches only the chosen node's two snapshot values.
#include "grid_sampler.h"

namespace jt {

struct NodeDev {
    const float* pos;       // (N, 3)
    int n;
    const float* data;      // (T, N, 3)
    long long snap_stride;  // N * 3
    const float* times;
    int T, interp, extrap;
};

// Argmin over all nodes for the 256 queries of a block (q may be a dummy for
// lanes past the end; every thread must call this: it synchronises).
__device__ inline int nn_block(const NodeDev& d, const float q[3], float* sp) {
    const int tid = threadIdx.x;
    float best = __int_as_float(0x7f800000);
    int bi = 0;
    bool any = false;
    for (int base = 0; base < d.n; base += 256) {
        const int j = base + tid;
        if (j < d.n) {
            sp[tid * 3] = d.pos[(size_t)j * 3];
            sp[tid * 3 + 1] = d.pos[(size_t)j * 3 + 1];
            sp[tid * 3 + 2] = d.pos[(size_t)j * 3 + 2];
        }
        __syncthreads();
        const int cnt = min(256, d.n - base);
        for (int k = 0; k < cnt; ++k) {
            float dx = q[0] - sp[k * 3], dy = q[1] - sp[k * 3 + 1], dz = q[2] - sp[k * 3 + 2];
            float d2 = (dx * dx + dy * dy) + dz * dz;
            // first minimum; a NaN distance is the minimum (jnp.argmin treats NaN as smallest)
            bool take = !any || (d2 < best) || (d2 != d2 && best == best);
            if (take) { best = d2; bi = base + k; any = true; }
        }
        __syncthreads();
    }
    return bi;
}

__global__ void __launch_bounds__(256)
k_nn_sample(NodeDev d, const float* __restrict__ Q, int m, float t, float* __restrict__ out, int* __restrict__ idx) {
    __shared__ float sp[256 * 3];
    const int i = blockIdx.x * 256 + threadIdx.x;
    float q[3] = {0.f, 0.f, 0.f};
    if (i < m) { q[0] = Q[(size_t)i * 3]; q[1] = Q[(size_t)i * 3 + 1]; q[2] = Q[(size_t)i * 3 + 2]; }
    const int k = nn_block(d, q, sp);
    if (i >= m) return;
    const Blend b = time_blend(d.times, d.T, t, d.interp, d.extrap);
    float v[3];
    corner<ST_F32x3>(d.data + (size_t)b.i0 * d.snap_stride, d.data + (size_t)b.i1 * d.snap_stride, b,
                     1.0f - b.a, k, v);
    out[(size_t)i * 3] = v[0];
    out[(size_t)i * 3 + 1] = v[1];
    out[(size_t)i * 3 + 2] = v[2];
    if (idx) idx[i] = k;
}

template <int INTEG>
__global__ void __launch_bounds__(256)
k_nn_track(NodeDev d, BcDev bc, float* __restrict__ xs, int n, const Blend* __restrict__ tab, int s0, int s1,
           float dt, float dth, float dt6, int rec_every, float* __restrict__ traj) {
    __shared__ float sp[256 * 3];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;                           // all threads stay: nn_block synchronises
    float p[3] = {0.f, 0.f, 0.f};
    if (live) { p[0] = xs[3 * (size_t)i]; p[1] = xs[3 * (size_t)i + 1]; p[2] = xs[3 * (size_t)i + 2]; }
    auto f = [&](const Blend& b, const float q[3], float out[3]) {
        const int k = nn_block(d, q, sp);
        corner<ST_F32x3>(d.data + (size_t)b.i0 * d.snap_stride, d.data + (size_t)b.i1 * d.snap_stride, b,
                         1.0f - b.a, k, out);
    };
    for (int s = s0; s < s1; ++s) {
        rk_step<INTEG>(f, tab, s, p, dt, dth, dt6);
        apply_bc(bc, p);
        if (live && traj != nullptr && (s + 1) % rec_every == 0) {
            float* o = traj + ((size_t)((s + 1) / rec_every) * (size_t)n + i) * 3;
            o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
        }
    }
    if (live) { xs[3 * (size_t)i] = p[0]; xs[3 * (size_t)i + 1] = p[1]; xs[3 * (size_t)i + 2] = p[2]; }
}

static int make_nodes(const jt_nodes_field* f, NodeDev& d) {
    JT_REQUIRE(f != nullptr, "nodes field descriptor is NULL");
    JT_REQUIRE(f->node_pos && f->data && f->times, "NULL pointer in nodes field");
    JT_REQUIRE(f->n_nodes >= 1 && f->n_nodes < (1LL << 31) && f->n_snap >= 1, "bad nodes field sizes");
    JT_REQUIRE(f->time_interp == JT_TIME_LINEAR || f->time_interp == JT_TIME_NEAREST, "bad time_interp");
    JT_REQUIRE(f->time_extrap >= JT_EXTRAP_CONSTANT && f->time_extrap <= JT_EXTRAP_ZERO, "bad time_extrap");
    d.pos = f->node_pos; d.n = (int)f->n_nodes; d.data = (const float*)f->data; d.snap_stride = f->n_nodes * 3;
    d.times = f->times; d.T = f->n_snap; d.interp = f->time_interp; d.extrap = f->time_extrap;
    return JT_OK;
}

}  // namespace jt

using namespace jt;

extern "C" size_t jt_track_nodes_workspace_size(int64_t n, int32_t n_steps) {
    (void)n;
    Arena a(nullptr, 0);
    a.take<Blend>((size_t)3 * (n_steps > 0 ? n_steps : 1));
    return a.used + 256;
}

extern "C" int jt_track_nodes(const jt_nodes_field* field, const jt_boundary* bc, int32_t integrator, float* x,
                              int64_t n, const float* step_times, float dt, int32_t n_steps, int32_t record_every,
                              float* traj, const jt_track_opts* opts, void* ws, size_t ws_bytes, void* stream) {
    NodeDev nd;
    int rc = make_nodes(field, nd);
    if (rc) return rc;
    BcDev d;
    rc = make_bc_dev(bc, d);
    if (rc) return rc;
    JT_REQUIRE(integrator >= JT_EULER && integrator <= JT_RK4, "bad integrator %d", integrator);
    JT_REQUIRE(n >= 0 && n < (1LL << 31), "n out of range");
    JT_REQUIRE(n_steps >= 0 && record_every >= 1, "bad n_steps / record_every");
    if (d.inlet == JT_INLET_HOST)
        return fail(JT_EUNSUPPORTED, "a host-drawn inlet boundary is not fused into tracking; step with an open "
                                     "boundary and apply it with jt_inlet_exits/jt_inlet_place");
    JT_REQUIRE(n == 0 || x != nullptr, "NULL x");
    JT_REQUIRE(n_steps == 0 || step_times != nullptr, "NULL step_times");
    hipStream_t s = (hipStream_t)stream;
    if (traj != nullptr && n > 0)
        JT_HIP(hipMemcpyAsync(traj, x, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToDevice, s));
    if (n == 0 || n_steps == 0) return JT_OK;
    Arena arena(ws, ws_bytes);
    Blend* tab = arena.take<Blend>((size_t)3 * n_steps);
    if (!tab || !arena.ok()) return fail(JT_ENOMEM, "workspace too small");
    GridDev g{};
    g.times = nd.times; g.T = nd.T; g.interp = nd.interp; g.extrap = nd.extrap;
    hipLaunchKernelGGL(k_blend_table, dim3((n_steps + 127) / 128), dim3(128), 0, s, g, step_times, n_steps, dt,
                       integrator, tab);
    JT_HIP(hipGetLastError());
    int spl = (opts && opts->steps_per_launch > 0) ? opts->steps_per_launch : 10;
    const float dth = 0.5f * dt, dt6 = dt / 6.0f;
    dim3 block(256), grid((unsigned)((n + 255) / 256));
    for (int s0 = 0; s0 < n_steps; s0 += spl) {
        int s1 = s0 + spl < n_steps ? s0 + spl : n_steps;
        if (integrator == JT_RK4)
            hipLaunchKernelGGL(k_nn_track<JT_RK4>, grid, block, 0, s, nd, d, x, (int)n, tab, s0, s1, dt, dth, dt6, record_every, traj);
        else if (integrator == JT_RK2)
            hipLaunchKernelGGL(k_nn_track<JT_RK2>, grid, block, 0, s, nd, d, x, (int)n, tab, s0, s1, dt, dth, dt6, record_every, traj);
        else
            hipLaunchKernelGGL(k_nn_track<JT_EULER>, grid, block, 0, s, nd, d, x, (int)n, tab, s0, s1, dt, dth, dt6, record_every, traj);
        JT_HIP(hipGetLastError());
    }
    return JT_OK;
}

extern "C" int jt_nodes_sample(const jt_nodes_field* field, const float* Q, int64_t m, float t, float* out,
                               int32_t* nearest, void* stream) {
    NodeDev d;
    int rc = make_nodes(field, d);
    if (rc) return rc;
    JT_REQUIRE(m >= 0 && m < (1LL << 31), "m out of range");
    if (m == 0) return JT_OK;
    JT_REQUIRE(Q && out, "NULL pointer");
    hipLaunchKernelGGL(k_nn_sample, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, (hipStream_t)stream, d, Q,
                       (int)m, t, out, nearest);
    JT_HIP(hipGetLastError());
    return JT_OK;
}
