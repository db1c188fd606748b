/*
 * jaxtrace_amd.h -- C ABI of the MI355X (gfx950) particle-advection and
 * density library.
 *
 * The reference JAXTrace is a pure-Python/JAX package with no FFI: its
 * boundary is a Python object protocol.  Each entry point below replaces one
 * XLA program the reference builds under jax.jit, and is bound from Python
 * with ctypes by the package `jaxtrace_amd` (the host-side mirror of the
 * reference interface).  Citations are relative to the reference repository.
 *
 * Conventions
 *   - every function returns a status (JT_OK == 0); on failure
 *     jt_last_error() returns a message for the calling thread;
 *   - all array arguments are caller-owned DEVICE pointers unless marked
 *     "host"; the library never frees or keeps them after the call returns;
 *   - `stream` is a hipStream_t (NULL = default stream); every call is
 *     asynchronous on it;
 *   - scratch memory comes from a caller-provided workspace sized by the
 *     matching *_workspace_size() query; no entry point calls hipMalloc;
 *   - positions are (N,3) float32, row-major (x,y,z per particle).
 */
#ifndef JAXTRACE_AMD_H
#define JAXTRACE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------- */
#define JT_OK            0
#define JT_EINVAL        1   /* bad argument            -> ValueError    */
#define JT_ENOMEM        2   /* workspace too small     -> MemoryError   */
#define JT_EHIP          3   /* HIP runtime error       -> RuntimeError  */
#define JT_EUNSUPPORTED  4   /* valid but not supported -> TypeError     */

/* ---- enums -------------------------------------------------------------- */
/* integrators: jaxtrace/integrators/{euler,rk2,rk4}.py */
#define JT_EULER 0
#define JT_RK2   1
#define JT_RK4   2

/* snapshot storage types (arithmetic is always float32) */
#define JT_DTYPE_F32  0
#define JT_DTYPE_BF16 1
#define JT_DTYPE_F16  2

/* spatial bounds handling: jaxtrace/fields/structured.py:165 */
#define JT_BOUNDS_CLAMP    0
#define JT_BOUNDS_ZERO     1
#define JT_BOUNDS_NAN      2
#define JT_BOUNDS_PERIODIC 3

/* temporal interpolation / extrapolation: jaxtrace/fields/time_series.py:149-153 */
#define JT_TIME_LINEAR  0
#define JT_TIME_NEAREST 1
#define JT_EXTRAP_CONSTANT 0
#define JT_EXTRAP_LINEAR   1
#define JT_EXTRAP_NAN      2
#define JT_EXTRAP_ZERO     3

/* per-axis boundary kinds: jaxtrace/tracking/boundary.py:101-210, 377-438 */
#define JT_BC_NONE       0
#define JT_BC_PERIODIC   1
#define JT_BC_REFLECTIVE 2
#define JT_BC_CLAMP      3
#define JT_BC_ABSORB     4

/* continuous-inlet replacement: jaxtrace/tracking/boundary.py:1444-1601 */
#define JT_INLET_OFF  0
#define JT_INLET_GRID 1   /* nearest inlet-grid node (:1249-1318)          */
#define JT_INLET_ZERO 2   /* 'grid' without concentrations: off-axis = 0    */

/* ---- descriptors (host structs, passed by pointer) ---------------------- */

/* Time series of snapshots on a uniform structured grid: the composition of
 * TimeSeriesField's temporal blend (time_series.py:155-245) with
 * StructuredGridSampler's trilinear interpolation (structured.py:167-250). */
typedef struct jt_grid_field {
    const void*  data;        /* device (n_snap, nx, ny, nz, ncomp), C order   */
    const float* times;       /* device (n_snap,) f32, non-decreasing          */
    int32_t n_snap;
    int32_t nx, ny, nz;
    int32_t dtype;            /* JT_DTYPE_*                                    */
    int32_t ncomp;            /* 3 (packed xyz) or 4 (xyz + pad)               */
    float   origin[3];        /* f32 grid_x[0], grid_y[0], grid_z[0]           */
    float   spacing[3];       /* f32 grid[1]-grid[0] (1 when n == 1)           */
    int32_t bounds_mode;      /* JT_BOUNDS_*                                   */
    int32_t time_interp;      /* JT_TIME_*                                     */
    int32_t time_extrap;      /* JT_EXTRAP_*                                   */
} jt_grid_field;

/* Boundary applied after every full step (tracker.py:238-239). */
typedef struct jt_boundary {
    int32_t axis_kind[3];     /* JT_BC_* per axis (mixed_boundary, :398-436)   */
    float   lo[3], hi[3];     /* f32 bounds (_ensure_bounds_shape, :39-86)     */
    int32_t absorb_rows;      /* 1: absorbing_boundary (:190-210), whole row   */
    float   outside_value;    /* value written by absorbing conditions         */
    int32_t inlet;            /* JT_INLET_*                                    */
    int32_t inlet_axis;       /* 0,1,2                                         */
    int32_t inlet_positive;   /* 1: enter at lo, exit past hi                  */
    float   inlet_backstop;   /* f32 lo + 0.001*(hi-lo) (or hi - ...), :1583   */
    const float* inlet_coords[2]; /* device linspace of the two other axes   */
    int32_t inlet_ncoords[2];
} jt_boundary;

/* Optional tracking knobs (NULL = defaults). */
typedef struct jt_track_opts {
    int32_t steps_per_launch; /* 0 = library default                           */
    int32_t sort_every;       /* re-sort particles by cell every k steps; 0 = never */
    int32_t block;            /* threads per block; 0 = default (256)          */
} jt_track_opts;

/* ---- library ------------------------------------------------------------- */
const char* jt_last_error(void);
int32_t     jt_version(void);

/* ---- structured tracking (hot path) --------------------------------------- */

/* Replaces one field_fn call: TimeSeriesField._sample_all_nodes_at_time_jax
 * (time_series.py:155-245) composed with StructuredGridSampler._trilinear_jax
 * (structured.py:167-250).  out[i] = u(x[i], t). */
int jt_grid_sample(const jt_grid_field* field, const float* x, int64_t n,
                   float t, float* out, void* stream);

/* Replaces the boundary closures of boundary.py (periodic/reflective/
 * clamping/absorbing factories :215-337, mixed_boundary :377-438,
 * continuous_inlet_boundary_factory :1444-1601).  In place on x. */
int jt_apply_boundary(const jt_boundary* bc, float* x, int64_t n, void* stream);

size_t jt_track_grid_workspace_size(int64_t n, int32_t n_steps,
                                    const jt_track_opts* opts);

/* Replaces ParticleTracker._compiled_simulate (tracker.py:228-246): n_steps
 * steps of the integrator from stage-origin times step_times[s] (the
 * reference's times[:-1]) with constant dt, boundary after each step.
 * x (n,3) holds x0 on entry and the final positions on return.  If traj is
 * not NULL it receives (1 + n_steps/record_every, n, 3) positions: x0 first,
 * then the state after every record_every-th step (original particle order). */
int jt_track_grid(const jt_grid_field* field, const jt_boundary* bc,
                  int32_t integrator, float* x, int64_t n,
                  const float* step_times, float dt, int32_t n_steps,
                  int32_t record_every, float* traj,
                  const jt_track_opts* opts, void* ws, size_t ws_bytes,
                  void* stream);

#ifdef __cplusplus
}
#endif
#endif /* JAXTRACE_AMD_H */
