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
#define JT_INLET_HOST 3   /* 'uniform'/'gaussian': the off-axis coordinates of
                             the exited particles are drawn on the host
                             (np.random, :1555-1571) and placed by
                             jt_inlet_place; not fused into the tracking
                             kernels, which reject it                         */

/* ---- descriptors (host structs, passed by pointer) ---------------------- */

/* Time series of snapshots on a uniform structured grid: the composition of
 * TimeSeriesField's temporal blend (time_series.py:155-245) with
 * StructuredGridSampler's trilinear interpolation (structured.py:167-250). */
typedef struct jt_grid_field {
    const void*  data;        /* device (n_snap, nx, ny, nz(+1), ncomp), C order */
    const float* times;       /* device (n_snap,) f32, non-decreasing          */
    int32_t n_snap;
    int32_t nx, ny, nz;
    int32_t dtype;            /* JT_DTYPE_*                                    */
    int32_t ncomp;            /* 3 (packed xyz) or 4: float32 (x, y, z, pad);
                                 16-bit (x, y, 0, z), the third half zero      */
    int32_t z_halo;           /* 1: each snapshot holds nz+1 z-planes; plane nz
                                 repeats plane 0 (periodic bounds) or plane
                                 nz-1 (other bounds), so the two z-corners of a
                                 cell are one contiguous load                  */
    float   origin[3];        /* f32 grid_x[0], grid_y[0], grid_z[0]           */
    float   spacing[3];       /* f32 grid[1]-grid[0] (1 when n == 1)           */
    int32_t bounds_mode;      /* JT_BOUNDS_*                                   */
    int32_t time_interp;      /* JT_TIME_*                                     */
    int32_t time_extrap;      /* JT_EXTRAP_*                                   */
    const float* pair_data;   /* optional, float32 fields with ncomp 4 and a z
                                 halo: consecutive-pair records, C order
                                 (n_snap-1, nx, ny, nz, 12): for pair (s, s+1)
                                 and node (i, j, k) the xyz of V_s(k),
                                 V_s(k+1), V_s+1(k), V_s+1(k+1) (k+1 the halo
                                 plane's node for k = nz-1), as
                                 jt_grid_pair_records writes them.  Tracking
                                 steps that blend (s, s+1) gather a cell row
                                 of both snapshots in one 48-byte record.
                                 NULL: not used.                               */
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
    int32_t n_launch_events;  /* size of launch_events (may be 0)              */
    void**  launch_events;    /* optional hipEvent_t array: events[2j] and
                                 events[2j+1] are recorded on the stream right
                                 before and after the j-th tracking-kernel
                                 launch (profiling hook; NULL = none)          */
} jt_track_opts;

/* ---- library ------------------------------------------------------------- */
const char* jt_last_error(void);
int32_t     jt_version(void);

/* ---- structured tracking (hot path) --------------------------------------- */

/* Replaces one field_fn call: TimeSeriesField._sample_all_nodes_at_time_jax
 * (time_series.py:155-245) composed with StructuredGridSampler._trilinear_jax
 * (structured.py:167-250).  out[i] = u(x[i], t). */
/* The consecutive-pair records of field (float32, ncomp 4, z halo) into
 * out, a device array of (n_snap-1) * nx * ny * nz * 12 floats (the
 * pair_data layout above; field->pair_data is ignored).  n_snap >= 2. */
int jt_grid_pair_records(const jt_grid_field* field, float* out, void* stream);

int jt_grid_sample(const jt_grid_field* field, const float* x, int64_t n,
                   float t, float* out, void* stream);

/* Replaces the boundary closures of boundary.py (periodic/reflective/
 * clamping/absorbing factories :215-337, mixed_boundary :377-438,
 * continuous_inlet_boundary_factory :1444-1601).  In place on x. */
int jt_apply_boundary(const jt_boundary* bc, float* x, int64_t n, void* stream);

/* Continuous inlet with host-drawn replacements (JT_INLET_HOST), split at
 * the point where continuous_inlet_boundary_factory's closure draws from
 * np.random (boundary.py:1524-1571).  jt_inlet_exits marks the particles
 * that left through the outlet face (:1526-1537) and gives each its rank
 * among them in particle order (rank[i], defined where x[i] exited) and the
 * count (*n_exited, device int32).  The host then draws the replacement
 * coordinates in the reference's order: n_exited values for the lower
 * off-axis, then n_exited for the higher one.  jt_inlet_place puts exited
 * particle i at the inlet face with draws[rank[i]] and
 * draws[n_exited + rank[i]], then applies the backstop and the off-axis
 * reflection to every particle (:1576-1597).  In place on x. */
size_t jt_inlet_exits_workspace_size(int64_t n);
int jt_inlet_exits(const jt_boundary* bc, const float* x, int64_t n, int32_t* rank, int32_t* n_exited,
                   void* ws, size_t ws_bytes, void* stream);
int jt_inlet_place(const jt_boundary* bc, float* x, int64_t n, const int32_t* rank, const float* draws,
                   int64_t n_exited, void* stream);

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

/* jt_track_grid without a workspace.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The stage-time
 * table lives in a static device array of the library (3 MiB per GPU, kept
 * until the process ends), so:
 *   - n_steps <= 65536 per call (JT_EINVAL beyond: split the run);
 *   - no cell sorting (opts->sort_every must be 0, JT_EUNSUPPORTED otherwise);
 *   - calls that overlap on one GPU (several streams or threads) race on that
 *     array: serialise them.
 * Same arguments, results and recording as jt_track_grid with sort_every = 0;
 * the call makes no allocation (no hipMalloc/hipFree, no allocating library
 * routine). */
int jt_track_grid_noalloc(const jt_grid_field* field, const jt_boundary* bc, int32_t integrator, float* x,
                          int64_t n, const float* step_times, float dt, int32_t n_steps, int32_t record_every,
                          float* traj, const jt_track_opts* opts, void* stream);

/* ---- hash-grid neighbours and SPH density ------------------------------- */
/* cell-key packing: neighbors.py:83-91.  The reference never enables
 * jax_enable_x64, so its int64 keys are int32 and `cx << 42` is 0: the
 * as-shipped 3-D key is (cy<<21)|cz (JT_KEY_X32).  JT_KEY_X64 is the
 * intended (cx<<42)|(cy<<21)|cz. */
#define JT_KEY_X32 0
#define JT_KEY_X64 1

/* SPH kernels: sph.py:227-262 */
#define JT_SPH_CUBIC    0
#define JT_SPH_WENDLAND 1
#define JT_SPH_GAUSSIAN 2

/* CSR hash grid of n points (HashGridNeighbors, neighbors.py:39-103).
 * origin = per-axis min of the points, inv = f32(1/max(cell_size,1e-12)),
 * dims = floor((max-origin)*inv)+1 per axis.  Cells are dense ids:
 * X64: (cx*dims[1]+cy)*dims[2]+cz; X32: cy*dims[2]+cz (2-D: cx*dims[1]+cy
 * and cy).  The caller allocates the arrays (jt_hashgrid_num_cells). */
typedef struct jt_hashgrid {
    int32_t dim;              /* 2 or 3                                        */
    int32_t key_mode;         /* JT_KEY_*                                      */
    float   origin[3];
    float   inv;
    int32_t dims[3];
    int64_t n_points;
    int32_t* order;           /* device (n): stable argsort of keys (pt_indices_sorted) */
    float*   sorted_pos;      /* device (n, dim): positions in cell order      */
    int32_t* cell_start;      /* device (n_cells)                              */
    int32_t* cell_end;        /* device (n_cells), exclusive                   */
} jt_hashgrid;

/* Per-axis min then max of points (n, dim) into out_minmax (2*dim floats,
 * device).  ws: device scratch of at least 24 bytes; each 24 bytes hold one
 * block's partial result, and up to JT_BOUNDS_WS_BYTES are used (a smaller
 * ws runs fewer blocks). */
#define JT_BOUNDS_WS_BYTES 24576
int jt_points_bounds(const float* P, int64_t n, int32_t dim, float* out_minmax,
                     void* ws, size_t ws_bytes, void* stream);
int64_t jt_hashgrid_num_cells(const jt_hashgrid* g);
size_t  jt_hashgrid_workspace_size(int64_t n);
/* Replaces HashGridNeighbors.__post_init__ (neighbors.py:61-103). */
int jt_hashgrid_build(const jt_hashgrid* g, const float* P, void* ws, size_t ws_bytes,
                      void* stream);
/* Replaces HashGridNeighbors.query_many (neighbors.py:109-168): (m, K) int32
 * indices and f32 distances, nearest first, stable in candidate order;
 * invalid slots carry pt_indices_sorted[0] and +inf. */
int jt_hashgrid_query(const jt_hashgrid* g, const float* Q, int64_t m, float radius,
                      int32_t per_cell_cap, int32_t max_neighbors, int32_t* out_idx,
                      float* out_dist, void* stream);
/* No reference counterpart: per query, stats[2*i] = candidates read (sum over
 * the probed cells of min(count, per_cell_cap)) and stats[2*i+1] = those within
 * radius; the input-determined counts of the SPH byte model and KDE pair
 * count, for the benchmark's roofline. */
int jt_hashgrid_query_stats(const jt_hashgrid* h, const float* Q, int64_t m, float radius, int32_t per_cell_cap,
                            int32_t* stats, void* stream);

/* Replaces SPHDensityEstimator._evaluate_density (sph.py:269-279): the sum of
 * m_j W(d_j, h_j) over the query's first max_neighbors neighbours within
 * radius.  masses/smoothing may be NULL (1 and h_mean). */
int jt_sph_eval(const jt_hashgrid* g, const float* Q, int64_t m, float radius,
                int32_t per_cell_cap, int32_t max_neighbors, const float* masses,
                const float* smoothing, float h_mean, int32_t kernel, float* out_rho,
                void* stream);

/* The same densities as jt_sph_eval for x32 keys with uniform masses and
 * smoothing (sph.py:269-279 with masses 1 and h_mean), the reference's
 * default for 3-D evaluation: the queries are ordered by their cell's (y, z)
 * column (2-D: y), whose candidates are the same for every query of it, and
 * a wave stages one column's candidates in LDS for 64 queries.  The terms
 * summed are jt_sph_eval's, in another order.  per_cell_cap <= 64.  ws:
 * device scratch of jt_sph_tiled_workspace_size(g, m) bytes. */
size_t jt_sph_tiled_workspace_size(const jt_hashgrid* g, int64_t m);
int jt_sph_eval_tiled(const jt_hashgrid* g, const float* Q, int64_t m, float radius,
                      int32_t per_cell_cap, int32_t max_neighbors, float h_mean, int32_t kernel,
                      float* out_rho, void* ws, size_t ws_bytes, void* stream);

/* ---- Gaussian KDE --------------------------------------------------------- */
/* Replaces KDEEstimator._evaluate_grid's all-pairs eval_chunk
 * (kde.py:244-270): out[i] = mean_n (2pi)^(-dim/2) exp(-0.5|Q_i-P_n|^2 f32(1/h^2)),
 * exhaustive over all n points (the reference algorithm). */
int jt_kde_eval(const float* P, int64_t n, int32_t dim, const float* Q, int64_t m,
                float h, float* out, void* stream);
/* jt_kde_eval with the particles split over more blocks when the queries
 * are few (the per-chunk sums in float64, then added in chunk order):
 * workspace jt_kde_eval_workspace_size(n, m) bytes (0 when one chunk fills
 * the chip; the call is then jt_kde_eval).  The estimator's evaluate(). */
size_t jt_kde_eval_workspace_size(int64_t n, int64_t m);
int jt_kde_eval_ws(const float* P, int64_t n, int32_t dim, const float* Q, int64_t m, float h, float* out,
                   void* ws, size_t ws_bytes, void* stream);
/* The same density on the structured grid of KDEEstimator.evaluate_3d
 * (kde.py:188-240; out is (ny, nx, nz), 'xy' meshgrid order), summing every
 * point within `cutoff` of each grid point (and some farther ones), found
 * through a 3-D JT_KEY_X64 hash grid of the points whose cell size is
 * >= cutoff / 8.  n_total is the number of points (the mean's divisor).
 * Partial sums are added with float atomics, so the last float32 bits may
 * differ between calls. */
int jt_kde_grid_cutoff(const jt_hashgrid* grid, int64_t n_total, const float* gx, int32_t nx,
                       const float* gy, int32_t ny, const float* gz, int32_t nz, float h,
                       float cutoff, float* out, void* stream);
/* jt_kde_grid_cutoff with a workspace (jt_kde_grid_cutoff_workspace_size
 * bytes): each block's partial densities go to their own slice and the
 * slices of a grid tile are added in a fixed order, so the result is the
 * same on every call and no float atomics are used (the estimator's path). */
size_t jt_kde_grid_cutoff_workspace_size(int32_t nx, int32_t ny, int32_t nz);
int jt_kde_grid_cutoff_ws(const jt_hashgrid* grid, int64_t n_total, const float* gx, int32_t nx,
                          const float* gy, int32_t ny, const float* gz, int32_t nz, float h,
                          float cutoff, float* out, void* ws, size_t ws_bytes, void* stream);
/* KDEEstimator.evaluate_2d's all-pairs density (kde.py:205-221, 244-270):
 * out (ny, nx) [iy][ix] = mean_n (2pi)^-1 exp(-0.5 |(gx[ix], gy[iy]) - P_n|^2
 * f32(1/h^2)) for the n points P (n, 2), every pair summed, on the bf16 matrix
 * cores (separable factors, each split into bf16 hi/lo halves).  Workspace:
 * jt_kde_grid2d_workspace_size(n, nx, ny) bytes (one partial tile per block,
 * summed in a fixed order: the same bits on every call). */
/* KDEEstimator.evaluate_3d's all-pairs density (kde.py:223-240, 244-270):
 * out (ny, nx, nz) = mean_n (2pi)^(-3/2) exp(-0.5 |q - P_n|^2 f32(1/h^2)) for
 * the n points P (n, 3) and the grid (gx, gy, gz) ('xy' meshgrid order),
 * every pair summed, on the bf16 matrix cores (the cut-off kernel's tiles
 * with chunks of the points instead of hash-grid columns).  Workspace:
 * jt_kde_grid3d_workspace_size(n, nx, ny, nz) bytes; the result is the same
 * on every call. */
size_t jt_kde_grid3d_workspace_size(int64_t n, int32_t nx, int32_t ny, int32_t nz);
int jt_kde_grid3d(const float* P, int64_t n, const float* gx, int32_t nx, const float* gy, int32_t ny,
                  const float* gz, int32_t nz, float h, float* out, void* ws, size_t ws_bytes, void* stream);
size_t jt_kde_grid2d_workspace_size(int64_t n, int32_t nx, int32_t ny);
int jt_kde_grid2d(const float* P, int64_t n, const float* gx, int32_t nx, const float* gy, int32_t ny,
                  float h, float* out, void* ws, size_t ws_bytes, void* stream);

/* ---- unstructured FEM field: element octree + P1 tetrahedra ---------------- */
/* Opaque host octree (OctreeMeshOptimized, octree_fem_interpolator_optimized.py:18-38). */
typedef struct jt_octree jt_octree;

/* Replaces build_octree_mesh_optimized (octree_fem_interpolator_optimized.py:41-215),
 * on the HOST: points (n_points,3) f32 and conn (n_elems,4) int32 are host
 * arrays.  The arrays it produces are bit-identical to the reference's. */
int  jt_octree_build(const float* points, int64_t n_points, const int32_t* conn, int64_t n_elems,
                     int32_t max_per_leaf, int32_t max_depth, jt_octree** out);
int  jt_octree_info(const jt_octree* t, int64_t* n_nodes, int32_t* max_elems, int32_t* tree_depth);
/* Copies the flattened arrays into caller-allocated HOST buffers (any may be NULL):
 * nodes_min/max (n_nodes,3) f32, nodes_elements (n_nodes,max_elems) int32 (-1 pad),
 * counts (n_nodes) int32, children (n_nodes,8) int32, is_leaf (n_nodes) u8,
 * element_bounds (n_elems,2,3) f32, centroids (n_elems,3) f32. */
int  jt_octree_export(const jt_octree* t, float* nodes_min, float* nodes_max, int32_t* nodes_elements,
                      int32_t* counts, int32_t* children, uint8_t* is_leaf, float* element_bounds,
                      float* centroids);
void jt_octree_free(jt_octree* t);

/* build_octree_mesh_optimized (octree_fem_interpolator_optimized.py:41-215)
 * on the GPU, arrays bit-identical to jt_octree_build's.  points (n_points, 3)
 * f32 and conn (n_elems, 4) int32 are DEVICE arrays; element_bounds
 * (n_elems, 2, 3) and centroids (n_elems, 3) f32 are written.  The tree (at
 * most max_nodes nodes; JT_ENOMEM naming max_nodes beyond it) stays in the
 * workspace (jt_octree_build_device_workspace_size(n_elems, max_nodes)
 * bytes) for jt_octree_export_device; *n_nodes, *max_elems (the widest
 * element row) and *tree_depth size its outputs.  The tree's shape depends
 * on the data, so this call synchronises `stream` (it reads two counts per
 * level). */
size_t jt_octree_build_device_workspace_size(int64_t n_elems, int64_t max_nodes);
int jt_octree_build_device(const float* points, int64_t n_points, const int32_t* conn, int64_t n_elems,
                           int32_t max_per_leaf, int32_t max_depth, int64_t max_nodes, float* element_bounds,
                           float* centroids, void* ws, size_t ws_bytes, int64_t* n_nodes, int32_t* max_elems,
                           int32_t* tree_depth, void* stream);
/* The flattened arrays of the tree jt_octree_build_device left in ws (same
 * n_elems, max_nodes), in depth-first pre-order, into DEVICE buffers:
 * nodes_min/max (n_nodes, 3) f32, nodes_elements (n_nodes, max_elems) int32
 * (-1 pad), counts (n_nodes) int32, children (n_nodes, 8) int32, is_leaf
 * (n_nodes) u8.  Asynchronous on stream. */
int jt_octree_export_device(const void* ws, size_t ws_bytes, int64_t n_elems, int64_t max_nodes, int64_t n_nodes,
                            int32_t max_elems, float* nodes_min, float* nodes_max, int32_t* nodes_elements,
                            int32_t* counts, int32_t* children, uint8_t* is_leaf, void* stream);

/* Device-resident octree field: the exported arrays plus nodal snapshots. */
typedef struct jt_octree_field {
    const float*   points;            /* (n_points, 3)             */
    const int32_t* conn;              /* (n_elems, 4)              */
    const float*   element_bounds;    /* (n_elems, 2, 3)           */
    const float*   nodes_min;         /* (n_nodes, 3)              */
    const float*   nodes_max;
    const int32_t* nodes_elements;    /* (n_nodes, max_elems)      */
    const int32_t* nodes_elem_counts; /* (n_nodes)                 */
    const int32_t* nodes_children;    /* (n_nodes, 8)              */
    const uint8_t* nodes_is_leaf;     /* (n_nodes)                 */
    int64_t n_points, n_elems, n_nodes;
    int32_t max_elems, max_depth;
    const void*  data;                /* (n_snap, n_points, 3) f32 */
    const float* times;               /* (n_snap,) f32             */
    int32_t n_snap, time_interp, time_extrap;
    const float* element_records;     /* optional (n_elems, 24) f32, 16-byte
                                         aligned, as jt_octree_element_records
                                         writes them: per element its box (min
                                         xyz, max xyz), the xyz of its 4 vertices,
                                         its 4 point ids (int32 bits), 2 zeros.
                                         The leaf scan then reads one contiguous
                                         96-byte record per candidate instead of
                                         box, connectivity and 4 vertex rows.
                                         NULL: not used.                        */
} jt_octree_field;

/* The per-element records of field (element_records layout above; the
 * field's element_records is ignored) into out, a device array of
 * n_elems * 24 floats, 16-byte aligned. */
int jt_octree_element_records(const jt_octree_field* field, float* out, void* stream);

/* Replaces OctreeFEMTimeSeriesFieldOptimized.sample_at_positions
 * (octree_fem_time_series_optimized.py:87-113).  found (nullable) receives 1
 * where a containing element was found, 0 where the fallback was used. */
int jt_octree_sample(const jt_octree_field* field, const float* x, int64_t n, float t, float* out,
                     int32_t* found, void* stream);
/* No reference counterpart: per query point, the three input-determined
 * counts of the point-location byte model (stats[3*i+0..2] = octree levels
 * descended, leaf candidates scanned up to the first hit, candidates whose
 * box contains the point), for the benchmark's roofline. */
int jt_octree_locate_stats(const jt_octree_field* field, const float* x, int64_t n, float t, int32_t* stats,
                           void* stream);
size_t jt_track_octree_workspace_size(int64_t n, int32_t n_steps, const jt_track_opts* opts);
/* jt_track_grid for an octree field (same arguments and recording). */
int jt_track_octree(const jt_octree_field* field, const jt_boundary* bc, int32_t integrator, float* x,
                    int64_t n, const float* step_times, float dt, int32_t n_steps, int32_t record_every,
                    float* traj, const jt_track_opts* opts, void* ws, size_t ws_bytes, void* stream);

/* ---- nearest-node TimeSeriesField ------------------------------------------ */
/* TimeSeriesField (time_series.py:82-473): snapshots (n_snap, n_nodes, 3) at
 * fixed node positions, sampled at the nearest node (ties: lowest index). */
typedef struct jt_nodes_field {
    const float* node_pos;    /* (n_nodes, 3) f32 */
    int64_t n_nodes;
    const void*  data;        /* (n_snap, n_nodes, 3) f32 */
    const float* times;       /* (n_snap,) f32 */
    int32_t n_snap, time_interp, time_extrap;
} jt_nodes_field;

/* Replaces TimeSeriesField.sample_at_positions (time_series.py:391-398):
 * _sample_all_nodes_at_time_jax then _nearest_neighbor_sample_jax (:249-259).
 * nearest (nullable) receives the chosen node index. */
int jt_nodes_sample(const jt_nodes_field* field, const float* Q, int64_t m, float t, float* out,
                    int32_t* nearest, void* stream);
size_t jt_track_nodes_workspace_size(int64_t n, int32_t n_steps);
/* jt_track_grid for a nearest-node field (same arguments and recording). */
int jt_track_nodes(const jt_nodes_field* field, const jt_boundary* bc, int32_t integrator, float* x,
                   int64_t n, const float* step_times, float dt, int32_t n_steps, int32_t record_every,
                   float* traj, const jt_track_opts* opts, void* ws, size_t ws_bytes, void* stream);

/* ---- adjoint RK4 (build-only: the reference has no differentiation) ------- */
size_t jt_track_grid_adjoint_workspace_size(int64_t n, int32_t n_steps, int32_t ckpt_every);
/* Reverse-mode gradient of n_steps RK4 steps of jt_track_grid (same field,
 * boundary, step_times and dt).  ckpt holds the forward positions at steps
 * 0, K, 2K, ... (K = ckpt_every; exactly jt_track_grid's trajectory with
 * record_every = K), (n_steps/K + 1, n, 3).  Given grad_xT = dL/dx_T, writes
 * grad_x0 = dL/dx_0 and ADDS dL/dU into grad_U, a float32
 * (n_snap, nx, ny, nz, 3) array the caller zeroes.  Atomic accumulation:
 * grad_U is reproducible to rounding, not bitwise.
 * host_step_times (n_steps) and host_field_times (n_snap) are HOST copies of
 * step_times and field->times with the same values: the library cuts its
 * reverse sweep where the snapshot pair changes from them, so the call never
 * reads the device and stays asynchronous on `stream`. */
int jt_track_grid_adjoint(const jt_grid_field* field, const jt_boundary* bc, const float* ckpt,
                          int64_t n, const float* step_times, const float* host_step_times,
                          const float* host_field_times, float dt, int32_t n_steps,
                          int32_t ckpt_every, const float* grad_xT, float* grad_x0, float* grad_U,
                          void* ws, size_t ws_bytes, void* stream);

/* Replaces the meshgrid expansion of uniform_grid_seeds
 * (jaxtrace/tracking/seeding.py:185-247): out (nx*ny*nz, 3) float32 with row
 * (i*ny + j)*nz + k = (xc[i], yc[j], zc[k]), np.meshgrid(..., indexing='ij')
 * order.  xc/yc/zc are the coordinate arrays the reference builds with
 * np.linspace (device pointers, nx/ny/nz floats). */
int jt_grid_seeds(const float* xc, int32_t nx, const float* yc, int32_t ny, const float* zc, int32_t nz,
                  float* out, void* stream);

/* Replaces random_seeds (jaxtrace/tracking/seeding.py:140-182): out (n, 3)
 * float32 = lo + u * (hi - lo) with u = jax.random.uniform(PRNGKey(rng_seed),
 * (n, 3)), the threefry2x32 bits in JAX's non-partitionable layout (the
 * default before JAX 0.5).  lo, hi: HOST float[3] (the validated bounds);
 * out: device.  n <= 1431655765 (3n words of one counter range). */
int jt_random_seeds(int64_t n, uint64_t rng_seed, const float* lo, const float* hi, float* out, void* stream);

/* The expansion of _generate_inlet_grid_particles
 * (jaxtrace/tracking/boundary.py:1321-1441): out (n, 3) float32 with the
 * flow-axis coordinate inlet_position and, on the two other axes in
 * increasing order, the point indices[i] of np.meshgrid(c0, c1, 'ij')
 * raveled (c0 (n0,), c1 (n1,): the np.linspace face axes).  indices (device
 * int32) are chosen on the host as the reference chooses them; an index
 * outside [0, n0*n1) gives NaN there. */
int jt_inlet_grid_particles(const int32_t* indices, int64_t n, int32_t axis, float inlet_position, const float* c0,
                            int32_t n0, const float* c1, int32_t n1, float* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* JAXTRACE_AMD_H */
