"""Benchmark: RK4 particle-steps/s at 16M particles, FP32 vs BF16 field.

Workload (one GPU): 2^24 particles uniform in the unit box (seed 3), RK4,
dt = 2e-3, x periodic / y,z reflective walls, through a 256^3 x 16-snapshot
divergence-free random Fourier field (64 modes, |k| <= 8, unit RMS speed,
seed 2) stored in float32 (headline) and bfloat16.  A "step" is one RK4 step
of every particle; inputs are resident in HBM when timing starts.

Prints ONE JSON line (rank 0).  Roofline: the tracking kernel's HBM bytes
per particle-step measured by rocprofv3 (FETCH_SIZE and WRITE_SIZE, scaled by
a plain copy's readings; profiles/kernel_counters.json) x the particle-steps of
one launch, over the launch's average duration timed with HIP events on the
launch stream, against 8 TB/s.  The algorithmic gather model (792 B per
particle-step in FP32, 408 B in BF16: 24 B of state plus 4 stages x 2
snapshots x 8 corners x 3 components) is reported beside it.  The other
sections (BF16, RK2/Euler, octree, adjoint, density, seeds, VTK ingestion)
report their own rates and rooflines the same way.
"""
from __future__ import annotations

import argparse
import ctypes as C
import hashlib
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_PARTICLES = 1 << 24
GRID_N = 256
N_SNAP = 16
DT = 2e-3
# KDE cut-off in bandwidths: the one the sampled cut-off parity test runs at
KDE_CUTOFF_SIGMAS = 7.0


def kernel_source_sha(files) -> str:
    """sha256 over the named kernel sources (paths relative to the repository)."""
    h = hashlib.sha256()
    for f in files:
        with open(os.path.join(ROOT, f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


COUNTERS = os.path.join(ROOT, "profiles", "kernel_counters.json")


def counter_entry(key):
    """rocprofv3 counter summary of one kernel from profiles/kernel_counters.json,
    if it was taken from the current sources of that kernel (sha256 match)."""
    try:
        with open(COUNTERS) as fh:
            prof = json.load(fh)
    except (OSError, ValueError):
        return None
    e = prof.get("kernels", {}).get(key)
    if not e or e.get("source_sha256") != kernel_source_sha(e.get("source_files", [])):
        return None
    return e


def hbm_roofline(key, units, ms, **extra):
    """HBM roofline of one kernel launch from measured traffic: the kernel's
    HBM bytes per unit (rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by a
    plain copy's readings) x the units the launch processed, over the launch
    time `ms`.  Counters are per unit, so they hold for any launch length."""
    e = counter_entry(key)
    out = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": None, "traffic": None}
    if e is None:
        out["note"] = f"no counter profile of {key} for the current kernel sources"
    else:
        traffic = e["hbm_bytes_per_unit"] * units
        ach = traffic / (ms * 1e-3) / 1e9
        out.update(achieved=ach, frac=ach / HBM_PEAK_GBS, traffic=traffic,
                   hbm_bytes_per_unit=e["hbm_bytes_per_unit"], unit_of_work=e["unit"],
                   l2_hit_rate=e.get("l2_hit_rate"), counters=f"profiles/kernel_counters.json#{key}")
        for k in ("ta_busy", "td_busy", "valu_insts_per_wave", "binding"):
            if k in e:
                out[k] = e[k]
    out.update(extra)
    return out


def algorithmic_bytes_per_particle_step(elem_bytes: int) -> int:
    return 24 + 4 * 2 * 8 * 3 * elem_bytes


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--particles", type=int, default=N_PARTICLES)
    ap.add_argument("--steps-per-launch", type=int, default=50)
    ap.add_argument("--sort-every", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bf16", action="store_true")
    ap.add_argument("--cpu-particles", type=int, default=1 << 20)
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the octree and density measurements")
    ap.add_argument("--only", choices=["track", "adjoint", "octree", "sph", "kde", "kde2d", "kde3d", "kdepts",
                                       "hashgrid", "nearest", "seeding", "vtk"],
                    help="run one section only and print its JSON (profiler runs)")
    ap.add_argument("--adjoint-reps", type=int, default=3,
                    help="adjoint repetitions per storage; the last one is reported")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (known HBM traffic) for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE readings")
    return ap.parse_args()


def launch_ranges(n_steps, spl, sort_every):
    """The [s0, s1) step ranges of the tracking launches one jt_track_grid
    call makes: spl steps at most, cut where a cell sort is due."""
    out, s0 = [], 0
    while s0 < n_steps:
        s1 = min(s0 + spl, n_steps)
        if sort_every > 0:
            s1 = min(s1, (s0 // sort_every + 1) * sort_every)
        out.append((s0, s1))
        s0 = s1
    return out


def run_tracking(lib, L, desc, bc, x, step_times, n_steps, spl, ws, stream, torch, record_events, sort_every=0,
                 integ=None):
    """One jt_track_grid call over [0, n_steps), as track_particles makes it
    (launches of up to spl steps, a cell sort every sort_every steps).  With
    record_events, returns (call_ms / launch, launch_ms, steps) per launch: the
    library records a pair of HIP events right around each tracking-kernel
    launch, on the stream it launches on."""
    o = L.TrackOpts()
    o.steps_per_launch = spl
    o.sort_every = sort_every
    o.block = 0
    ranges = launch_ranges(n_steps, spl, sort_every)
    if record_events:
        kev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * len(ranges))]
        for e in kev:
            e.record()                     # creates the underlying HIP event
        evarr = (C.c_void_p * len(kev))(*[e.cuda_event for e in kev])
        o.launch_events = C.cast(evarr, C.POINTER(C.c_void_p))
        o.n_launch_events = len(kev)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
    rc = lib.jt_track_grid(C.byref(desc), C.byref(bc), L.JT_RK4 if integ is None else integ, L.ptr(x), x.shape[0],
                           L.ptr(step_times),
                           C.c_float(np.float32(DT)), n_steps, 1, None, C.byref(o), L.ptr(ws), ws.numel(), stream)
    L.check(rc, "jt_track_grid")
    if not record_events:
        return []
    e1.record()
    torch.cuda.synchronize()
    per_call = e0.elapsed_time(e1) / max(len(ranges), 1)
    return [(per_call, kev[2 * k].elapsed_time(kev[2 * k + 1]), b - a) for k, (a, b) in enumerate(ranges)]


def measure(field, bc_obj, x0, args, torch, L, lib, sort_every=0, integ=None, steps=None, warmup=None):
    """Warm-up, then the timed steps, each one jt_track_grid call (launches of
    spl steps: the steps per launch, capped at the sort interval)."""
    dev = field.device
    n = x0.shape[0]
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    total = warmup + steps
    times = (np.arange(total + 1, dtype=np.float64) * DT).astype(np.float32)
    step_times = torch.from_numpy(times).to(dev)
    desc = field.descriptor()
    bc = bc_obj.descriptor(dev)
    o = L.TrackOpts()
    o.sort_every = sort_every
    ws_bytes = int(lib.jt_track_grid_workspace_size(n, total, C.byref(o)))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    stream = L.stream_handle()
    spl = launch_steps(args, sort_every)
    x = x0.clone()
    run_tracking(lib, L, desc, bc, x, step_times, warmup, spl, ws, stream, torch, False, sort_every, integ)
    torch.cuda.synchronize()
    x = x0.clone()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    durs = run_tracking(lib, L, desc, bc, x, step_times, steps, spl, ws, stream, torch, True, sort_every, integ)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, durs, x


def launch_steps(args, sort_every):
    spl = args.steps_per_launch
    return min(spl, sort_every) if sort_every > 0 else spl


def cpu_baseline(host_data, times, grid, x0_host, n_part, n_steps):
    """The oracle's NumPy restatement of the reference algorithm (blend every
    node, then trilinear; rk4_step; mixed boundary), single process."""
    from oracle import boundary as OB
    from oracle import fields as OF
    from oracle import integrators as OI
    meta = OF.GridMeta(grid, grid, grid)
    x = x0_host[:n_part].copy()
    fn = lambda p, t: OF.structured_series_sample(host_data, times, meta, p, t, "periodic")
    t0 = time.perf_counter()
    for s in range(n_steps):
        x = OI.rk4_step(x, np.float32(s * DT), DT, fn)
        x = OB.per_axis(x, [[0, 0, 0], [1, 1, 1]], ["periodic", "reflective", "reflective"])
    el = time.perf_counter() - t0
    return n_part * n_steps / el, el


# CPU baselines of the other rows are timed beside their GPU lines unless
# --no-cpu-baseline is given (main sets this).
CPU_BASELINES = {"on": True}


def cpu_host_info():
    """Host facts that go with every CPU baseline: the cores this process may
    use, the machine's count, the CPU model and the NumPy thread pools."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    pools = None
    try:
        import threadpoolctl
        pools = [{"api": d.get("internal_api"), "threads": d.get("num_threads")}
                 for d in threadpoolctl.threadpool_info()]
    except Exception:
        pass
    env = {k: os.environ[k] for k in ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
           if k in os.environ}
    return {"host_cpus": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count(), "cpu_model": model,
            "numpy_thread_pools": pools, "thread_env": env}


def cpu_sph_baseline(P, Q, h):
    """The oracle's hash grid (the reference's HashGridNeighbors build: keys,
    stable argsort, CSR) on all particles, then its query_many and the cubic
    spline sum for a sample of the query grid, single process."""
    from oracle import density as OD
    t0 = time.perf_counter()
    grid = OD.HashGrid(P, h, 64, "x32")
    build_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    idx, dist = grid.query_many(Q, 2.0 * h, 512)
    w = OD.cubic_spline(dist, np.full(dist.shape, np.float32(h), np.float32), 3)
    (np.where(idx >= 0, w, np.float32(0))).sum(axis=1, dtype=np.float32)
    q_s = time.perf_counter() - t0
    info = cpu_host_info()
    return ({"value": len(P) / build_s, "unit": "points/s", "cores": 1, "kind": "port",
             "sample": f"HashGrid build of all {len(P)} points, x32 keys, cell h ({build_s:.1f} s)", **info},
            {"value": len(Q) / q_s, "unit": "queries/s", "cores": 1, "kind": "port",
             "sample": f"{len(Q)} grid queries over the {len(P)}-particle hash grid: 27-probe candidate lists, "
                       f"512-neighbour selection, cubic-spline sum ({q_s:.1f} s)", **info})


def cpu_kde_baseline(P, Q, h, total="all 2^24 x 256^3 pairs"):
    """The reference's exhaustive KDE sum (the oracle's kde_eval) for a
    particle sample and a grid-point sample, single process."""
    from oracle import density as OD
    t0 = time.perf_counter()
    OD.kde_eval(Q, P, h)
    el = time.perf_counter() - t0
    return {"value": len(P) * len(Q) / el, "unit": "pairs/s", "cores": 1, "kind": "port",
            "sample": f"exhaustive Gaussian sum, {len(P)} particles x {len(Q)} grid points ({el:.1f} s); "
                      f"the reference evaluates {total}", **cpu_host_info()}


def cpu_octree_baseline(field, pts, conn, data, times, x0, n_steps=2):
    """The reference's octree RK4 path restated by the oracle (every node
    blended per stage, fixed-depth descent, leaf scan with barycentric tests,
    continuous grid inlet), on the octree arrays the device built (equal to
    the oracle's build, tested), for a particle sample, single process."""
    from oracle import boundary as OB
    from oracle import integrators as OI
    from oracle import octree as OO
    tree = {k: field._dev[k].cpu().numpy() for k in ("nodes_min", "nodes_max", "nodes_elements", "nodes_children",
                                                     "element_bounds")}
    tree["nodes_is_leaf"] = field._dev["nodes_is_leaf"].cpu().numpy().astype(bool)
    fn = lambda p, t: OO.sample(tree, pts, conn, data, times, p, t, field.max_depth)
    x = np.array(x0, np.float32)
    t0 = time.perf_counter()
    for s in range(n_steps):
        x = OI.rk4_step(x, np.float32(s * 1e-3), 1e-3, fn)
        x = OB.continuous_inlet(x, [[0, 0, 0], [1, 1, 1]], "x", "positive", "grid", {"x": 64, "y": 64, "z": 64})
    el = time.perf_counter() - t0
    return {"value": len(x0) * n_steps / el, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{len(x0)} particles x {n_steps} RK4 steps of the same workload, octree arrays from the "
                      f"device build ({el:.1f} s)", **cpu_host_info()}


def cpu_adjoint_baseline(data, times, g, x0, tgt, n_steps=10):
    """Forward plus backward on the host: the oracle's float64 restatement of
    the tracker (torch float64 autograd, every node blended per stage) for a
    particle sample, with the same loss; the reference has no derivative of
    its own."""
    import torch
    from oracle import adjoint_ref as A
    from oracle import fields as OF
    U = torch.from_numpy(np.asarray(data, np.float64)).requires_grad_(True)
    X = torch.from_numpy(np.asarray(x0, np.float64)).requires_grad_(True)
    meta = OF.GridMeta(g, g, g)
    st = np.arange(n_steps, dtype=np.float32) * np.float32(DT)
    t0 = time.perf_counter()
    xT = A.track(U, times, meta, X, st, DT, ["periodic"] * 3)
    loss = ((xT - torch.from_numpy(np.asarray(tgt, np.float64))) ** 2).sum(1).mean()
    t1 = time.perf_counter()
    loss.backward()
    t2 = time.perf_counter()
    return {"value": len(x0) * n_steps / (t2 - t0), "unit": "particle-steps/s (forward + backward)", "cores": 1,
            "kind": "port", "forward_s": t1 - t0, "backward_s": t2 - t1,
            "sample": f"{len(x0)} particles x {n_steps} RK4 steps, float64 torch autograd on the host "
                      f"(every node blended per stage), loss mean |x_T - x*|^2", **cpu_host_info()}


def trajectory_error_vs_fp64(torch, jt, data, times, grid, x0, bc, n_steps=1000, every=10, sample=1024,
                             report=(10, 20, 50, 100, 200, 500, 1000)):
    """Trajectory error of the GPU's own output on the headline workload over
    the full run: all particles tracked on the GPU for n_steps RK4 steps from
    t = 0 with float32 and bfloat16 snapshot storage (a frame every `every`
    steps).  For `sample` particles (seed 11), |x_gpu - x_fp64| per frame in
    units of the domain size (1), against float64 tracks (float64 arithmetic
    on the stored values, oracle.adjoint_ref.track_np) of the same particles
    through (a) the float32 field and (b) the field as stored in bfloat16.
    (a) is storage rounding and float32 arithmetic together; (b) isolates the
    arithmetic.  The field is chaotic, so any difference grows exponentially:
    the line reports the max and median at the `report` steps and the last
    frame at which the max over the sample is still within 1e-4."""
    from oracle import adjoint_ref as A
    from oracle import fields as OF
    meta = OF.GridMeta(grid, grid, grid)
    kinds = ["periodic", "reflective", "reflective"]
    n = x0.shape[0]
    idx = np.sort(np.random.default_rng(11).choice(n, sample, replace=False))
    tidx = torch.from_numpy(idx).to(x0.device)
    tg = np.linspace(0.0, n_steps * DT, n_steps + 1, dtype=np.float32)
    gpu = {}
    for storage in ("float32", "bfloat16"):
        f = jt.StructuredTimeSeriesField(data, times, grid, grid, grid, bounds_handling="periodic",
                                         storage=storage, layout="halo")
        tr = jt.create_tracker("rk4", f, bc, record_every=every, return_device=True)
        traj = tr.track_particles(x0, (0.0, n_steps * DT), n_steps + 1, dt=DT)
        gpu[storage] = traj.positions[:, tidx].cpu().numpy().astype(np.float64)
        del traj, tr, f
        torch.cuda.empty_cache()
    host32 = data.cpu().numpy()
    hostbf = data.to(torch.bfloat16).to(torch.float32).cpu().numpy()
    dt = float(np.float32(DT))
    steps = np.arange(0, n_steps + 1, every)
    out = {"report_steps": list(report)}
    for ref_name, host in (("fp64_on_f32_field", host32), ("fp64_on_bf16_stored_field", hostbf)):
        x = x0[tidx].cpu().numpy().astype(np.float64)
        ref = [x]
        for k in range(0, n_steps, every):
            x = A.track_np(host, times, meta, x, tg[k:k + every], dt, kinds)
            ref.append(x)
        ref = np.stack(ref)
        for storage in ("float32", "bfloat16"):
            e = np.abs(gpu[storage] - ref).max(axis=2)          # (frames, sample)
            emax, emed = e.max(axis=1), np.median(e, axis=1)
            first_out = steps[emax > 1e-4]
            out[f"{storage}_vs_{ref_name}"] = {
                "max": [float(emax[r // every]) for r in report],
                "median": [float(emed[r // every]) for r in report],
                "max_within_1e-4_up_to_step": int(first_out[0] - every) if first_out.size else int(n_steps)}
    out["sample"] = (f"{sample} particles of the headline workload, {n_steps} RK4 steps (t = 0 .. "
                     f"{n_steps * DT:g}), |x_gpu - x_fp64| per particle (max over axes), domain size 1")
    return out


def bench_adjoint(torch, jt, W, dev, n_part=1 << 20, steps=500, K=32, reps=3):
    """Adjoint workload: 128^3 x 8 random Fourier field (as above, t in [0,1]),
    2^20 particles (seed 7), RK4 dt = 2e-3, 500 steps, periodic box,
    L = mean |x_T - x*|^2 (x* seed 8), checkpoints every 32 steps.  Returns
    forward and backward wall times."""
    from jaxtrace_amd import adjoint as ADJ
    data, times, g = W.random_fourier(n=128, n_snap=8, t_end=1.0, seed=2, device=dev)
    gen = torch.Generator(device="cpu").manual_seed(7)
    x0 = torch.rand((n_part, 3), generator=gen).to(dev)
    tgt = torch.from_numpy(np.random.default_rng(8).random((n_part, 3)).astype(np.float32)).to(dev)
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    out = {}
    for storage in ("float32", "bfloat16"):
        res = {}
        for rep in range(reps):                   # the earlier repetitions warm up
            U = data.clone().requires_grad_(True)
            X = x0.clone().requires_grad_(True)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            xT = ADJ.track_rk4(U, times, g, g, g, X, (0.0, steps * DT), steps + 1, dt=DT, boundary=bc,
                               storage=storage, checkpoint_every=K)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            loss = ((xT - tgt) ** 2).sum(1).mean()
            loss.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            res = {"fwd_s": t1 - t0, "bwd_s": t2 - t1, "ratio_bwd_over_fwd": (t2 - t1) / (t1 - t0)}
        out[storage] = res
    if CPU_BASELINES["on"]:
        out["cpu_baseline"] = cpu_adjoint_baseline(data.cpu().numpy(), times, g, x0[:1 << 14].cpu().numpy(),
                                                   tgt[:1 << 14].cpu().numpy())
    out["workload"] = (f"128^3 x 8 random Fourier, {n_part} particles, RK4 dt=2e-3, {steps} steps, "
                       f"checkpoints every {K}, loss mean |x_T - x*|^2")
    return out


VALU_LANE_OPS = 256 * 4 * 16 * 2.4e9      # non-packed FP32 VALU: 256 CUs x 4 SIMD16 x 2.4 GHz
MFMA_F32_TFLOPS = 157.3                    # dense FP32 MFMA peak of MI355X
MFMA_BF16_TFLOPS = 2500.0                 # dense BF16 MFMA peak of MI355X


def _events(torch):
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def bench_integrators(torch, L, lib, field, bc, x0, args, steps=200, warmup=25):
    """RK2 and Euler (rk2.py:117-180, euler.py:111-157) through the same fused
    kernel on the headline workload: FP32 field, 2^24 particles, mixed
    boundary, launches of `steps_per_launch` steps with the cell sort.  Timed:
    `steps` steps after `warmup`.  Roofline: measured HBM bytes per
    particle-step (profiles/kernel_counters.json); the gather model (24 B of
    state + per stage 2 snapshots x 8 corners x 12 B: RK2 408 B, Euler 216 B
    per particle-step) is reported beside it: registers, L1 and L2 serve most
    of those gathers."""
    n = x0.shape[0]
    spl = launch_steps(args, args.sort_every)
    out = {}
    for name, integ, stages in (("rk2", L.JT_RK2, 2), ("euler", L.JT_EULER, 1)):
        wall, durs, _ = measure(field, bc, x0, args, torch, L, lib, args.sort_every, integ=integ, steps=steps,
                                warmup=warmup)
        full = [r for r in durs if r[2] == spl] or durs
        launch_ms = float(np.mean([r[1] for r in full]))
        bpps = 24 + stages * 2 * 8 * 12
        gather = bpps * n * full[0][2] / (launch_ms * 1e-3) / 1e9
        out[name] = {"value": n * steps / wall, "unit": "particle-steps/s", "launch_ms": launch_ms,
                     "steps_per_launch": full[0][2],
                     "roofline": hbm_roofline(f"k_track_{name}_f32", n * full[0][2], launch_ms,
                                              gather_model={"bytes_per_particle_step": bpps, "GBs": gather})}
    out["workload"] = f"the headline FP32 workload, {steps} steps after {warmup}"
    return out


def bench_stream(torch, jt, field, bc_obj, x0, steps=100, every=25):
    """Streamed trajectory output (TrackerOptions.output_path) on the headline
    workload: frames recorded every `every` steps go to a .npy file in the
    temp directory while tracking runs.  Timed beside the same run kept on
    the device and the in-memory run with its one host copy at the end."""
    import tempfile
    path = os.path.join(tempfile.gettempdir(), "jt_stream_bench.npy")
    span = (0.0, steps * DT)

    def run(**kw):
        torch.cuda.synchronize()
        t = time.perf_counter()
        tr = jt.create_tracker("rk4", field, bc_obj, record_every=every, **kw)
        tr = tr.track_particles(x0, span, steps + 1, dt=DT)
        torch.cuda.synchronize()
        return time.perf_counter() - t, tr

    run(return_device=True)
    t_dev, _ = run(return_device=True)
    t_mem, mem = run()
    t_st, st = run(output_path=path)
    same = bool(np.array_equal(np.asarray(st.positions), mem.positions))
    nbytes = int(mem.positions.nbytes)
    del mem, st
    os.remove(path)
    return {"device_only_ms": t_dev * 1e3, "in_memory_ms": t_mem * 1e3, "streamed_ms": t_st * 1e3,
            "frames": 1 + steps // every, "bytes": nbytes, "streamed_GBs": nbytes / t_st / 1e9,
            "streamed_equals_in_memory": same,
            "workload": f"the headline FP32 workload, {x0.shape[0]} particles, {steps} steps, "
                        f"a frame every {every} steps"}


def bench_seeding(torch, jt, W, L, lib, dev, res=256, reps=20):
    """uniform_grid_seeds' meshgrid expansion (jt_grid_seeds) for res^3 seeds
    in the unit box, coordinate arrays resident: 12 B written per seed."""
    from jaxtrace_amd import seeding as S
    ax = [torch.from_numpy(a).to(dev) for a in S.grid_axes(res, [[0, 0, 0], [1, 1, 1]])]
    n = res ** 3
    out = torch.empty((n, 3), dtype=torch.float32, device=dev)

    def call():
        L.check(lib.jt_grid_seeds(L.ptr(ax[0]), res, L.ptr(ax[1]), res, L.ptr(ax[2]), res, L.ptr(out),
                                  L.stream_handle()), "jt_grid_seeds")

    call()
    torch.cuda.synchronize()
    e0, e1 = _events(torch)
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    gbs = n * 12 / (ms * 1e-3) / 1e9
    out_d = {"value": n / (ms * 1e-3), "unit": "seeds/s", "ms": ms, "seeds": n,
             "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": gbs / HBM_PEAK_GBS, "bytes_per_seed": 12},
             "workload": f"uniform_grid_seeds {res}^3 in the unit box, mean of {reps} calls"}
    del out
    out_d["random_seeds"] = _bench_random_seeds(torch, L, lib, dev, n, reps)
    out_d["inlet_grid_particles"] = _bench_inlet(torch, L, lib, dev, 1 << 22, reps)
    if CPU_BASELINES["on"]:
        from oracle import seeding as OS
        info = cpu_host_info()
        t0 = time.perf_counter()
        OS.uniform_grid_seeds(res, [[0, 0, 0], [1, 1, 1]])
        el = time.perf_counter() - t0
        out_d["cpu_baseline"] = {"value": n / el, "unit": "seeds/s", "cores": 1, "kind": "port",
                                 "sample": f"the oracle's uniform_grid_seeds {res}^3 (NumPy meshgrid, {el:.2f} s)",
                                 **info}
        nr = 1 << 20
        t0 = time.perf_counter()
        OS.random_seeds(nr, [[0, 0, 0], [1, 1, 1]], 42)
        el = time.perf_counter() - t0
        out_d["random_seeds"]["cpu_baseline"] = {
            "value": nr / el, "unit": "seeds/s", "cores": 1, "kind": "port",
            "sample": f"the oracle's threefry2x32 random_seeds for {nr} seeds (NumPy, {el:.2f} s)", **info}
        ni = 1 << 16
        t0 = time.perf_counter()
        OS.inlet_grid_particles(ni, 0.0, 0, [[0, 0, 0], [1, 1, 1]], {"x": 64, "y": 64, "z": 64})
        el = time.perf_counter() - t0
        out_d["inlet_grid_particles"]["cpu_baseline"] = {
            "value": ni / el, "unit": "particles/s", "cores": 1, "kind": "port",
            "sample": f"the oracle's inlet face grid for {ni} particles (the reference's index loops, {el:.2f} s)",
            **info}
    return out_d


def _timed(torch, call, reps):
    """Mean HIP-event time of reps calls on the current stream, after one."""
    call()
    torch.cuda.synchronize()
    e0, e1 = _events(torch)
    e0.record()
    for _ in range(reps):
        call()
    e1.record()
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def _bench_random_seeds(torch, L, lib, dev, n, reps):
    """random_seeds' JAX branch (seeding.py:140-174) on the GPU: n seeds of
    jax.random.uniform(PRNGKey(42), (n, 3)) scaled to the unit box
    (jt_random_seeds: one threefry2x32-20 block per two output words).
    12 B written per seed; the 20 rounds are VALU work."""
    out = torch.empty((n, 3), dtype=torch.float32, device=dev)
    lo = (C.c_float * 3)(0.0, 0.0, 0.0)
    hi = (C.c_float * 3)(1.0, 1.0, 1.0)

    def call():
        L.check(lib.jt_random_seeds(n, 42, lo, hi, L.ptr(out), L.stream_handle()), "jt_random_seeds")

    ms = _timed(torch, call, reps)
    gbs = n * 12 / (ms * 1e-3) / 1e9
    return {"value": n / (ms * 1e-3), "unit": "seeds/s", "ms": ms, "seeds": n,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_seed": 12,
                         "note": "the 20 threefry rounds per two words (integer VALU work) cost more than the "
                                 "12 B of writes per seed"},
            "workload": f"random_seeds n = {n}, rng_seed 42, unit box, mean of {reps} calls"}


def _bench_inlet(torch, L, lib, dev, n, reps):
    """_generate_inlet_grid_particles' face-grid expansion
    (boundary.py:1321-1441) for n particles on a 64 x 64 inlet face (the
    octree workload's concentrations), indices resident: 4 B read and 12 B
    written per particle."""
    c = torch.linspace(0.0, 1.0, 64, dtype=torch.float32, device=dev)
    idx = torch.randint(0, 64 * 64, (n,), generator=torch.Generator(device="cpu").manual_seed(5),
                        dtype=torch.int32).to(dev)
    out = torch.empty((n, 3), dtype=torch.float32, device=dev)

    def call():
        L.check(lib.jt_inlet_grid_particles(L.ptr(idx), n, 0, C.c_float(0.0), L.ptr(c), 64, L.ptr(c), 64, L.ptr(out),
                                            L.stream_handle()), "jt_inlet_grid_particles")

    ms = _timed(torch, call, reps)
    gbs = n * 16 / (ms * 1e-3) / 1e9
    return {"value": n / (ms * 1e-3), "unit": "particles/s", "ms": ms, "particles": n,
            "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs / HBM_PEAK_GBS, "bytes_per_particle": 16},
            "workload": f"inlet face grid 64 x 64 on x = 0, {n} particles, mean of {reps} calls"}


def bench_vtk(torch, jt, W, L, lib, dev, n=48, steps=4):
    """The on-disk input path (vtk_reader.py:426-942, example_workflow.py:
    436-660): a series of `steps` .pvtu files, each naming two .vtu pieces
    (raw appended binary) of the n^3 x 6 jittered-tet mesh with the
    Poiseuille-vortex velocity as 'Displacement', written to a temporary
    directory; then load_octree_field (host parse of every file, GPU octree
    build, uploads).  Host I/O bound: reported as file MB/s, with the parse
    and the field construction timed apart."""
    import shutil
    import tempfile
    from jaxtrace_amd import vtk_io as V
    pts, conn = W.kuhn_mesh(n, 0.2, 4)
    pts, conn = pts.astype(np.float32), conn.astype(np.int32)
    half = len(conn) // 2
    pieces = []
    for cc in (conn[:half], conn[half:]):
        used, inv = np.unique(cc, return_inverse=True)
        pieces.append((pts[used], inv.reshape(cc.shape).astype(np.int32)))
    allp = np.concatenate([pieces[0][0], pieces[1][0]])
    times = (np.arange(steps) / 14.0).astype(np.float32)
    vel = W.poiseuille_vortex(allp, times).astype(np.float32)
    d = tempfile.mkdtemp(prefix="jt_vtk_")
    try:
        na = len(pieces[0][0])
        for k in range(steps):
            srcs = []
            for j, (pp, cc) in enumerate(pieces):
                v = vel[k, :na] if j == 0 else vel[k, na:]
                f = f"p{j}_{k}.vtu"
                V.write_vtu(os.path.join(d, f), pp, cc, {"Displacement": v}, encoding="appended")
                srcs.append(f)
            V.write_pvtu(os.path.join(d, f"flow_{k:04d}.pvtu"), srcs, ["Displacement"])
        nbytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        pattern = os.path.join(d, "flow_*.pvtu")
        t0 = time.perf_counter()
        s = V.read_octree_series(pattern, max_timesteps=steps)
        t_parse = time.perf_counter() - t0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = jt.OctreeFEMTimeSeriesFieldOptimized(s["velocity_data"], s["times"], s["positions"], s["connectivity"])
        torch.cuda.synchronize()
        t_field = time.perf_counter() - t0
        t0 = time.perf_counter()
        f2 = jt.load_octree_field(pattern, max_timesteps=steps)
        torch.cuda.synchronize()
        t_all = time.perf_counter() - t0
        assert f2.connectivity.shape == (len(conn), 4)
        del f, f2
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"value": nbytes / t_parse / 1e6, "unit": "file MB/s (parse)", "file_bytes": nbytes,
            "parse_s": t_parse, "field_build_s": t_field, "load_octree_field_s": t_all,
            "workload": f"{steps} .pvtu files x 2 raw-appended .vtu pieces, {n}^3 x 6 tets ({len(conn)}), "
                        f"{len(allp)} points, 'Displacement' per file; host cores parse, the GPU builds the octree"}


def bench_nearest(torch, jt, W, L, lib, dev, m=1 << 20):
    """Nearest-node sampling of TimeSeriesField (time_series.py:249-259,
    391-412): the 32^3 lattice nodes of the ABC flow (4 snapshots) as the
    node set, 2^20 uniform queries at t = 0.5.  Timed: one
    sample_at_positions call (jt_nodes_sample, exhaustive as the reference).
    Rate: query-node pairs per second, against a VALU bound of 13 lane-ops
    per pair (the squared distance 8, the first-minimum test 5)."""
    data, times, g = W.abc_flow()
    T = data.shape[0]
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    f = jt.TimeSeriesField(data.reshape(T, -1, 3), times, pos, device=dev)
    q = torch.from_numpy(W.uniform_particles(m, 9)).to(dev)
    f.sample_at_positions(q, 0.5)
    torch.cuda.synchronize()
    e0, e1 = _events(torch)
    e0.record()
    f.sample_at_positions(q, 0.5)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    pairs = float(m) * pos.shape[0]
    rate = pairs / (ms * 1e-3)
    peak = VALU_LANE_OPS / 13.0
    extra = {}
    if CPU_BASELINES["on"]:
        from oracle import fields as OF
        qh = q[:4096].cpu().numpy()
        t0 = time.perf_counter()
        OF.nearest_node_sample(data.reshape(T, -1, 3), times, pos, qh, 0.5)
        el = time.perf_counter() - t0
        extra["cpu_baseline"] = {"value": len(qh) * pos.shape[0] / el, "unit": "pairs/s", "cores": 1, "kind": "port",
                                 "sample": f"{len(qh)} queries against all {pos.shape[0]} nodes, the oracle's "
                                           f"exhaustive argmin ({el:.1f} s)", **cpu_host_info()}
    return {**extra, "value": rate, "unit": "pairs/s", "ms": ms, "queries": m, "nodes": int(pos.shape[0]),
            "roofline": {"bound": "valu", "achieved": rate, "peak": peak, "unit": "pairs/s", "frac": rate / peak},
            "workload": "ABC flow 32^3 nodes x 4 snapshots, 2^20 uniform queries, nearest node (exhaustive)"}


def bench_octree(torch, jt, W, L, lib, dev, steps=20):
    """Unstructured FEM workload: the unit box split into 128^3 hexes x 6 tets
    (interior nodes jittered by up to 0.2 h, seed 4: 12,582,912 tets on 129^3
    nodes), 8 snapshots of the Poiseuille-plus-vortex field, 2^22 uniform
    particles (seed 5), RK4 dt = 1e-3 with the continuous inlet at x = 0
    (64 x 64 grid replacement), particles Hilbert-sorted every 10 steps.
    Timed: `steps` steps after a warm-up run (the whole call, cell sorts
    included).  Roofline: measured HBM bytes per particle-step
    (profiles/kernel_counters.json); beside it the point-location byte model
    24 + 4 [L 29 + c 28 + a 64 + 96] B per particle-step, with L, c, a the
    mean levels descended, leaf candidates scanned and box hits counted by
    jt_octree_locate_stats at the initial positions."""
    pts, conn = W.kuhn_mesh(128, 0.2, 4)
    times = (np.arange(8) / 14.0).astype(np.float32)
    data = W.poiseuille_vortex(pts, times)
    t0 = time.perf_counter()
    f = jt.OctreeFEMTimeSeriesFieldOptimized(data, times, pts, conn)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t0
    bc = jt.continuous_inlet_boundary_factory([[0, 0, 0], [1, 1, 1]], "x", "positive", "grid",
                                              {"x": 64, "y": 64, "z": 64})
    n = 1 << 22
    x0 = torch.from_numpy(W.uniform_particles(n, 5)).to(dev)
    tr = jt.create_tracker("rk4", f, bc, return_device=True, record_final_only=True)
    span = (0.0, steps * 1e-3)
    tr.track_particles(x0, span, steps + 1, dt=1e-3)
    e0, e1 = _events(torch)
    torch.cuda.synchronize()
    e0.record()
    tr.track_particles(x0, span, steps + 1, dt=1e-3)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    rate = n * steps / (ms * 1e-3)
    stats = torch.empty((n, 3), dtype=torch.int32, device=dev)
    L.check(lib.jt_octree_locate_stats(C.byref(f.descriptor()), L.ptr(x0), n, C.c_float(0.0), L.ptr(stats),
                                       L.stream_handle()), "jt_octree_locate_stats")
    Lm, cm, am = (float(v) for v in stats.double().mean(0).cpu())
    bpps = 24 + 4 * (Lm * 29 + cm * 28 + am * 64 + 96)
    extra = {}
    if CPU_BASELINES["on"]:
        extra["cpu_baseline"] = cpu_octree_baseline(f, pts, conn, data, times, x0[:4096].cpu().numpy())
    return {**extra, "value": rate, "unit": "particle-steps/s", "ms": ms, "steps": steps, "particles": n,
            "octree_build_s": build_s,
            "roofline": hbm_roofline("k_oct_track", n * steps, ms,
                                     locate_model={"bytes_per_particle_step": bpps, "GBs": bpps * rate / 1e9,
                                                   "levels": Lm, "candidates_scanned": cm, "box_hits": am}),
            "workload": "128^3 x 6 jittered tets, 8 snapshots, 2^22 particles, RK4 dt=1e-3, continuous inlet"}


def bench_density(torch, jt, W, L, lib, dev, kde_res=256, sample=1 << 16, parts=("sph", "kde", "kde2d", "kde3d", "kdepts")):
    """Density workload: 2^24 particles from the 64-component Gaussian mixture
    (seed 6).
    SPH: the 256^3 evaluate_3d grid as queries, 3-D cubic spline, h = 2/256,
    hash grid of cell h with 64 slots per cell, 512 neighbours, x32 keys (the
    reference as shipped); timed: the estimator's evaluation call for x32
    keys and uniform masses (jt_sph_eval_tiled: queries ordered by their
    cell's (y, z) column, one column's candidates staged in LDS per wave of
    64 queries, the top-K pass for queries with more than 512 valid
    candidates); the hash-grid construction (HashGridNeighbors on device
    points: bounds, keys, counting-sort passes, cell table, sorted positions;
    byte model 36 B per point); and the estimator's whole setup.
    Roofline: measured HBM bytes per query of the jt_sph_eval_tiled call (its
    kernels) and per point of the build (profiles/kernel_counters.json);
    beside them the candidate model 9 x 8 + C_q/3 x 16 + 4 B per query (C_q
    the reference's candidate count over its 27 probes, from
    jt_hashgrid_query_stats on a sample of queries; x32 keys make 9 distinct
    cells, each read once) and the 36 B per point build model.
    KDE: Scott bandwidth, evaluate_3d grid of kde_res^3 points, cut-off
    KDE_CUTOFF_SIGMAS h, the cut-off the sampled GPU parity test
    (tests/test_gpu_density.py::test_kde_cutoff_tiles_sampled) holds to 1e-5
    of the float64 exhaustive sum; every dropped term is below
    (2 pi)^(-3/2) exp(-c^2/2), so their mean is too, and the line reports
    that bound relative to the density floor (1e-3 of the maximum); timed: the
    jt_kde_grid_cutoff_ws call (tile kernel and slice sum).  Rate: pairs within the cut-off per second
    (counted by jt_hashgrid_query_stats on a sample).  Roofline: the dense
    BF16 MFMA peak against 8 flops (four bfloat16 products of the split
    float32 operands) per (grid point, particle) pair of each 16 x 16 x 32 tile and
    the particles within R of the tile's box, the products the kernel's
    matrix cores evaluate (estimated from 512 sampled tiles).
    KDE 2-D, all-pairs 3-D and evaluate(points): see _bench_kde2d,
    _bench_kde3d and _bench_kde_points.
    parts: which of 'hashgrid', 'sph', 'kde', 'kde2d', 'kde3d', 'kdepts' to run."""
    F = np.float32
    P = W.gaussian_mixture(1 << 24, seed=6)
    Pd = torch.from_numpy(P).to(dev)
    h = 2.0 / 256
    a = np.linspace(0.0, 1.0, 256, dtype=F)
    X, Y, Z = np.meshgrid(a, a, a, indexing="xy")
    Qd = torch.from_numpy(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(F)).to(dev)
    del X, Y, Z
    M = int(Qd.shape[0])
    out = {}
    if "hashgrid" in parts and "sph" not in parts:
        jt.HashGridNeighbors(Pd, h, 64, "x32", dev)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        jt.HashGridNeighbors(Pd, h, 64, "x32", dev)
        torch.cuda.synchronize()
        hg_ms = (time.perf_counter() - t0) * 1e3
        out["hashgrid_build"] = {"ms": hg_ms, "roofline": hbm_roofline("hashgrid_build", len(P), hg_ms)}
        return out
    if "sph" not in parts:
        if "kde" in parts:
            out = _bench_kde(torch, jt, L, lib, dev, P, Pd, kde_res, out)
        if "kde2d" in parts:
            out = _bench_kde2d(torch, jt, L, lib, dev, P, out)
        if "kde3d" in parts:
            out = _bench_kde3d(torch, jt, L, lib, dev, P, out)
        if "kdepts" in parts:
            out = _bench_kde_points(torch, jt, L, lib, dev, P, out)
        return out
    # the hash grid alone (x32 keys, cell h): bounds, keys, radix sort, cell
    # table, sorted positions; after one warm-up build
    jt.HashGridNeighbors(Pd, h, 64, "x32", dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    jt.HashGridNeighbors(Pd, h, 64, "x32", dev)
    torch.cuda.synchronize()
    hg_ms = (time.perf_counter() - t0) * 1e3
    e0, e1 = _events(torch)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    est = jt.SPHDensityEstimator(Pd, dimensions="3d", smoothing_length=h, kernel_type="cubic_spline",
                                 per_cell_cap=64, max_neighbors=512, normalize=True, key_mode="x32")
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t0) * 1e3
    rho = torch.empty(M, dtype=torch.float32, device=dev)
    radius = F(2.0 * est._mean_h)
    ws_bytes = int(lib.jt_sph_tiled_workspace_size(C.byref(est._hash._desc), M))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)

    def sph():
        # the estimator's path for x32 keys and uniform masses: queries ordered
        # by cell column, one column's candidates staged in LDS per wave
        L.check(lib.jt_sph_eval_tiled(C.byref(est._hash._desc), L.ptr(Qd), M, C.c_float(radius), 64, 512,
                                      C.c_float(F(est._mean_h)), L.JT_SPH["cubic_spline"], L.ptr(rho), L.ptr(ws),
                                      ws_bytes, L.stream_handle()), "jt_sph_eval_tiled")
    sph()
    torch.cuda.synchronize()
    e0.record()
    sph()
    e1.record()
    torch.cuda.synchronize()
    sph_ms = e0.elapsed_time(e1)
    idx = torch.from_numpy(np.random.default_rng(12).choice(M, sample, replace=False)).to(dev)
    Qs = Qd[idx].contiguous()
    st = torch.empty((sample, 2), dtype=torch.int32, device=dev)
    L.check(lib.jt_hashgrid_query_stats(C.byref(est._hash._desc), L.ptr(Qs), sample, C.c_float(radius), 64,
                                        L.ptr(st), L.stream_handle()), "jt_hashgrid_query_stats")
    cq = float(st[:, 0].double().mean())
    # x32 keys: the direct pass reads the 9 distinct probe cells of the
    # 27-probe stencil once (the other 18 repeat them)
    bpq = 9 * 8 + cq / 3 * 16 + 4
    out["sph"] = {"value": M / (sph_ms * 1e-3), "unit": "queries/s", "ms": sph_ms, "queries": M,
                  "estimator_setup_ms": build_ms,
                  "hashgrid_build": {"ms": hg_ms,
                                     "roofline": hbm_roofline("hashgrid_build", len(P), hg_ms,
                                                              model={"bytes_per_point": 36,
                                                                     "GBs": 36.0 * len(P) / (hg_ms * 1e-3) / 1e9})},
                  "roofline": hbm_roofline("sph_eval", M, sph_ms,
                                           candidate_model={"bytes_per_query": bpq, "GBs": bpq * M / (sph_ms * 1e-3) / 1e9,
                                                            "mean_candidates": cq,
                                                            "mean_within_radius": float(st[:, 1].double().mean())}),
                  "workload": "2^24 Gaussian-mixture particles, 256^3 query grid, cubic spline h=2/256, x32 keys"}
    del est, rho, ws
    if CPU_BASELINES["on"]:
        qs = np.random.default_rng(15).choice(M, 32768, replace=False)
        hb, sb = cpu_sph_baseline(P, Qd[torch.from_numpy(qs).to(dev)].cpu().numpy(), h)
        out["sph"]["cpu_baseline"] = sb
        out["sph"]["hashgrid_build"]["cpu_baseline"] = hb
    if "kde" in parts:
        out = _bench_kde(torch, jt, L, lib, dev, P, Pd, kde_res, out)
    if "kde2d" in parts:
        out = _bench_kde2d(torch, jt, L, lib, dev, P, out)
    if "kde3d" in parts:
        out = _bench_kde3d(torch, jt, L, lib, dev, P, out)
    if "kdepts" in parts:
        out = _bench_kde_points(torch, jt, L, lib, dev, P, out)
    return out


def _bench_kde_points(torch, jt, L, lib, dev, P, out, res=50):
    """KDEEstimator.evaluate(points) (kde.py:128-184, 244-270) as the
    reference's workflow calls it: a res x res grid of (x, y) points on the
    z = 0.5 plane, given as 3-D points, over the 2^24 mixture particles,
    every pair: jt_kde_eval_ws (the particles split over enough blocks to
    fill the chip, float64 chunk sums added in order; the estimator's path)
    beside jt_kde_eval (one thread per point over every particle)."""
    F = np.float32
    kde = jt.KDEEstimator(torch.from_numpy(P).to(dev), normalize=False)
    a = np.linspace(0.0, 1.0, res, dtype=F)
    X, Y = np.meshgrid(a, a)
    Q = np.stack([X.ravel(), Y.ravel(), np.full(res * res, 0.5, F)], 1).astype(F)
    e0, e1 = _events(torch)
    kde.evaluate(Q)
    Qd = torch.from_numpy(Q).to(dev)
    Pd = kde._dev_points(kde.P)
    n, m = len(P), len(Q)
    ws_bytes = int(lib.jt_kde_eval_workspace_size(n, m))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    o1 = torch.empty(m, dtype=torch.float32, device=dev)
    o2 = torch.empty(m, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    e0.record()
    L.check(lib.jt_kde_eval_ws(L.ptr(Pd), n, 3, L.ptr(Qd), m, C.c_float(F(kde.h)), L.ptr(o1), L.ptr(ws), ws_bytes,
                               L.stream_handle()), "jt_kde_eval_ws")
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    e0.record()
    L.check(lib.jt_kde_eval(L.ptr(Pd), n, 3, L.ptr(Qd), m, C.c_float(F(kde.h)), L.ptr(o2), L.stream_handle()),
            "jt_kde_eval")
    e1.record()
    torch.cuda.synchronize()
    ref_ms = e0.elapsed_time(e1)
    a1, a2 = o1.cpu().numpy(), o2.cpu().numpy()
    # 28 VALU instructions per pair in the kernel's inner loop (3-D: the
    # differences, squares and sums, the scale, expf's range reduction and
    # exp, the float64 conversion and add), against MI355X's published FP32
    # vector rate counted in lane-operations (157.3 TFLOP/s = 78.65 T FMA)
    valu_per_pair, peak_tops = 28, 78.65
    ach = valu_per_pair * float(n) * m / (ms * 1e-3) / 1e12
    out["kde_points"] = {"value": float(n) * m / (ms * 1e-3), "unit": "pairs/s", "ms": ms, "points": m,
                         "particles": n, "one_thread_per_point_ms": ref_ms,
                         "roofline": {"bound": "valu", "achieved": ach, "peak": peak_tops, "unit": "T lane-ops/s",
                                      "frac": ach / peak_tops, "valu_per_pair": valu_per_pair,
                                      "note": "VALU instructions per pair counted in the built kernel's inner loop; "
                                              "peak: the published FP32 vector rate as lane-operations"},
                         "max_rel_diff_vs_one_thread_per_point": float(np.max(np.abs(a1 - a2) / np.maximum(a2, 1e-30))),
                         "parity_test": "tests/test_gpu_density.py::test_kde_eval_chunked",
                         "workload": f"evaluate() at a {res}x{res} grid of points on the z = 0.5 plane over the 2^24 "
                                     f"Gaussian-mixture particles, Scott bandwidth, every pair"}
    if CPU_BASELINES["on"]:
        out["kde_points"]["cpu_baseline"] = cpu_kde_baseline(P[:1 << 16], Q, float(kde.h),
                                                             total=f"all 2^24 x {m} pairs")
    return out


def _bench_kde3d(torch, jt, L, lib, dev, P, out, res=100):
    """KDEEstimator.evaluate_3d's default (no cut-off: every pair, kde.py:223-240,
    244-270) at the reference's default resolution 100 over the 2^24 mixture
    particles, Scott bandwidth: jt_kde_grid3d (the 16 x 16 x 32 tile kernel
    over chunks of the particles, bf16 hi/lo products on the matrix cores;
    timed: the call) beside jt_kde_eval (one thread per grid point over every
    particle).  Roofline: 8 flops per pair against the dense bf16 MFMA peak
    (the tiles' padding past the grid's end is not counted)."""
    F = np.float32
    kde = jt.KDEEstimator(torch.from_numpy(P).to(dev), resolution=res, bounds=((0, 1),) * 3, normalize=False)
    e0, e1 = _events(torch)
    kde.density_grid_3d()
    torch.cuda.synchronize()
    e0.record()
    D = kde.density_grid_3d()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    n = len(P)
    pairs = float(n) * res ** 3
    x, y, z = kde._axes_3d()
    X, Y, Z = np.meshgrid(x, y, z, indexing="xy")
    Gd = torch.from_numpy(np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(F)).to(dev)
    ref = torch.empty(res ** 3, dtype=torch.float32, device=dev)
    e0.record()
    L.check(lib.jt_kde_eval(L.ptr(kde._dev_points(kde.P)), n, 3, L.ptr(Gd), res ** 3, C.c_float(F(kde.h)), L.ptr(ref),
                            L.stream_handle()), "jt_kde_eval")
    e1.record()
    torch.cuda.synchronize()
    ref_ms = e0.elapsed_time(e1)
    Dn, Rn = D.cpu().numpy().ravel(), ref.cpu().numpy()
    m = Rn >= 1e-3 * Rn.max()
    ach = 8.0 * pairs / (ms * 1e-3) / 1e12
    out["kde_3d_all_pairs"] = {
        "value": pairs / (ms * 1e-3), "unit": "pairs/s", "ms": ms, "grid": [res] * 3, "particles": n, "pairs": pairs,
        "bandwidth": float(kde.h), "per_point_kernel_ms": ref_ms,
        "max_rel_diff_vs_per_point_kernel_above_floor": float(np.max(np.abs(Dn[m] - Rn[m]) / Rn[m])),
        "parity_test": "tests/test_gpu_density.py::test_full_size_kde_grid3d",
        "roofline": {"bound": "mfma", "achieved": ach, "peak": MFMA_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / MFMA_BF16_TFLOPS, "flops_per_pair": 8,
                     "note": "dense bfloat16 MFMA peak; every pair summed, each term as four bf16 products"},
        "workload": f"2^24 Gaussian-mixture particles, {res}^3 grid (the reference's default resolution), "
                    f"Scott bandwidth, all pairs"}
    if CPU_BASELINES["on"]:
        G = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        out["kde_3d_all_pairs"]["cpu_baseline"] = cpu_kde_baseline(
            P[:1 << 16], G[np.random.default_rng(18).choice(len(G), 4096, replace=False)], float(kde.h),
            total=f"all 2^24 x {res}^3 pairs")
    return out


def _bench_kde2d(torch, jt, L, lib, dev, P, out, res=256):
    """KDEEstimator.evaluate_2d's all-pairs sum (kde.py:205-221, 244-270) on
    the (x, y) of the 2^24 mixture particles, a res^2 grid over the unit
    square, Scott bandwidth of those 2-D points: jt_kde_grid2d (both factors
    per particle, bf16 hi/lo, on the matrix cores; timed: the call, tile
    kernel and slice sum) beside jt_kde_eval (one thread per grid point over
    every particle, float64 sums).  Every pair is summed, so the pair count
    is n x res^2; roofline: 8 flops per pair (four bfloat16 products) against
    the dense bf16 MFMA peak."""
    F = np.float32
    P2 = np.ascontiguousarray(P[:, :2])
    kde = jt.KDEEstimator(torch.from_numpy(P2).to(dev), resolution=res, bounds=((0, 1), (0, 1)), normalize=False)
    x = np.linspace(0.0, 1.0, res, dtype=F)
    e0, e1 = _events(torch)
    kde.density_grid_2d(x, x, P2)
    torch.cuda.synchronize()
    e0.record()
    D = kde.density_grid_2d(x, x, P2)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    n = len(P2)
    pairs = float(n) * res * res
    # the one-thread-per-point kernel on the same grid, for comparison
    X, Y = np.meshgrid(x, x, indexing="xy")
    Gd = torch.from_numpy(np.stack([X.ravel(), Y.ravel()], 1).astype(F)).to(dev)
    Pd2 = kde._dev_points(P2)
    ref = torch.empty(res * res, dtype=torch.float32, device=dev)
    e0.record()
    L.check(lib.jt_kde_eval(L.ptr(Pd2), n, 2, L.ptr(Gd), res * res, C.c_float(F(kde.h)), L.ptr(ref),
                            L.stream_handle()), "jt_kde_eval")
    e1.record()
    torch.cuda.synchronize()
    ref_ms = e0.elapsed_time(e1)
    Dn, Rn = D.cpu().numpy().ravel(), ref.cpu().numpy()
    m = Rn >= 1e-3 * Rn.max()
    ach = 8.0 * pairs / (ms * 1e-3) / 1e12
    out["kde_2d"] = {"value": pairs / (ms * 1e-3), "unit": "pairs/s", "ms": ms, "grid": [res, res], "particles": n,
                     "pairs": pairs, "bandwidth": float(kde.h),
                     "per_point_kernel_ms": ref_ms,
                     "max_rel_diff_vs_per_point_kernel_above_floor": float(np.max(np.abs(Dn[m] - Rn[m]) / Rn[m])),
                     "parity_test": "tests/test_gpu_density.py::test_full_size_kde_grid2d",
                     "roofline": {"bound": "mfma", "achieved": ach, "peak": MFMA_BF16_TFLOPS, "unit": "TFLOP/s",
                                  "frac": ach / MFMA_BF16_TFLOPS, "flops_per_pair": 8,
                                  "hbm": hbm_roofline("k_kde2d_tile", 1, ms),
                                  "note": "dense bfloat16 MFMA peak; every pair is summed, each term as the four "
                                          "bf16 products of the split float32 factors"},
                     "workload": f"the 2^24 Gaussian-mixture particles' (x, y), {res}^2 grid on the unit square, "
                                 f"Scott bandwidth, all pairs"}
    if CPU_BASELINES["on"]:
        out["kde_2d"]["cpu_baseline"] = cpu_kde_baseline(
            P2[:1 << 16], np.stack([X.ravel(), Y.ravel()], 1)[np.random.default_rng(17).choice(res * res, 8192,
                                                                                                replace=False)],
            float(kde.h), total=f"all 2^24 x {res}^2 pairs")
    return out


def _bench_kde(torch, jt, L, lib, dev, P, Pd, kde_res, out):
    """The KDE half of bench_density."""
    F = np.float32
    e0, e1 = _events(torch)
    kde = jt.KDEEstimator(Pd, bandwidth_rule="scott", normalize=True, resolution=kde_res,
                          bounds=((0, 1), (0, 1), (0, 1)), cutoff_sigmas=KDE_CUTOFF_SIGMAS)
    x, y, z = kde._axes_3d()
    R = F(KDE_CUTOFF_SIGMAS * kde.h)
    hg = jt.HashGridNeighbors(Pd, float(R) / 2, 1, "x64", dev)
    gx, gy, gz = (torch.from_numpy(np.ascontiguousarray(v)).to(dev) for v in (x, y, z))
    D = torch.empty((len(y), len(x), len(z)), dtype=torch.float32, device=dev)

    ws_bytes = int(lib.jt_kde_grid_cutoff_workspace_size(len(x), len(y), len(z)))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)

    def kd():
        L.check(lib.jt_kde_grid_cutoff_ws(C.byref(hg._desc), kde.N, L.ptr(gx), len(x), L.ptr(gy), len(y), L.ptr(gz),
                                          len(z), C.c_float(F(kde.h)), C.c_float(R), L.ptr(D), L.ptr(ws), ws_bytes,
                                          L.stream_handle()), "jt_kde_grid_cutoff_ws")
    kd()
    torch.cuda.synchronize()
    e0.record()
    kd()
    e1.record()
    torch.cuda.synchronize()
    kde_ms = e0.elapsed_time(e1)
    d_max = float(D.max())
    tail = (2 * math.pi) ** -1.5 * math.exp(-0.5 * KDE_CUTOFF_SIGMAS ** 2) / (1e-3 * d_max)
    Xk, Yk, Zk = np.meshgrid(x, y, z, indexing="xy")
    Qk = np.stack([Xk.ravel(), Yk.ravel(), Zk.ravel()], 1).astype(F)
    mk = len(Qk)
    sel = np.random.default_rng(13).choice(mk, min(2048, mk), replace=False)
    Qks = torch.from_numpy(Qk[sel]).to(dev)
    stk = torch.empty((len(sel), 2), dtype=torch.int32, device=dev)
    hgs = jt.HashGridNeighbors(Pd, float(R), 1, "x64", dev)        # cells of R: the 27 probes cover R
    L.check(lib.jt_hashgrid_query_stats(C.byref(hgs._desc), L.ptr(Qks), len(sel), C.c_float(R), 1 << 30,
                                        L.ptr(stk), L.stream_handle()), "jt_hashgrid_query_stats")
    pairs = float(stk[:, 1].double().mean()) * mk
    rate = pairs / (kde_ms * 1e-3)
    del ws
    # tile pairs: (grid points of a 16 x 16 x 32 tile) x (particles within R
    # of the tile's box), the products the kernel's matrix cores evaluate,
    # summed over 512 tiles drawn at random (seed 14) and scaled to all tiles
    ext = (16, 16, 32)
    nt = [(len(v) + e - 1) // e for v, e in zip((x, y, z), ext)]
    n_tiles = nt[0] * nt[1] * nt[2]
    trng = np.random.default_rng(14)
    pts = 0.0
    for _ in range(512):
        tb = [int(trng.integers(0, k)) for k in nt]
        seg = [v[t * e:min(t * e + e, len(v))] for v, t, e in zip((x, y, z), tb, ext)]
        lo_t = torch.tensor([float(g_.min()) for g_ in seg], device=dev)
        hi_t = torch.tensor([float(g_.max()) for g_ in seg], device=dev)
        d = torch.clamp(torch.maximum(lo_t - Pd, Pd - hi_t), min=0.0)
        inside = int(((d * d).sum(1) <= float(R) * float(R)).sum())
        pts += inside * len(seg[0]) * len(seg[1]) * len(seg[2])
    tile_pairs = pts * n_tiles / 512
    # 4 bfloat16 products (hi/lo halves of both operands) of 2 flops per tile pair
    ach = 8.0 * tile_pairs / (kde_ms * 1e-3) / 1e12
    out["kde"] = {"value": rate, "unit": "pairs/s", "ms": kde_ms, "grid": [kde_res] * 3, "pairs_within_cutoff": pairs,
                  "bandwidth": float(kde.h), "cutoff": float(R), "cutoff_sigmas": KDE_CUTOFF_SIGMAS,
                  "cutoff_parity_test": "tests/test_gpu_density.py::test_kde_cutoff_tiles_sampled",
                  "dropped_tail_bound_rel_at_floor": tail,
                  "roofline": {"bound": "mfma", "achieved": ach, "peak": MFMA_BF16_TFLOPS, "unit": "TFLOP/s",
                               "frac": ach / MFMA_BF16_TFLOPS, "tile_pairs": tile_pairs,
                               "tile_pairs_sample": "512 random 16 x 16 x 32 tiles (seed 14), scaled to all tiles",
                               "hbm": hbm_roofline("k_kde_tile_mfma", 1, kde_ms),
                               "flops_per_tile_pair": 8,
                               "note": "dense bfloat16 MFMA peak; each float32 operand is split into bfloat16 hi/lo "
                                       "halves and the four products are summed in float32"},
                  "workload": f"2^24 Gaussian-mixture particles, {kde_res}^3 grid, Scott bandwidth, "
                              f"cut-off {KDE_CUTOFF_SIGMAS} h"}
    if CPU_BASELINES["on"]:
        out["kde"]["cpu_baseline"] = cpu_kde_baseline(P[:1 << 16], Qk[np.random.default_rng(16).choice(mk, 4096,
                                                                                                      replace=False)],
                                                      float(kde.h))
    return out


def copy_calibration(torch):
    """A plain copy of 2^28 floats (1 GiB read + 1 GiB written), five times."""
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": a.numel() * 4, "copies": 5}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 "
                                   "(tracking shards particles with no exchange, so the "
                                   "one-GPU number is one shard at its real shape)"}))
        sys.exit(2)
    CPU_BASELINES["on"] = not args.no_cpu_baseline
    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    import jaxtrace_amd as jt
    from jaxtrace_amd import _lib as L
    from jaxtrace_amd import workloads as W
    lib = L.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(0)
    if args.only and args.only != "track":
        fn = {"adjoint": lambda: bench_adjoint(torch, jt, W, dev, reps=args.adjoint_reps),
              "octree": lambda: bench_octree(torch, jt, W, L, lib, dev),
              "sph": lambda: bench_density(torch, jt, W, L, lib, dev, parts=("sph",)),
              "kde": lambda: bench_density(torch, jt, W, L, lib, dev, parts=("kde",)),
              "kde2d": lambda: bench_density(torch, jt, W, L, lib, dev, parts=("kde2d",)),
              "kde3d": lambda: bench_density(torch, jt, W, L, lib, dev, parts=("kde3d",)),
              "kdepts": lambda: bench_density(torch, jt, W, L, lib, dev, parts=("kdepts",)),
              "hashgrid": lambda: bench_density(torch, jt, W, L, lib, dev, parts=("hashgrid",)),
              "nearest": lambda: bench_nearest(torch, jt, W, L, lib, dev),
              "seeding": lambda: bench_seeding(torch, jt, W, L, lib, dev),
              "vtk": lambda: bench_vtk(torch, jt, W, L, lib, dev)}[args.only]
        print(json.dumps({"only": args.only, "result": fn()}), flush=True)
        return

    t_gen = time.perf_counter()
    data, times, grid = W.random_fourier(n=GRID_N, n_snap=N_SNAP, t_end=2.0, seed=2, device=dev)
    g = torch.Generator(device="cpu").manual_seed(3)
    x0_host = torch.rand((args.particles, 3), generator=g, dtype=torch.float32).numpy()
    x0 = torch.from_numpy(x0_host).to(dev)
    field32 = jt.StructuredTimeSeriesField(data, times, grid, grid, grid, bounds_handling="periodic",
                                           layout="halo")
    bc = jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], {"x": "periodic", "y": "reflective", "z": "reflective"})
    gen_s = time.perf_counter() - t_gen

    wall32, durs32, xf32 = measure(field32, bc, x0, args, torch, L, lib, args.sort_every)
    n = args.particles
    rate32 = n * args.steps / wall32
    ms_step = wall32 * 1e3 / args.steps
    spl = launch_steps(args, args.sort_every)
    full = [r for r in durs32 if r[2] == spl] or durs32
    avg_call_ms = float(np.mean([r[0] for r in full]))
    avg_launch_ms = float(np.mean([r[1] for r in full]))
    steps_per = full[0][2]
    bpps32 = algorithmic_bytes_per_particle_step(4)
    gather32 = bpps32 * n * steps_per / (avg_launch_ms * 1e-3) / 1e9

    result = {
        "metric": "RK4 particle-steps/s at 16M particles, FP32 vs BF16 field; adjoint/fwd ratio",
        "value": rate32,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded random Fourier field generated on device, uniform particles)",
        "config": {"workload": "random-Fourier 256^3 x 16 snapshots, RK4, dt=2e-3, "
                               "x periodic / y,z reflective, FP32 field",
                   "particles": n, "grid": [GRID_N] * 3, "snapshots": N_SNAP, "dt": DT,
                   "steps_per_launch": spl, "cell_sort_every": args.sort_every,
                   "snapshot_layout": "z-halo, 4 components per node" + (
                       ", and consecutive-pair records (both snapshots' z-pairs of a node in 48 B)"
                       if field32._desc.pair_data else "")},
        "roofline": hbm_roofline(
            "k_track_rk4_f32", n * steps_per, avg_launch_ms,
            kernel="k_track<RK4,f32,pair records,fast path>" if field32._desc.pair_data else
            "k_track<RK4,f32x4,z-halo,fast path>",
            launch_ms=avg_launch_ms, call_ms_with_cell_sort=avg_call_ms, particle_steps_per_launch=n * steps_per,
            gather_model={"bytes_per_particle_step": bpps32, "GBs": gather32,
                          "note": "24 B of state + 4 stages x 2 snapshots x 8 corners x 12 B: the corner "
                                  "gathers the algorithm makes, most of them served by registers, L1 and L2"}),
    }

    if not args.no_bf16:
        field16 = field32.with_storage("bfloat16", layout="halo")
        wall16, durs16, xf16 = measure(field16, bc, x0, args, torch, L, lib, args.sort_every)
        rate16 = n * args.steps / wall16
        full16 = [r for r in durs16 if r[2] == spl] or durs16
        avg16 = float(np.mean([r[1] for r in full16]))
        g16 = algorithmic_bytes_per_particle_step(2) * n * full16[0][2] / (avg16 * 1e-3) / 1e9
        result["bf16"] = {"value": rate16, "speedup_vs_f32": rate16 / rate32, "launch_ms": avg16,
                          "roofline": hbm_roofline("k_track_rk4_bf16", n * full16[0][2], avg16,
                                                   gather_model={"bytes_per_particle_step":
                                                                 algorithmic_bytes_per_particle_step(2),
                                                                 "GBs": g16})}
        del field16

    if args.only == "track":
        result["rk2_euler"] = bench_integrators(torch, L, lib, field32, bc, x0, args, steps=50, warmup=25)
        print(json.dumps(result), flush=True)
        return
    if not args.no_extras:
        try:
            result["rk2_euler"] = bench_integrators(torch, L, lib, field32, bc, x0, args)
        except Exception as e:  # reported, never fatal
            result["rk2_euler"] = {"error": str(e)}
        try:
            result["trajectory_stream"] = bench_stream(torch, jt, field32, bc, x0)
        except Exception as e:  # reported, never fatal
            result["trajectory_stream"] = {"error": str(e)}

    # trajectory error of the GPU output against float64 over the whole run
    try:
        if args.no_extras:
            raise RuntimeError("skipped (--no-extras)")
        result["trajectory_err_vs_fp64"] = trajectory_error_vs_fp64(torch, jt, data, times, grid, x0, bc)
    except Exception as e:  # reported, never fatal
        result["trajectory_err_vs_fp64"] = {"error": str(e)}

    if not args.no_cpu_baseline:
        host = data.cpu().numpy()
        npart = min(args.cpu_particles, n)
        rate_cpu, el = cpu_baseline(host, times, grid, x0_host, npart, args.cpu_steps)
        result["cpu_baseline"] = {
            "value": rate_cpu, "unit": "particle-steps/s", "cores": 1, "kind": "port",
            "sample": f"{npart} particles x {args.cpu_steps} RK4 steps of the same workload "
                      f"(NumPy restatement of the reference algorithm: every node blended per "
                      f"stage, then trilinear), {el:.1f} s",
            **cpu_host_info()}
    try:
        if args.no_adjoint:
            raise RuntimeError("skipped (--no-adjoint)")
        adj = bench_adjoint(torch, jt, W, dev)
        result["adjoint"] = adj
        result["adjoint_fwd_ratio"] = adj["float32"]["ratio_bwd_over_fwd"]
    except Exception as e:  # reported, never fatal
        result["adjoint"] = {"error": str(e)}
        result["adjoint_fwd_ratio"] = None
    if not args.no_extras:
        for key, fn in (("octree", bench_octree), ("density", bench_density), ("nearest_node", bench_nearest),
                        ("seeding", bench_seeding), ("vtk_ingestion", bench_vtk)):
            try:
                result[key] = fn(torch, jt, W, L, lib, dev)
            except Exception as e:  # reported, never fatal
                result[key] = {"error": str(e)}
    result["setup_s"] = gen_s
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
