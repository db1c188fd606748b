"""jt_track_grid_noalloc (no workspace: the stage-time table lives in the
library's static device scratch) against jt_track_grid on the same inputs,
both through the C ABI: two calls in a row, the second with fewer particles
and steps (the scratch keeps the first call's longer table, which must not
leak into the second), every recorded position bitwise equal.  Also the
argument errors that need no kernel launch."""
import ctypes as C

import numpy as np
import pytest
import torch

import jaxtrace_amd as jt
from jaxtrace_amd import _lib as L
from jaxtrace_amd import workloads as W

pytestmark = pytest.mark.gpu
F32 = np.float32


def _calls(field, bc_obj, integ, x0, n_steps, dt, rec, dev):
    lib = L.load()
    desc = field.descriptor()
    bc = bc_obj.descriptor(dev)
    tg = np.linspace(0.0, n_steps * dt, n_steps + 1, dtype=F32)
    st = torch.from_numpy(tg[:-1].copy()).to(dev)
    n = x0.shape[0]
    n_rec = 1 + n_steps // rec
    o = L.TrackOpts()
    o.sort_every = 0
    xa = torch.from_numpy(x0).to(dev)
    xb = xa.clone()
    ta = torch.empty((n_rec, n, 3), dtype=torch.float32, device=dev)
    tb = torch.empty_like(ta)
    wsb = int(lib.jt_track_grid_workspace_size(n, n_steps, C.byref(o)))
    ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
    L.check(lib.jt_track_grid(C.byref(desc), C.byref(bc), integ, L.ptr(xa), n, L.ptr(st), C.c_float(F32(dt)),
                              n_steps, rec, L.ptr(ta), C.byref(o), L.ptr(ws), wsb, L.stream_handle()),
            "jt_track_grid")
    torch.cuda.synchronize()
    before = torch.cuda.memory_allocated(dev)
    L.check(lib.jt_track_grid_noalloc(C.byref(desc), C.byref(bc), integ, L.ptr(xb), n, L.ptr(st),
                                      C.c_float(F32(dt)), n_steps, rec, L.ptr(tb), C.byref(o), L.stream_handle()),
            "jt_track_grid_noalloc")
    torch.cuda.synchronize()
    assert torch.cuda.memory_allocated(dev) == before
    return xa.cpu().numpy(), xb.cpu().numpy(), ta.cpu().numpy(), tb.cpu().numpy()


def test_noalloc_matches_workspace_entry_point(gpu):
    data, times, g = W.abc_flow()
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic")
    bc = jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], {"x": "periodic", "y": "reflective", "z": "reflective"})
    # first call: 20,000 particles x 300 steps (RK4); second: 3,001 particles
    # x 41 steps (RK2) on the same GPU's scratch
    for n, steps, rec, integ, seed in ((20_000, 300, 7, L.JT_RK4, 1), (3_001, 41, 5, L.JT_RK2, 2)):
        x0 = W.uniform_particles(n, seed)
        xa, xb, ta, tb = _calls(f, bc, integ, x0, steps, 0.003, rec, gpu)
        np.testing.assert_array_equal(tb, ta)
        np.testing.assert_array_equal(xb, xa)
        assert not np.array_equal(ta[-1], ta[0])


def test_noalloc_rejects_what_it_cannot_do(gpu):
    lib = L.load()
    data, times, g = W.abc_flow()
    f = jt.StructuredTimeSeriesField(data, times, g, g, g, bounds_handling="periodic")
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]]).descriptor(gpu)
    x = torch.zeros((4, 3), dtype=torch.float32, device=gpu)
    st = torch.zeros(70_000, dtype=torch.float32, device=gpu)
    o = L.TrackOpts()
    o.sort_every = 0
    rc = lib.jt_track_grid_noalloc(C.byref(f.descriptor()), C.byref(bc), L.JT_RK4, L.ptr(x), 4, L.ptr(st),
                                   C.c_float(0.01), 65537, 1, None, C.byref(o), L.stream_handle())
    assert rc == L.JT_EINVAL and b"at most 65536 steps" in lib.jt_last_error()
    o.sort_every = 10
    rc = lib.jt_track_grid_noalloc(C.byref(f.descriptor()), C.byref(bc), L.JT_RK4, L.ptr(x), 4, L.ptr(st),
                                   C.c_float(0.01), 10, 1, None, C.byref(o), L.stream_handle())
    assert rc == L.JT_EUNSUPPORTED and b"sort" in lib.jt_last_error()
