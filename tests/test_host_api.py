"""Host-side behaviour of the reference-shaped API that needs no GPU:
argument checking and error types (tracker.py:810-855, boundary.py:39-86,
377-438, 1444-1601)."""
import numpy as np
import pytest

import jaxtrace_amd as jt
from jaxtrace_amd import _lib as L


def test_unknown_integrator_is_value_error():
    with pytest.raises(ValueError):
        jt.create_tracker("rk5", object(), jt.periodic_boundary([[0, 0, 0], [1, 1, 1]]))


def test_python_callables_are_rejected_with_type_error():
    bc = jt.periodic_boundary([[0, 0, 0], [1, 1, 1]])
    with pytest.raises(TypeError):
        jt.create_tracker("rk4", lambda x, t: x, bc)


def test_bounds_validation():
    with pytest.raises(ValueError):
        jt.periodic_boundary([[0, 0, 0], [1, 0, 1]])
    with pytest.raises(ValueError):
        jt.periodic_boundary([0, 1, 2])
    b = jt.reflective_boundary([0, 1, 0, 2, 0, 3])
    np.testing.assert_array_equal(b.bounds, [[0, 0, 0], [1, 2, 3]])
    assert b.__name__ == "_reflective_bc"


def test_mixed_boundary_kinds_and_errors():
    b = jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], {"y": "reflective", "z": "clamping"})
    assert b.axis_kind == (L.JT_BC_PERIODIC, L.JT_BC_REFLECTIVE, L.JT_BC_CLAMP)
    with pytest.raises(ValueError):
        jt.mixed_boundary([[0, 0, 0], [1, 1, 1]], {"x": "sticky"})


def test_continuous_inlet_descriptor():
    with pytest.raises(ValueError):
        jt.continuous_inlet_boundary_factory([[0, 0, 0], [1, 1, 1]], flow_axis="w")
    b = jt.continuous_inlet_boundary_factory([[0, 0, 0], [1, 1, 1]], "x", "positive", "grid",
                                             {"x": 64, "y": 64, "z": 32})
    assert b.inlet["mode"] == L.JT_INLET_GRID and b.inlet["axis"] == 0
    # the reference passes f32 bound scalars, so NumPy 2 evaluates linspace in f32
    f0, f1 = np.float32(0), np.float32(1)
    np.testing.assert_array_equal(b.inlet["coords"][1], np.linspace(f0, f1, 32, dtype=np.float32))
    assert b.inlet["backstop"] == np.float32(0.001)
    for dist in ("uniform", "gaussian"):
        r = jt.continuous_inlet_boundary_factory([[0, 0, 0], [1, 1, 1]], inlet_distribution=dist)
        assert r.inlet["mode"] == L.JT_INLET_HOST and r.host_inlet and r.inlet["distribution"] == dist
    assert not b.host_inlet


def test_tracker_options_defaults_match_reference():
    o = jt.TrackerOptions()
    assert o.max_batch_size == 100_000 and o.min_batch_size == 100 and o.max_memory_gb == 8.0
    assert o.record_every == 1 and not o.return_device


def test_streamed_frame_writer_ranges_and_layout(tmp_path):
    """The streamed-output writer puts frame ranges at their byte offsets of a
    standard (frames, n, 3) float32 .npy file, in any order."""
    from jaxtrace_amd.tracking import _NpyFrames
    rng = np.random.default_rng(0)
    want = rng.standard_normal((7, 5, 3)).astype(np.float32)
    w = _NpyFrames(str(tmp_path / "t.npy"), 7, 5)
    w.write(3, want[3:7])
    w.write(0, want[0])
    w.write(1, want[1:3])
    w.close()
    got = np.load(tmp_path / "t.npy")
    assert got.dtype == np.float32 and got.shape == (7, 5, 3)
    np.testing.assert_array_equal(got, want)
    e = _NpyFrames(str(tmp_path / "e.npy"), 4, 0)
    e.write(0, np.zeros((1, 0, 3), np.float32))
    e.write(1, np.zeros((3, 0, 3), np.float32))
    e.close()
    assert np.load(tmp_path / "e.npy").shape == (4, 0, 3)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says why."""
    import json
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]
