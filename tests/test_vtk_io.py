"""VTK XML unstructured-grid ingestion (jaxtrace_amd.vtk_io), on the CPU.

Pinned three ways: hand-written files whose bytes follow the VTK XML format
directly (ascii cells of mixed types; an inline base64 array with a UInt32
byte-count header; a big-endian UInt64-header array), write -> read round
trips over every encoding the reader takes, and the reference loader's file
rules (vtk_reader.py:496-541 discovery and sort, example_workflow.py:436-610
stable-mesh skip, stride and "_<n>.pvtu" times) on small series of files.
"""
import os

import numpy as np
import pytest

from jaxtrace_amd import vtk_io as V
from jaxtrace_amd import workloads as W

F32 = np.float32


def _mesh(n=3):
    pts, conn = W.kuhn_mesh(n, 0.2, 1)
    return pts.astype(F32), conn.astype(np.int32)


HAND_ASCII = """<?xml version="1.0"?>
<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian">
  <UnstructuredGrid>
    <Piece NumberOfPoints="6" NumberOfCells="3">
      <Points>
        <DataArray type="Float32" NumberOfComponents="3" format="ascii">
          0 0 0  1 0 0  0 1 0  0 0 1  1 1 1  2 0 0
        </DataArray>
      </Points>
      <Cells>
        <DataArray type="Int32" Name="connectivity" format="ascii">0 1 2 3  1 2 4  1 5 2 4</DataArray>
        <DataArray type="Int32" Name="offsets" format="ascii">4 7 11</DataArray>
        <DataArray type="UInt8" Name="types" format="ascii">10 5 10</DataArray>
      </Cells>
      <PointData Vectors="Velocity">
        <DataArray type="Float64" Name="Velocity" NumberOfComponents="2" format="ascii">
          1 2 3 4 5 6 7 8 9 10 11 12
        </DataArray>
      </PointData>
    </Piece>
  </UnstructuredGrid>
</VTKFile>
"""


def test_hand_written_ascii_mixed_cells(tmp_path):
    f = tmp_path / "hand.vtu"
    f.write_text(HAND_ASCII)
    g = V.read_vtu(str(f))
    assert g["points"].shape == (6, 3) and g["points"].dtype == F32
    np.testing.assert_array_equal(g["points"][4], [1, 1, 1])
    np.testing.assert_array_equal(g["types"], [10, 5, 10])
    # the triangle (type 5) is dropped, the tets keep their point order
    np.testing.assert_array_equal(V.tetra_connectivity(g), [[0, 1, 2, 3], [1, 5, 2, 4]])
    v = g["point_data"]["Velocity"]
    assert v.shape == (6, 2) and v.dtype == np.float64
    np.testing.assert_array_equal(V._ensure_3d(v)[2], [5, 6, 0])


def test_hand_written_binary_headers(tmp_path):
    # 'CAAAAAAAgD8AAABA' = base64 of 08 00 00 00 | 00 00 80 3f | 00 00 00 40:
    # a little-endian UInt32 byte count of 8 and the Float32 values 1.0, 2.0.
    # 'AAAAAAAAAAwAAAAAAAAAAQAAAAI=' = a big-endian UInt64 count of 12 and the
    # big-endian Int32 values 0, 1, 2.
    le = """<VTKFile type="UnstructuredGrid" byte_order="LittleEndian" header_type="UInt32"><UnstructuredGrid>
<Piece NumberOfPoints="2" NumberOfCells="0"><Points>
<DataArray type="Float32" NumberOfComponents="3" format="ascii">0 0 0 1 1 1</DataArray></Points>
<PointData><DataArray type="Float32" Name="s" format="binary">
  CAAAAAAAgD8AAABA
</DataArray></PointData></Piece></UnstructuredGrid></VTKFile>"""
    be = """<VTKFile type="UnstructuredGrid" byte_order="BigEndian" header_type="UInt64"><UnstructuredGrid>
<Piece NumberOfPoints="3" NumberOfCells="0"><Points>
<DataArray type="Float32" NumberOfComponents="3" format="ascii">0 0 0 1 1 1 2 2 2</DataArray></Points>
<PointData><DataArray type="Int32" Name="ids" format="binary">AAAAAAAAAAwAAAAAAAAAAQAAAAI=</DataArray>
</PointData></Piece></UnstructuredGrid></VTKFile>"""
    (tmp_path / "le.vtu").write_text(le)
    (tmp_path / "be.vtu").write_text(be)
    np.testing.assert_array_equal(V.read_vtu(str(tmp_path / "le.vtu"))["point_data"]["s"], [1.0, 2.0])
    ids = V.read_vtu(str(tmp_path / "be.vtu"))["point_data"]["ids"]
    np.testing.assert_array_equal(ids, [0, 1, 2])
    assert V.tetra_connectivity(V.read_vtu(str(tmp_path / "be.vtu"))).shape == (0, 4)


ENCODINGS = [("ascii", "UInt32", False, False)] + [
    (enc, ht, comp, be) for enc in ("binary", "appended", "appended-base64")
    for ht in ("UInt32", "UInt64") for comp in (False, True) for be in (False, True)]


@pytest.mark.parametrize("enc,header,compress,big_endian", ENCODINGS)
def test_roundtrip_every_encoding(tmp_path, enc, header, compress, big_endian):
    pts, conn = _mesh(3)
    rng = np.random.default_rng(1)
    vel = rng.standard_normal((len(pts), 3)).astype(F32)
    s64 = rng.standard_normal(len(pts))
    f = str(tmp_path / "g.vtu")
    V.write_vtu(f, pts, conn, {"Velocity": vel, "s": s64}, encoding=enc, header_type=header,
                compress=compress, big_endian=big_endian)
    g = V.read_vtu(f)
    np.testing.assert_array_equal(g["points"], pts)
    np.testing.assert_array_equal(V.tetra_connectivity(g), conn)
    np.testing.assert_array_equal(g["point_data"]["Velocity"], vel)
    np.testing.assert_array_equal(g["point_data"]["s"], s64)


def test_compressed_multi_block(tmp_path):
    # more than one 32 KiB zlib block, the last one partial
    pts, conn = _mesh(20)
    vel = np.random.default_rng(2).standard_normal((len(pts), 3)).astype(F32)
    assert vel.nbytes > 2 * (1 << 15) and vel.nbytes % (1 << 15)
    for enc in ("binary", "appended", "appended-base64"):
        f = str(tmp_path / f"{enc}.vtu")
        V.write_vtu(f, pts, conn, {"Velocity": vel}, encoding=enc, compress=True)
        g = V.read_vtu(f)
        np.testing.assert_array_equal(g["point_data"]["Velocity"], vel)
        np.testing.assert_array_equal(V.tetra_connectivity(g), conn)


def test_pvtu_concatenates_pieces(tmp_path):
    pa, ca = _mesh(2)
    pb, cb = _mesh(3)
    pb = pb + F32(5)
    va = np.full((len(pa), 3), 1, F32)
    vb = np.full((len(pb), 3), 2, F32)
    os.makedirs(tmp_path / "parts")
    V.write_vtu(str(tmp_path / "parts" / "p0.vtu"), pa, ca, {"Velocity": va})
    V.write_vtu(str(tmp_path / "parts" / "p1.vtu"), pb, cb, {"Velocity": vb}, encoding="appended")
    V.write_pvtu(str(tmp_path / "s_0.pvtu"), ["parts/p0.vtu", "parts/p1.vtu"], ["Velocity"])
    g = V.read_pvtu(str(tmp_path / "s_0.pvtu"))
    np.testing.assert_array_equal(g["points"], np.concatenate([pa, pb]))
    np.testing.assert_array_equal(V.tetra_connectivity(g), np.concatenate([ca, cb + len(pa)]))
    np.testing.assert_array_equal(g["point_data"]["Velocity"], np.concatenate([va, vb]))
    with pytest.raises(ValueError):
        V.read_unstructured(str(tmp_path / "parts" / "p0.vtk"))


def _series(tmp_path, names, pts, conn, vel_of, field="Velocity", npts_of=None):
    for i, name in enumerate(names):
        p = pts if npts_of is None else pts[:npts_of(i)]
        c = conn if npts_of is None else conn[np.all(conn < len(p), axis=1)]
        V.write_vtu(str(tmp_path / name), p, c, {field: vel_of(i, len(p))})


def test_reader_discovery_natural_sort_and_pvtu_preference(tmp_path):
    pts, conn = _mesh(2)
    vel_of = lambda i, n: np.full((n, 3), i, F32)
    names = [f"flow_{k}.vtu" for k in (10, 2, 1, 33)]
    _series(tmp_path, names, pts, conn, vel_of)
    r = V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "flow_*.vtu"))
    assert [os.path.basename(f) for f in r.time_files] == ["flow_1.vtu", "flow_2.vtu", "flow_10.vtu",
                                                            "flow_33.vtu"]
    r = V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "flow_*.vtu"), max_time_steps=2)
    assert [os.path.basename(f) for f in r.time_files] == ["flow_10.vtu", "flow_33.vtu"]
    s = r.load_time_series()
    assert s["velocity_data"].shape == (2, len(pts), 3) and s["velocity_data"].dtype == F32
    np.testing.assert_array_equal(s["velocity_data"][:, 0, 0], [0, 3])      # written as indices 0 and 3
    np.testing.assert_array_equal(s["times"], [0, 1])
    assert s["velocity_field_name"] == "Velocity"
    assert r.get_bounds() == tuple(float(v) for v in np.concatenate([pts.min(0), pts.max(0)]))
    with pytest.raises(IndexError):
        r.load_single_timestep(2)
    # a .pvtu next to the .vtu files is preferred
    V.write_pvtu(str(tmp_path / "flow_5.pvtu"), ["flow_2.vtu"], ["Velocity"])
    r = V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "flow_*"))
    assert [os.path.basename(f) for f in r.time_files] == ["flow_5.pvtu"]
    np.testing.assert_array_equal(r[0], np.full((len(pts), 3), 1, F32))
    with pytest.raises(ValueError, match="No VTK files"):
        V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "none_*.vtu"))


def test_reader_velocity_detection_and_padding(tmp_path):
    pts, conn = _mesh(2)
    V.write_vtu(str(tmp_path / "a_0.vtu"), pts, conn,
                {"pressure": np.zeros(len(pts), F32), "U": np.ones((len(pts), 2), F32)})
    r = V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "a_*.vtu"))
    v = r.load_single_timestep(0)
    assert r.velocity_field_name == "U"
    np.testing.assert_array_equal(v[:, 2], 0)
    np.testing.assert_array_equal(v[:, :2], 1)
    V.write_vtu(str(tmp_path / "b_0.vtu"), pts, conn, {"pressure": np.zeros(len(pts), F32)})
    with pytest.raises(ValueError, match="Velocity field not found"):
        V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "b_*.vtu")).load_single_timestep(0)


def test_reader_cache_limit(tmp_path):
    pts, conn = _mesh(2)
    _series(tmp_path, [f"c_{k}.vtu" for k in range(5)], pts, conn, lambda i, n: np.full((n, 3), i, F32))
    r = V.VTKUnstructuredTimeSeriesReader(str(tmp_path / "c_*.vtu"), cache_size_limit=2)
    for k in range(5):
        r.load_single_timestep(k)
    assert sorted(r.velocity_cache) == [3, 4]


def test_octree_series_times_stride_and_skip(tmp_path):
    pts, conn = _mesh(3)
    names = [f"run_{k:04d}.pvtu" for k in range(0, 120, 10)]               # 12 files, 0 .. 110
    os.makedirs(tmp_path / "d")
    for i, name in enumerate(names):
        V.write_vtu(str(tmp_path / "d" / f"{i}.vtu"), pts, conn,
                    {"Displacement": np.full((len(pts), 3), i, F32), "Velocity": np.zeros((len(pts), 3), F32)})
        V.write_pvtu(str(tmp_path / name), [f"d/{i}.vtu"], ["Displacement"])
    s = V.read_octree_series(str(tmp_path / "run_*.pvtu"), max_timesteps=4, skip_initial_timesteps=1)
    # 11 left after the skip, stride 11 // 4 = 2: files 1, 3, 5, 7
    np.testing.assert_array_equal(s["times"], [10, 30, 50, 70])
    np.testing.assert_array_equal(s["velocity_data"][:, 0, 0], [1, 3, 5, 7])   # 'Displacement' first
    np.testing.assert_array_equal(s["positions"], pts)
    np.testing.assert_array_equal(s["connectivity"], conn)
    assert s["connectivity"].dtype == np.int32


def test_octree_series_stable_mesh_detection(tmp_path):
    pts, conn = _mesh(4)
    # the first three files are smaller (a refinement phase), then 9 of one size
    npts = lambda i: len(pts) - 7 if i < 3 else len(pts)
    names = [f"amr_{k:03d}.vtu" for k in range(12)]
    _series(tmp_path, names, pts, conn, lambda i, n: np.full((n, 3), i, F32), npts_of=npts)
    s = V.read_octree_series(str(tmp_path / "amr_*.vtu"), max_timesteps=40)
    assert len(s["files"]) == 9 and os.path.basename(s["files"][0]) == "amr_003.vtu"
    np.testing.assert_array_equal(s["times"], np.arange(9))                 # no "_<n>.pvtu": indices
    with pytest.raises(ValueError, match="Inconsistent mesh sizes"):
        V.read_octree_series(str(tmp_path / "amr_*.vtu"), use_stable_mesh_only=False)
    with pytest.raises(FileNotFoundError):
        V.read_octree_series(str(tmp_path / "none_*.vtu"))
