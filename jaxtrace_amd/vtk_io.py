"""VTK XML unstructured-grid time series (.vtu / .pvtu) into the octree field.

Mirrors the reference's on-disk input path:
  * VTKUnstructuredTimeSeriesReader (jaxtrace/io/vtk_reader.py:426-942):
    file discovery (.pvtu preferred over .vtu, natural sort, the last
    max_time_steps), velocity-field auto-detection, load_single_timestep and
    load_time_series in the (T, N, 3) float32 layout with times = step indices;
  * the workflow's octree loader (example_workflow.py:436-660):
    points and VTK_TETRA connectivity from the first file, one velocity array
    per file ('Displacement' first), times from the "_<n>.pvtu" suffix, into
    OctreeFEMTimeSeriesFieldOptimized.

The reference reads through the `vtk` library, which is not part of this
build.  The files are parsed here by the VTK XML format itself: DataArrays in
"ascii", inline "binary" (base64 with a UInt32 or UInt64 byte-count header,
optionally vtkZLibDataCompressor blocks) and "appended" (raw or base64)
encodings, little- or big-endian.  A .pvtu's pieces are concatenated in file
order, points and cells alike (cell point ids shifted by the points before
the piece), as vtkXMLPUnstructuredGridReader appends them.  write_vtu /
write_pvtu produce the same encodings (tests and fixtures).
"""
from __future__ import annotations

import base64
import glob
import os
import re
import xml.etree.ElementTree as ET
import zlib
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

F32 = np.float32
VTK_TETRA = 10

_DTYPES = {"Int8": "i1", "UInt8": "u1", "Int16": "i2", "UInt16": "u2", "Int32": "i4", "UInt32": "u4",
           "Int64": "i8", "UInt64": "u8", "Float32": "f4", "Float64": "f8"}

# vtk_reader.py:558-604: names tried, in order, when the velocity field is not given
VELOCITY_NAMES = ['velocity', 'Velocity', 'VELOCITY', 'U', 'u', 'vel', 'Vel', 'VEL',
                  'Displacement', 'displacement', 'DISPLACEMENT', 'VelocityField', 'velocity_field',
                  'FlowVelocity', 'flow_velocity', 'v', 'V', 'fluid_velocity', 'Result', 'RESULT',
                  'NodeVel', 'NODEVEL']


# ---------------------------------------------------------------------------
# Reading one .vtu
# ---------------------------------------------------------------------------

class _Doc:
    """A parsed VTK XML file: root element, byte order, header type,
    compressor and the raw appended-data bytes."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            raw = fh.read()
        self.path = path
        app = None
        i = raw.find(b"<AppendedData")
        if i >= 0:
            j = raw.find(b">", i)
            head = raw[i:j + 1]
            k = raw.rfind(b"</AppendedData>")
            body = raw[j + 1:k]
            u = body.find(b"_")
            app = (head, body[u + 1:])
            raw = raw[:j + 1] + b"</AppendedData>" + raw[k + len(b"</AppendedData>"):]
        self.root = ET.fromstring(raw)
        if self.root.tag != "VTKFile":
            raise ValueError(f"{path}: not a VTK XML file")
        self.endian = "<" if self.root.get("byte_order", "LittleEndian") == "LittleEndian" else ">"
        self.header = _DTYPES[self.root.get("header_type", "UInt32")]
        self.compressed = self.root.get("compressor") == "vtkZLibDataCompressor"
        self.appended = None
        if app is not None:
            enc = self.root.find("AppendedData").get("encoding", "raw")
            data = app[1]
            if enc == "base64":
                data = data.strip()
            self.appended = (enc, data)

    def _hdr(self, buf: bytes, count: int) -> Tuple[np.ndarray, int]:
        dt = np.dtype(self.endian + self.header)
        return np.frombuffer(buf[:count * dt.itemsize], dtype=dt).astype(np.int64), count * dt.itemsize

    def _decode_block(self, data: bytes) -> Tuple[bytes, int]:
        """One encoded array (raw bytes): returns (payload, bytes consumed)."""
        if not self.compressed:
            (n,), off = self._hdr(data, 1)
            return data[off:off + n], off + n
        (nb, bsize, last), off = self._hdr(data, 3)
        sizes, off2 = self._hdr(data[off:], int(nb))
        pos = off + off2
        out = []
        for s in sizes:
            out.append(zlib.decompress(data[pos:pos + s]))
            pos += int(s)
        return b"".join(out), pos

    def _b64_array(self, text: bytes) -> bytes:
        """The payload of one base64-encoded array starting at text[0]:
        uncompressed, header and data are one base64 stream; compressed, the
        header and the concatenated blocks are two streams."""
        hb = np.dtype(self.header).itemsize
        first = base64.b64decode(text[:((hb + 2) // 3) * 4 + 4])[:hb]
        n = int(np.frombuffer(first, dtype=self.endian + self.header)[0])
        if not self.compressed:
            full = base64.b64decode(text[:((hb + n + 2) // 3) * 4])
            return full[hb:hb + n]
        hlen = ((3 + n) * hb + 2) // 3 * 4
        head = base64.b64decode(text[:hlen])[:(3 + n) * hb]
        sizes = np.frombuffer(head[3 * hb:], dtype=self.endian + self.header).astype(np.int64)
        ctot = int(sizes.sum())
        payload = base64.b64decode(text[hlen:hlen + ((ctot + 2) // 3) * 4])[:ctot]
        return self._decode_block(head + payload)[0]

    def array(self, el: ET.Element) -> np.ndarray:
        """A DataArray element as a (tuples, components) or flat array."""
        dt = np.dtype(self.endian + _DTYPES[el.get("type")])
        ncomp = int(el.get("NumberOfComponents", "1"))
        fmt = el.get("format", "ascii")
        if fmt == "ascii":
            vals = np.array((el.text or "").split(), dtype=np.float64 if dt.kind == "f" else np.int64)
            a = vals.astype(dt.newbyteorder("="))
        elif fmt == "binary":
            payload = self._b64_array("".join((el.text or "").split()).encode())
            a = np.frombuffer(payload, dtype=dt).astype(dt.newbyteorder("="))
        elif fmt == "appended":
            if self.appended is None:
                raise ValueError(f"{self.path}: appended DataArray without AppendedData")
            enc, data = self.appended
            off = int(el.get("offset", "0"))
            payload = self._b64_array(data[off:]) if enc == "base64" else self._decode_block(data[off:])[0]
            a = np.frombuffer(payload, dtype=dt).astype(dt.newbyteorder("="))
        else:
            raise ValueError(f"{self.path}: unsupported DataArray format {fmt!r}")
        return a.reshape(-1, ncomp) if ncomp > 1 else a


def read_vtu(path: str) -> Dict[str, object]:
    """One UnstructuredGrid file: {'points' (N, 3) f32, 'connectivity',
    'offsets', 'types' (the cells as stored), 'point_data' {name: array}};
    several Pieces are concatenated."""
    doc = _Doc(path)
    grid = doc.root.find("UnstructuredGrid")
    if grid is None:
        raise ValueError(f"{path}: no UnstructuredGrid element")
    pts, conn, offs, types, pdata = [], [], [], [], {}
    base_pt = 0
    base_conn = 0
    for piece in grid.findall("Piece"):
        npt = int(piece.get("NumberOfPoints", "0"))
        ncell = int(piece.get("NumberOfCells", "0"))
        pel = piece.find("Points/DataArray")
        p = doc.array(pel).reshape(npt, -1) if npt else np.zeros((0, 3))
        pts.append(p)
        cells = {el.get("Name"): el for el in piece.findall("Cells/DataArray")}
        if ncell:
            c = doc.array(cells["connectivity"]).astype(np.int64).ravel()
            o = doc.array(cells["offsets"]).astype(np.int64).ravel()
            t = doc.array(cells["types"]).astype(np.int64).ravel()
            conn.append(c + base_pt)
            offs.append(o + base_conn)
            types.append(t)
            base_conn += len(c)
        for el in piece.findall("PointData/DataArray"):
            pdata.setdefault(el.get("Name"), []).append(doc.array(el))
        base_pt += npt
    points = np.concatenate(pts) if pts else np.zeros((0, 3))
    return {"points": points, "connectivity": np.concatenate(conn) if conn else np.zeros(0, np.int64),
            "offsets": np.concatenate(offs) if offs else np.zeros(0, np.int64),
            "types": np.concatenate(types) if types else np.zeros(0, np.int64),
            "point_data": {k: np.concatenate(v) for k, v in pdata.items()}}


def read_pvtu(path: str) -> Dict[str, object]:
    """A PUnstructuredGrid file: its Pieces' .vtu files (paths relative to the
    .pvtu) concatenated in order, point ids of later pieces shifted."""
    root = ET.parse(path).getroot()
    grid = root.find("PUnstructuredGrid")
    if grid is None:
        raise ValueError(f"{path}: no PUnstructuredGrid element")
    parts = [read_vtu(os.path.join(os.path.dirname(path), p.get("Source"))) for p in grid.findall("Piece")]
    if not parts:
        raise ValueError(f"{path}: no pieces")
    base_pt, base_conn = 0, 0
    conn, offs = [], []
    for p in parts:
        conn.append(p["connectivity"] + base_pt)
        offs.append(p["offsets"] + base_conn)
        base_pt += len(p["points"])
        base_conn += len(p["connectivity"])
    names = [k for k in parts[0]["point_data"] if all(k in p["point_data"] for p in parts)]
    return {"points": np.concatenate([p["points"] for p in parts]), "connectivity": np.concatenate(conn),
            "offsets": np.concatenate(offs), "types": np.concatenate([p["types"] for p in parts]),
            "point_data": {k: np.concatenate([p["point_data"][k] for p in parts]) for k in names}}


def read_unstructured(path: str) -> Dict[str, object]:
    low = path.lower()
    if low.endswith(".pvtu"):
        return read_pvtu(path)
    if low.endswith(".vtu"):
        return read_vtu(path)
    raise ValueError(f"Unsupported VTK file format: {path}")


def tetra_connectivity(grid: Dict[str, object]) -> np.ndarray:
    """(M, 4) int32 point ids of the VTK_TETRA cells, in cell order
    (example_workflow.py:523-532)."""
    types = grid["types"]
    offs = grid["offsets"]
    conn = grid["connectivity"]
    ends = offs
    starts = np.concatenate([[0], offs[:-1]]) if len(offs) else offs
    sel = np.nonzero(types == VTK_TETRA)[0]
    if len(sel) == 0:
        return np.zeros((0, 4), np.int32)
    if np.any(ends[sel] - starts[sel] != 4):
        raise ValueError("a VTK_TETRA cell without 4 point ids")
    idx = starts[sel][:, None] + np.arange(4)[None, :]
    return conn[idx].astype(np.int32)


# vtk_reader.py:54-103
def _ensure_3d(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=F32)
    if a.ndim == 1:
        a = a[:, None]
    if a.ndim != 2:
        raise ValueError(f"expected a 2-D array (N, D), got shape {a.shape}")
    out = np.zeros((a.shape[0], 3), F32)
    k = min(3, a.shape[1])
    out[:, :k] = a[:, :k]
    return out


def _natural_sort_key(filename: str) -> tuple:
    """vtk_reader.py:106-117: the numbers in the file's base name."""
    base = os.path.basename(filename)
    nums = tuple(int(x) for x in re.findall(r"\d+", base))
    return nums if nums else (base,)


# ---------------------------------------------------------------------------
# The reference reader's API
# ---------------------------------------------------------------------------

class VTKUnstructuredTimeSeriesReader:
    """vtk_reader.py:426-942: a time series of .pvtu (preferred) or .vtu files
    on one fixed point set, velocities as (T, N, 3) float32."""

    def __init__(self, file_pattern: str, max_time_steps: Optional[int] = None,
                 velocity_field_name: Optional[str] = None, cache_size_limit: Optional[int] = None):
        self.file_pattern = file_pattern
        self.max_time_steps = max_time_steps
        self.velocity_field_name = velocity_field_name
        self.cache_size_limit = 15 if cache_size_limit is None else int(cache_size_limit)
        self.time_files = self._find_time_files()
        if not self.time_files:
            raise ValueError(f"No VTK files found matching pattern: {file_pattern}")
        self.grid_points = None
        self.connectivity = None
        self.velocity_cache: Dict[int, np.ndarray] = {}
        self._grid_bounds = None
        self._n_grid_points = None

    def _find_time_files(self) -> List[str]:
        """vtk_reader.py:496-541."""
        files = glob.glob(self.file_pattern) or glob.glob(os.path.abspath(self.file_pattern)) or \
            glob.glob(os.path.expanduser(self.file_pattern))
        pvtu = [f for f in files if f.lower().endswith(".pvtu")]
        vtu = [f for f in files if f.lower().endswith(".vtu")]
        sel = pvtu if pvtu else vtu
        sel.sort(key=_natural_sort_key)
        if self.max_time_steps is not None and len(sel) > self.max_time_steps:
            sel = sel[-self.max_time_steps:]
        return sel

    def _detect_velocity_field(self, point_data: Dict[str, np.ndarray]) -> str:
        """vtk_reader.py:558-603."""
        if self.velocity_field_name is not None:
            return self.velocity_field_name
        for name in VELOCITY_NAMES:
            if name in point_data:
                self.velocity_field_name = name
                return name
        avail = [f"{k} ({1 if v.ndim == 1 else v.shape[1]}D)" for k, v in point_data.items()]
        raise ValueError(f"Velocity field not found. Available point data fields: {avail}\n"
                         f"Specify velocity_field_name explicitly if using non-standard field names.")

    def load_single_timestep(self, time_idx: int) -> np.ndarray:
        """vtk_reader.py:605-681: (N, 3) float32 velocity of one file."""
        if time_idx < 0 or time_idx >= len(self.time_files):
            raise IndexError(f"Time index {time_idx} out of range [0, {len(self.time_files) - 1}]")
        if time_idx in self.velocity_cache:
            return self.velocity_cache[time_idx]
        grid = read_unstructured(self.time_files[time_idx])
        if self.grid_points is None:
            self.grid_points = _ensure_3d(grid["points"])
            self.connectivity = tetra_connectivity(grid)
            self._n_grid_points = self.grid_points.shape[0]
            mn, mx = self.grid_points.min(axis=0), self.grid_points.max(axis=0)
            self._grid_bounds = tuple(float(v) for v in (mn[0], mn[1], mn[2], mx[0], mx[1], mx[2]))
        name = self._detect_velocity_field(grid["point_data"])
        if name not in grid["point_data"]:
            raise ValueError(f"Velocity field '{name}' not found in {self.time_files[time_idx]}")
        vel = _ensure_3d(grid["point_data"][name])
        if vel.shape[0] != self._n_grid_points:
            raise ValueError(f"Velocity field size mismatch in {self.time_files[time_idx]}: "
                             f"expected {self._n_grid_points}, got {vel.shape[0]}")
        self.velocity_cache[time_idx] = vel
        while len(self.velocity_cache) > self.cache_size_limit:
            del self.velocity_cache[next(iter(self.velocity_cache))]
        return vel

    def load_time_series(self, start_idx: int = 0, end_idx: Optional[int] = None) -> Dict[str, object]:
        """vtk_reader.py:697-765: velocity_data (T, N, 3), positions (N, 3),
        times = the step indices as float32."""
        end_idx = len(self.time_files) if end_idx is None else min(end_idx, len(self.time_files))
        n_t = end_idx - start_idx
        if n_t <= 0:
            raise ValueError(f"Invalid time range: [{start_idx}, {end_idx})")
        if self.grid_points is None:
            self.load_single_timestep(start_idx)
        vel = np.zeros((n_t, self._n_grid_points, 3), F32)
        for i, k in enumerate(range(start_idx, end_idx)):
            vel[i] = self.load_single_timestep(k)
        return {"velocity_data": vel, "positions": self.grid_points,
                "times": np.arange(start_idx, end_idx, dtype=F32), "source": "vtk_unstructured_series",
                "file_pattern": self.file_pattern, "velocity_field_name": self.velocity_field_name}

    def get_bounds(self) -> Tuple[float, ...]:
        if self._grid_bounds is None:
            self.load_single_timestep(0)
        return self._grid_bounds

    def __len__(self) -> int:
        return len(self.time_files)

    def __getitem__(self, i: int) -> np.ndarray:
        return self.load_single_timestep(i)


def read_octree_series(data_pattern: str, max_timesteps: int = 40, skip_initial_timesteps: int = 0,
                       use_stable_mesh_only: bool = True) -> Dict[str, object]:
    """example_workflow.py:436-610, the file side of the octree loader: the
    sorted files of the pattern, minus the first skip_initial_timesteps and,
    with more than 10 left, minus those before the first of the first 10 whose
    point count is the most common among them; then max_timesteps of them at
    stride len // max_timesteps.  Points and VTK_TETRA connectivity come from
    the first file; per file the first of 'Displacement', 'displacement',
    'Velocity', 'velocity' is the velocity (2 components padded with a zero z)
    and the time is the n of a "_<n>.pvtu" suffix, else the file's index.
    Returns velocity_data (T, N, 3) f32, times (T,) f32, positions (N, 3) f32,
    connectivity (M, 4) int32 and the files read."""
    files = sorted(glob.glob(data_pattern))
    if not files:
        raise FileNotFoundError(f"No files found: {data_pattern}")
    if skip_initial_timesteps > 0:
        files = files[skip_initial_timesteps:]
    if use_stable_mesh_only and len(files) > 10:
        sizes = [len(read_unstructured(f)["points"]) for f in files[:10]]
        common = max(set(sizes), key=lambda s: (sizes.count(s), -sizes.index(s)))
        start = sizes.index(common)
        if start > 0:
            files = files[start:]
    stride = max(1, len(files) // max_timesteps)
    files = files[::stride][:max_timesteps]
    first = read_unstructured(files[0])
    points = np.asarray(first["points"], F32)
    conn = tetra_connectivity(first)
    vel, times = [], []
    for idx, fname in enumerate(files):
        g = first if idx == 0 else read_unstructured(fname)
        arr = None
        for name in ("Displacement", "displacement", "Velocity", "velocity"):
            if name in g["point_data"]:
                arr = g["point_data"][name]
                break
        if arr is None:
            raise ValueError(f"No velocity field found in {fname}")
        v = np.asarray(arr, F32)
        if v.ndim == 2 and v.shape[1] == 2:
            v = np.column_stack([v, np.zeros(v.shape[0])]).astype(F32)
        vel.append(v)
        m = re.search(r"_(\d+)\.pvtu$", fname)
        times.append(float(m.group(1)) if m else float(idx))
    sizes = {v.shape[0] for v in vel}
    if len(sizes) > 1:
        raise ValueError(f"Inconsistent mesh sizes: {len(sizes)} different sizes found {sorted(sizes)}. "
                         f"JAXTrace octree FEM requires fixed mesh topology. "
                         f"Use --skip-timesteps or prepare data with constant mesh.")
    return {"velocity_data": np.array(vel, dtype=F32), "times": np.array(times, dtype=F32),
            "positions": points, "connectivity": conn, "files": files}


def load_octree_field(data_pattern: str, max_timesteps: int = 40, skip_initial_timesteps: int = 0,
                      use_stable_mesh_only: bool = True, max_elements_per_leaf: int = 32,
                      max_octree_depth: int = 12, **field_kwargs):
    """example_workflow.py:436-660: read_octree_series into
    OctreeFEMTimeSeriesFieldOptimized (interpolation 'linear', extrapolation
    'constant'; the element octree is built on the GPU)."""
    from .fields import OctreeFEMTimeSeriesFieldOptimized
    s = read_octree_series(data_pattern, max_timesteps, skip_initial_timesteps, use_stable_mesh_only)
    field_kwargs.setdefault("interpolation", "linear")
    field_kwargs.setdefault("extrapolation", "constant")
    return OctreeFEMTimeSeriesFieldOptimized(s["velocity_data"], s["times"], s["positions"], s["connectivity"],
                                             max_elements_per_leaf=max_elements_per_leaf,
                                             max_depth=max_octree_depth, **field_kwargs)


# ---------------------------------------------------------------------------
# Writing (the same encodings; fixtures and tests)
# ---------------------------------------------------------------------------

_NAMES = {v: k for k, v in _DTYPES.items()}


def _encode(a: np.ndarray, header: str, compress: bool, endian: str) -> bytes:
    """Header (byte count, or the zlib block table) and payload of one array."""
    raw = np.ascontiguousarray(a).astype(a.dtype.newbyteorder(endian)).tobytes()
    hdt = np.dtype(endian + header)
    if not compress:
        return np.array([len(raw)], hdt).tobytes() + raw
    bsize = 1 << 15
    blocks = [raw[i:i + bsize] for i in range(0, len(raw), bsize)] or [b""]
    comp = [zlib.compress(b) for b in blocks]
    last = len(blocks[-1]) if len(raw) % bsize else (bsize if raw else 0)
    head = np.array([len(blocks), bsize, last] + [len(c) for c in comp], hdt).tobytes()
    return head + b"".join(comp)


def write_vtu(path: str, points: np.ndarray, tets: np.ndarray, point_data: Optional[Dict[str, np.ndarray]] = None,
              encoding: str = "binary", header_type: str = "UInt32", compress: bool = False,
              big_endian: bool = False) -> None:
    """One Piece of VTK_TETRA cells, arrays as 'ascii', inline 'binary'
    (base64), 'appended' (raw) or 'appended-base64'."""
    point_data = point_data or {}
    endian = ">" if big_endian else "<"
    pts = np.asarray(points, F32)
    tets = np.asarray(tets, np.int64)
    enc = encoding
    app_parts: List[bytes] = []
    app_len = [0]
    hdt = _DTYPES[header_type]

    def da(name, a, ncomp):
        a = np.asarray(a)
        attrs = f'type="{_NAMES[a.dtype.str[1:]]}" Name="{name}" NumberOfComponents="{ncomp}"'
        if enc == "ascii":
            body = " ".join(repr(float(v)) if a.dtype.kind == "f" else str(int(v)) for v in a.ravel())
            return f'<DataArray {attrs} format="ascii">{body}</DataArray>'
        blob = _encode(a, hdt, compress, endian)
        if enc in ("binary", "appended-base64"):
            if compress:
                nb = int(np.frombuffer(blob[:np.dtype(hdt).itemsize], endian + hdt)[0])
                hl = (3 + nb) * np.dtype(hdt).itemsize
                blob = base64.b64encode(blob[:hl]) + base64.b64encode(blob[hl:])
            else:
                blob = base64.b64encode(blob)
            if enc == "binary":
                return f'<DataArray {attrs} format="binary">{blob.decode()}</DataArray>'
        off = app_len[0]
        app_parts.append(blob)
        app_len[0] += len(blob)
        return f'<DataArray {attrs} format="appended" offset="{off}"/>'

    xml = [f'<?xml version="1.0"?>\n<VTKFile type="UnstructuredGrid" version="1.0" '
           f'byte_order="{"BigEndian" if big_endian else "LittleEndian"}" header_type="{header_type}"'
           + (' compressor="vtkZLibDataCompressor"' if compress else "") + ">\n<UnstructuredGrid>\n",
           f'<Piece NumberOfPoints="{len(pts)}" NumberOfCells="{len(tets)}">\n<Points>'
           + da("Points", pts, 3) + "</Points>\n<Cells>"
           + da("connectivity", tets.ravel(), 1)
           + da("offsets", (4 * np.arange(1, len(tets) + 1)).astype(np.int64), 1)
           + da("types", np.full(len(tets), VTK_TETRA, np.uint8), 1) + "</Cells>\n<PointData>"]
    for name, v in point_data.items():
        v = np.asarray(v)
        xml.append(da(name, v, 1 if v.ndim == 1 else v.shape[1]))
    xml.append("</PointData>\n</Piece>\n</UnstructuredGrid>\n")
    out = "".join(xml).encode()
    if app_parts:
        mode = "base64" if enc == "appended-base64" else "raw"
        out += f'<AppendedData encoding="{mode}">_'.encode() + b"".join(app_parts) + b"</AppendedData>\n"
    out += b"</VTKFile>\n"
    with open(path, "wb") as fh:
        fh.write(out)


def write_pvtu(path: str, sources: Sequence[str], point_data_names: Sequence[str] = ()) -> None:
    """A PUnstructuredGrid naming its piece files (relative paths)."""
    pd = "".join(f'<PDataArray type="Float32" Name="{n}" NumberOfComponents="3"/>' for n in point_data_names)
    pieces = "".join(f'<Piece Source="{s}"/>' for s in sources)
    with open(path, "w") as fh:
        fh.write('<?xml version="1.0"?>\n<VTKFile type="PUnstructuredGrid" version="1.0" '
                 'byte_order="LittleEndian" header_type="UInt32">\n<PUnstructuredGrid GhostLevel="0">\n'
                 f'<PPoints><PDataArray type="Float32" NumberOfComponents="3"/></PPoints>\n'
                 f'<PPointData>{pd}</PPointData>\n{pieces}\n</PUnstructuredGrid>\n</VTKFile>\n')
