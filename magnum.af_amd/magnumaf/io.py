"""State files: the VTK image-data (.vti) form of a magnetization and a step-exact resume file.

.vti (State::write_vti / _vti_reader, state.cpp:139-141, vtk_io.cpp:52-100): the reference stores
m as cell data of an image grid with nx+1, ny+1, nz+1 points, spacing (dx, dy, dz), origin 0,
one Float64 array of 3 components per cell, x fastest.  write_vti produces that grid with the
array, named ImageScalars as VTK names the reference's scalars, in raw appended form (UInt64
byte-count header); read_vti reads such files and the ascii
and uncompressed inline-binary encodings of other writers.

Resume files (.npz) hold what a continuation needs to take the same steps as a run that was never
interrupted: m bit for bit, t, the step count, and the integrator's scheme, parameters and
controller state (next step size, previous error, reject flag).  The reference has no such file:
its State can be rebuilt from a .vti, but the integrator then restarts from h = 1.01e-15.
"""
import base64
import re
import struct
import xml.etree.ElementTree as ET

import numpy as np


def _vti_name(path):
    # vtk_io.cpp:86-87: ".vti" is appended unless present
    return path if str(path).endswith(".vti") else str(path) + ".vti"


def write_vti(path, m, spacing, name="ImageScalars"):
    """m: array (nx, ny, nz, 3); spacing (dx, dy, dz) in m.  Returns the file name written."""
    m = np.asarray(m, dtype=np.float64)
    if m.ndim != 4 or m.shape[3] != 3:
        raise ValueError("write_vti: m must have shape (nx, ny, nz, 3)")
    nx, ny, nz = m.shape[:3]
    dx, dy, dz = (float(v) for v in spacing)
    data = np.ascontiguousarray(m.transpose(2, 1, 0, 3)).astype("<f8").tobytes()
    ext = f"0 {nx} 0 {ny} 0 {nz}"
    head = (
        '<?xml version="1.0"?>\n'
        '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n'
        f'  <ImageData WholeExtent="{ext}" Origin="0 0 0" Spacing="{dx!r} {dy!r} {dz!r}">\n'
        f'    <Piece Extent="{ext}">\n'
        "      <PointData>\n      </PointData>\n"
        f'      <CellData Scalars="{name}">\n'
        f'        <DataArray type="Float64" Name="{name}" NumberOfComponents="3" format="appended" offset="0"/>\n'
        "      </CellData>\n    </Piece>\n  </ImageData>\n"
        '  <AppendedData encoding="raw">\n   _'
    )
    fn = _vti_name(path)
    with open(fn, "wb") as f:
        f.write(head.encode())
        f.write(struct.pack("<Q", len(data)))
        f.write(data)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return fn


def read_vti(path):
    """Returns (m (nx, ny, nz, 3) float64, spacing (dx, dy, dz)) from the first 3-component cell
    array of an image-data file."""
    raw = open(_vti_name(path), "rb").read()
    appended, app_enc = b"", None
    k = raw.find(b"<AppendedData")
    if k >= 0:
        tag_end = raw.index(b">", k)
        mt = re.search(rb'encoding="([a-z0-9]+)"', raw[k:tag_end])
        app_enc = mt.group(1).decode() if mt else None
        us = raw.index(b"_", tag_end) + 1
        appended = raw[us:]
        raw = raw[:k] + b"</VTKFile>"
    root = ET.fromstring(raw.decode())
    if root.get("compressor"):
        raise ValueError("read_vti: compressed .vti files are not supported")
    big = root.get("byte_order", "LittleEndian") == "BigEndian"
    hdr = "Q" if root.get("header_type", "UInt32") == "UInt64" else "I"
    img = root.find("ImageData")
    e = [int(v) for v in img.get("WholeExtent").split()]
    nx, ny, nz = e[1] - e[0], e[3] - e[2], e[5] - e[4]
    spacing = tuple(float(v) for v in img.get("Spacing").split())
    arr = None
    for da in img.iter("DataArray"):
        if int(da.get("NumberOfComponents", "1")) == 3:
            arr = da
            break
    if arr is None:
        raise ValueError("read_vti: no 3-component cell array")
    dt = {"Float64": "f8", "Float32": "f4"}[arr.get("type")]
    dt = (">" if big else "<") + dt
    bo = ">" if big else "<"
    fmt = arr.get("format")
    n = nx * ny * nz * 3
    if fmt == "ascii":
        v = np.array(arr.text.split(), dtype=np.float64)
    elif fmt == "binary":
        blob = base64.b64decode(re.sub(r"\s+", "", arr.text))
        hs = struct.calcsize(hdr)
        nbytes = struct.unpack(bo + hdr, blob[:hs])[0]
        v = np.frombuffer(blob[hs:hs + nbytes], dtype=dt)
    elif fmt == "appended":
        off = int(arr.get("offset", "0"))
        hs = struct.calcsize(hdr)
        if app_enc != "raw":
            raise ValueError(f"read_vti: appended data encoding {app_enc!r} is not supported")
        nbytes = struct.unpack(bo + hdr, appended[off:off + hs])[0]
        v = np.frombuffer(appended[off + hs:off + hs + nbytes], dtype=dt)
    else:
        raise ValueError(f"read_vti: unknown data format {fmt!r}")
    if v.size != n:
        raise ValueError(f"read_vti: expected {n} values, found {v.size}")
    m = v.astype(np.float64).reshape(nz, ny, nx, 3).transpose(2, 1, 0, 3)
    return np.ascontiguousarray(m), spacing


def stream_line(t, mean):
    """One line of the reference's state stream, 't<TAB>mx<TAB>my<TAB>mz' with the default
    six significant digits of std::ostream (state.cpp:26-30, af_overloads.hpp:57-80)."""
    return f"{t:g}\t{mean[0]:g}\t{mean[1]:g}\t{mean[2]:g}"
