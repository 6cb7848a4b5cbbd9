"""State files on the CPU: the .vti form of a magnetization (vtk_io.cpp:52-100 layout) and the
reference's 't<TAB>mx<TAB>my<TAB>mz' stream line (state.cpp:26-30)."""
import base64
import struct
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from magnumaf import io


def test_vti_round_trip_and_layout(tmp_path):
    rng = np.random.default_rng(3)
    m = rng.standard_normal((5, 4, 3, 3))
    fn = io.write_vti(tmp_path / "state", m, (1e-9, 2e-9, 3e-9))
    assert fn.endswith(".vti")  # vtk_io.cpp:86-87
    back, spacing = io.read_vti(fn)
    assert np.array_equal(back, m) and spacing == (1e-9, 2e-9, 3e-9)
    # cell data of an (nx+1, ny+1, nz+1)-point grid, one Float64 3-vector per cell, x fastest
    raw = open(fn, "rb").read()
    head = raw[: raw.index(b"<AppendedData")].decode() + "</VTKFile>"
    root = ET.fromstring(head)
    img = root.find("ImageData")
    assert img.get("WholeExtent") == "0 5 0 4 0 3"
    da = img.find("Piece/CellData/DataArray")
    assert da.get("type") == "Float64" and da.get("NumberOfComponents") == "3"
    blob = raw[raw.index(b"_", raw.index(b"<AppendedData")) + 1:]
    nbytes = struct.unpack("<Q", blob[:8])[0]
    vals = np.frombuffer(blob[8:8 + nbytes], dtype="<f8")
    assert nbytes == m.size * 8
    x, y, z = 2, 1, 2
    assert np.array_equal(vals[3 * (x + 5 * (y + 4 * z)):][:3], m[x, y, z])


@pytest.mark.parametrize("fmt", ["ascii", "binary"])
def test_vti_reads_other_encodings(tmp_path, fmt):
    m = np.arange(2 * 3 * 2 * 3, dtype=np.float64).reshape(2, 3, 2, 3) / 7.0
    flat = m.transpose(2, 1, 0, 3).ravel()
    if fmt == "ascii":
        body = " ".join(repr(float(v)) for v in flat)
    else:
        data = flat.astype("<f8").tobytes()
        body = base64.b64encode(struct.pack("<I", len(data)) + data).decode()
    xml = ('<?xml version="1.0"?><VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">'
           '<ImageData WholeExtent="0 2 0 3 0 2" Origin="0 0 0" Spacing="1 1 1"><Piece Extent="0 2 0 3 0 2">'
           '<CellData Scalars="ImageScalars"><DataArray type="Float64" Name="ImageScalars" '
           f'NumberOfComponents="3" format="{fmt}">{body}</DataArray></CellData></Piece></ImageData></VTKFile>')
    (tmp_path / "o.vti").write_text(xml)
    back, _ = io.read_vti(tmp_path / "o.vti")
    assert np.array_equal(back, m)


def test_vti_rejects_compressed(tmp_path):
    xml = ('<?xml version="1.0"?><VTKFile type="ImageData" compressor="vtkZLibDataCompressor">'
           '<ImageData WholeExtent="0 1 0 1 0 1" Spacing="1 1 1"></ImageData></VTKFile>')
    (tmp_path / "c.vti").write_text(xml)
    with pytest.raises(ValueError):
        io.read_vti(tmp_path / "c.vti")


def test_stream_line_matches_ostream_default():
    # std::ostream prints doubles with six significant digits (%g)
    assert io.stream_line(1e-11, (0.123456789, -1.0, 2.5e-7)) == "1e-11\t0.123457\t-1\t2.5e-07"
