"""VTK output (visualization/write2vtk.jl): the writer's files read back
with an independent XML + numpy reader; no GPU needed."""
import os
import xml.etree.ElementTree as ET

import numpy as np

from mfrbf import vtk


def _tree_before_data(path):
    raw = open(path, "rb").read()
    return ET.fromstring(raw[:raw.index(b"<AppendedData")].decode("ascii") + "</VTKFile>")


def test_vtu_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    N = 257
    pts = rng.random((N, 2))
    rho = rng.random(N)
    mom = rng.standard_normal((N, 2))
    code = rng.integers(0, 3, N).astype(np.int8)
    idx = np.arange(1, N + 1, dtype=np.int64)
    path = vtk.write_vtu(str(tmp_path / "a.vtu"), pts, [("density", rho), ("momentum", mom), ("eps_scalar", code),
                                                         ("index", idx)],
                         [("time", 0.125), ("solver_version", "v1-test")])
    p, pd, fd = vtk.read_vtu(path)
    assert np.array_equal(p[:, :2], pts) and np.all(p[:, 2] == 0.0)
    assert np.array_equal(pd["density"], rho)
    assert np.array_equal(pd["momentum"], mom)
    assert np.array_equal(pd["eps_scalar"], code) and pd["eps_scalar"].dtype == np.int8
    assert np.array_equal(pd["index"], idx)
    assert float(fd["time"][0]) == 0.125 and fd["solver_version"] == "v1-test"
    # one vertex cell per point (VTK_VERTEX = 1), offsets 1..N
    root = _tree_before_data(path)
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == str(N) and piece.get("NumberOfCells") == str(N)
    names = [e.get("Name") for e in piece.find("Cells")]
    assert names == ["connectivity", "offsets", "types"]
    assert root.get("header_type") == "UInt64" and root.get("byte_order") == "LittleEndian"


def test_vtu_appended_block_layout(tmp_path):
    """Each appended block is a UInt64 byte count and the raw little-endian
    values, at the offset its DataArray names (counted from after '_')."""
    vals = np.array([1.5, -2.0, 3.25])
    path = vtk.write_vtu(str(tmp_path / "b.vtu"), np.zeros((3, 3)), [("v", vals)])
    raw = open(path, "rb").read()
    base = raw.index(b"_", raw.index(b"<AppendedData")) + 1
    el = [e for e in _tree_before_data(path).iter("DataArray") if e.get("Name") == "v"][0]
    o = base + int(el.get("offset"))
    assert int.from_bytes(raw[o:o + 8], "little") == 24
    assert np.frombuffer(raw[o + 8:o + 32], "<f8").tolist() == vals.tolist()


def test_coordinates_writer_and_names(tmp_path):
    """trixi2vtk(coordinates; prefix, filename, custom quantities...)
    (write2vtk.jl:281-301) and the "<type>_<k>" system names
    (write2vtk.jl:45-52)."""
    pts = np.array([[0.0, 1.0, 2.0], [0.0, 0.5, 1.0]])  # 2 x N, as the reference passes them
    path = vtk.trixi2vtk(pts, output_directory=str(tmp_path), prefix="p", filename="coords",
                         mass=np.array([1.0, 2.0, 3.0]))
    assert os.path.basename(path) == "p_coords.vtu"
    p, pd, _ = vtk.read_vtu(path)
    assert np.array_equal(p[:, 0], [0.0, 1.0, 2.0]) and np.array_equal(pd["index"], [1, 2, 3])
    assert np.array_equal(pd["mass"], [1.0, 2.0, 3.0])

    class A:
        pass

    class B:
        pass

    assert vtk.system_names([A(), B(), A()]) == ["A_1", "B_1", "A_2"]
