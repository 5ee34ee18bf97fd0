"""VTK output: trixi2vtk (visualization/write2vtk.jl:113-279, 303-351).

One VTK XML UnstructuredGrid file (.vtu) per call, one vertex cell per node,
appended raw binary data with UInt64 block headers, plus a ParaView
collection (.pvd) that lists the file at time t, as WriteVTK.jl's
vtk_grid / paraview_collection do for the reference.

Point data follow the reference's write2vtk! methods (write2vtk.jl:303-351):
- CompressibleEulerEquations2D: density, density_energy, momentum (2
  components), pressure, velocity (2 components);
- SourceUpwindViscosityTominec / SourceResidualViscosityTominec: eps and
  eps_scalar (the code of the branch that set eps), and for the residual
  source approx_du.  The residual source's eps_uw, eps_rv and residual are
  intermediates of the device kernel that are not kept, so they are not
  written;
- SourceIGR: sigma;
- always index (1-based, as the reference's eachelement) and the field data
  time and solver_version.
The state is read back from the device; the files are written on the host.
"""
import os
import subprocess

import numpy as np

from .equations import CompressibleEulerEquations2D
from .sources import SourceIGR, SourceResidualViscosityTominec, SourceUpwindViscosityTominec

_TYPES = {np.dtype(np.float64): "Float64", np.dtype(np.float32): "Float32",
          np.dtype(np.int64): "Int64", np.dtype(np.int32): "Int32",
          np.dtype(np.int8): "Int8", np.dtype(np.uint8): "UInt8"}
VTK_VERTEX = 1


def vtkname(system):
    """vtkname (write2vtk.jl:42): the system's type name."""
    return type(system).__name__


def system_names(systems):
    """system_names (write2vtk.jl:45-52): "<type>_<k>" for the k-th system of
    that type."""
    names = [vtkname(s) for s in systems]
    return ["%s_%d" % (n, names[:i + 1].count(n)) for i, n in enumerate(names)]


def assemble_systems(semi):
    """assemble_systems (write2vtk.jl:3-12): the equations, then the sources."""
    return (semi.equations,) + tuple(semi.source_terms.values())


def get_git_hash():
    """get_git_hash (write2vtk.jl:54-70), for this package's directory."""
    d = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    try:
        r = subprocess.run(["git", "describe", "--tags", "--always", "--first-parent", "--dirty"], cwd=d,
                           stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, timeout=10)
        return r.stdout.strip() if r.returncode == 0 and r.stdout.strip() else "UnknownVersion"
    except (OSError, subprocess.SubprocessError):
        return "UnknownVersion"


def _ascii_string(s):
    # VTK XML String arrays in ascii format: byte values, each string ends with 0
    return " ".join(str(b) for b in s.encode("utf-8")) + " 0"


def write_vtu(path, points, point_data=(), field_data=()):
    """Writes an UnstructuredGrid of vertex cells.  points: (N, 2) or (N, 3);
    point_data: (name, array) pairs, arrays (N,) or (N, ncomp); field_data:
    (name, value) pairs, value a number, an array or a str.  Returns path."""
    pts = np.asarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] not in (2, 3):
        raise ValueError("points must be (N, 2) or (N, 3)")
    N = pts.shape[0]
    if pts.shape[1] == 2:
        pts = np.c_[pts, np.zeros(N)]
    blocks = []
    offset = 0

    def add(arr):
        nonlocal offset
        arr = np.ascontiguousarray(arr)
        if arr.dtype not in _TYPES:
            raise TypeError("unsupported array type %s" % arr.dtype)
        raw = arr.astype(arr.dtype.newbyteorder("<"), copy=False).tobytes()
        blocks.append(np.uint64(len(raw)).tobytes() + raw)
        o = offset
        offset += 8 + len(raw)
        return o

    def head(kind, name, arr, ncomp=None, tuples=None):
        a = np.asarray(arr)
        o = add(a)
        attrs = 'type="%s"' % _TYPES[a.dtype]
        if name:
            attrs += ' Name="%s"' % name
        if ncomp:
            attrs += ' NumberOfComponents="%d"' % ncomp
        if tuples:
            attrs += ' NumberOfTuples="%d"' % tuples
        return '%s<DataArray %s format="appended" offset="%d"/>' % (kind, attrs, o)

    fd = []
    for name, val in field_data:
        if isinstance(val, str):
            fd.append('      <DataArray type="String" Name="%s" NumberOfTuples="1" format="ascii">%s</DataArray>'
                      % (name, _ascii_string(val)))
        else:
            v = np.atleast_1d(np.asarray(val))
            fd.append(head("      ", name, v, tuples=v.size))
    p_head = head("        ", None, pts, ncomp=3)
    conn = head("        ", "connectivity", np.arange(N, dtype=np.int64))
    offs = head("        ", "offsets", np.arange(1, N + 1, dtype=np.int64))
    types = head("        ", "types", np.full(N, VTK_VERTEX, dtype=np.uint8))
    pd = []
    for name, arr in point_data:
        a = np.asarray(arr)
        if a.shape[0] != N:
            raise ValueError("point data %s has %d values for %d points" % (name, a.shape[0], N))
        pd.append(head("        ", name, a, ncomp=(a.shape[1] if a.ndim == 2 else None)))
    xml = ['<?xml version="1.0"?>',
           '<VTKFile type="UnstructuredGrid" version="1.0" byte_order="LittleEndian" header_type="UInt64">',
           '  <UnstructuredGrid>']
    if fd:
        xml += ["    <FieldData>"] + fd + ["    </FieldData>"]
    xml += ['    <Piece NumberOfPoints="%d" NumberOfCells="%d">' % (N, N),
            "      <Points>", p_head, "      </Points>",
            "      <Cells>", conn, offs, types, "      </Cells>"]
    if pd:
        xml += ["      <PointData>"] + pd + ["      </PointData>"]
    xml += ["    </Piece>", "  </UnstructuredGrid>", '  <AppendedData encoding="raw">']
    with open(path, "wb") as f:
        f.write(("\n".join(xml) + "\n   _").encode("ascii"))
        for b in blocks:
            f.write(b)
        f.write(b"\n  </AppendedData>\n</VTKFile>\n")
    return path


def _write_pvd(path, entries):
    lines = ['<?xml version="1.0"?>',
             '<VTKFile type="Collection" version="1.0" byte_order="LittleEndian">', "  <Collection>"]
    for t, f in entries:
        lines.append('    <DataSet timestep="%s" group="" part="0" file="%s"/>' % (repr(float(t)), f))
    lines += ["  </Collection>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


def _read_pvd(path):
    import xml.etree.ElementTree as ET
    if not os.path.exists(path):
        return []
    return [(float(d.get("timestep")), d.get("file")) for d in ET.parse(path).getroot().iter("DataSet")]


def _system_data(system, u, semi):
    """write2vtk! for one system (write2vtk.jl:303-351)."""
    out = []
    if isinstance(system, CompressibleEulerEquations2D):
        rho, m1, m2, E = u
        v1, v2 = m1 / rho, m2 / rho
        p = (system.gamma - 1.0) * (E - 0.5 * (m1 * v1 + m2 * v2))
        out += [("density", rho), ("density_energy", E), ("momentum", np.c_[m1, m2]),
                ("pressure", p), ("velocity", np.c_[v1, v2])]
    elif isinstance(system, (SourceUpwindViscosityTominec, SourceResidualViscosityTominec)):
        eps, code = semi.ctx.viscosity()
        out += [("eps", eps), ("eps_scalar", code)]
        if isinstance(system, SourceResidualViscosityTominec):
            out.append(("approx_du", semi.ctx.approx_du().T))
    elif isinstance(system, SourceIGR):
        out.append(("sigma", semi.ctx.igr_sigma()))
    return out


def trixi2vtk(*args, iter=None, output_directory="out", prefix="", write_meta_data=True,
              max_coordinates=np.inf, filename="coordinates", **custom_quantities):
    """trixi2vtk(u_ode, semi, t; iter, output_directory, prefix,
    write_meta_data, max_coordinates) writes the solution
    (write2vtk.jl:113-279); trixi2vtk(coordinates; output_directory, prefix,
    filename, custom_quantities...) writes a point set (write2vtk.jl:281-301).
    Returns the .vtu path.  u_ode is the (nvar, N) state; the viscosity
    fields come from the semidiscretization's last rhs!."""
    os.makedirs(output_directory, exist_ok=True)
    if len(args) == 1:
        pts = np.asarray(args[0], dtype=np.float64).T  # (d, N) as the reference's coordinates
        name = filename if prefix == "" else "%s_%s" % (prefix, filename)
        data = [("index", np.arange(1, pts.shape[0] + 1, dtype=np.int64))]
        data += [(k, np.asarray(v)) for k, v in custom_quantities.items() if v is not None]
        return write_vtu(os.path.join(output_directory, name + ".vtu"), pts, data)
    u_ode, semi, t = args
    systems = assemble_systems(semi)
    system_name = system_names(systems)[0]
    pre = "" if prefix == "" else prefix + "_"
    post = "" if iter is None else "_%s" % iter
    base = pre + system_name + post
    u = np.asarray(u_ode, dtype=np.float64).reshape(semi.nvariables, -1)
    pts = np.array(semi.domain.pd.points, dtype=np.float64)
    if np.abs(pts.max()) > max_coordinates or np.abs(pts.min()) > max_coordinates:
        print("Warning: At least one particle's absolute coordinates exceed `max_coordinates` "
              "and have been clipped")
        pts = np.clip(pts, -max_coordinates, max_coordinates)
    data = []
    for s in systems:
        data += _system_data(s, u, semi)
    data.append(("index", np.arange(1, u.shape[1] + 1, dtype=np.int64)))
    fields = [("time", float(t))]
    if write_meta_data:
        fields.append(("solver_version", get_git_hash()))
    path = write_vtu(os.path.join(output_directory, base + ".vtu"), pts, data, fields)
    # the collection restarts unless iter > 0 (paraview_collection(...; append = iter > 0))
    pvd = os.path.join(output_directory, pre + system_name + ".pvd")
    entries = _read_pvd(pvd) if (iter is not None and iter > 0) else []
    entries.append((t, os.path.basename(path)))
    _write_pvd(pvd, entries)
    return path


def read_vtu(path):
    """Reads a file written by write_vtu: (points (N, 3), point data dict,
    field data dict).  For tests and post-processing."""
    import xml.etree.ElementTree as ET
    raw = open(path, "rb").read()
    start = raw.index(b'<AppendedData encoding="raw">')
    us = raw.index(b"_", start) + 1
    root = ET.fromstring(raw[:start].decode("ascii") + "</VTKFile>")
    inv = {v: k for k, v in _TYPES.items()}

    def arr(el):
        o = us + int(el.get("offset"))
        n = int(np.frombuffer(raw, np.uint64, 1, o)[0])
        a = np.frombuffer(raw, inv[el.get("type")].newbyteorder("<"), n // inv[el.get("type")].itemsize, o + 8)
        nc = int(el.get("NumberOfComponents", "1"))
        return a.reshape(-1, nc) if nc > 1 else a

    piece = root.find("UnstructuredGrid/Piece")
    pts = arr(piece.find("Points/DataArray"))
    pd = piece.find("PointData")
    pdata = {el.get("Name"): arr(el) for el in (pd if pd is not None else [])}
    fdata = {}
    fd = root.find("UnstructuredGrid/FieldData")
    for el in (fd if fd is not None else []):
        if el.get("type") == "String":
            b = [int(x) for x in el.text.split()]
            fdata[el.get("Name")] = bytes(b[:b.index(0)]).decode("utf-8")
        else:
            fdata[el.get("Name")] = arr(el)
    return pts, pdata, fdata
