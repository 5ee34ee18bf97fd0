"""Host-side logic and the C-ABI library (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from mfrbf import _lib as L


def _header_symbols():
    src = open(os.path.join(ROOT, "include", "mfrbf.h")).read()
    return sorted(set(re.findall(r"\b(mf_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_header_symbol():
    import ctypes

    lib = ctypes.CDLL(L.LIB_PATH)
    syms = _header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(syms) == L.exported_symbols()


def test_header_constants_match_binding():
    src = open(os.path.join(ROOT, "include", "mfrbf.h")).read()
    for name, val in re.findall(r"#define (MF_[A-Z0-9_]+) (\d+)", src):
        assert getattr(L, name) == int(val), name


def test_context_fails_loudly_without_device():
    from conftest import gpu_available

    if gpu_available():
        pytest.skip("a GPU is present")
    with pytest.raises(L.MfError):
        L.Context(0)


def test_basis_stencil_size_rule():
    from mfrbf import PointCloudBasis, Point2D, PolyharmonicSpline, RBF, setproperties

    assert PointCloudBasis(Point2D(), 2).nv == 15
    assert PointCloudBasis(Point2D(), 3, approximation_type=RBF(PolyharmonicSpline(5))).nv == 20
    assert PointCloudBasis(Point2D(), 4).nv == 30
    b = setproperties(PointCloudBasis(Point2D(), 3, approximation_type=RBF(PolyharmonicSpline(5))), nv=30)
    assert b.nv == 30 and b.phs_power == 5 and b.npoly == 10


def test_medusa_reader_and_empty_group_quirk(tmp_path):
    from mfrbf import read_medusa_file

    pos, interior, bidx, bnorm, normals, ranges, types = read_medusa_file(
        os.path.join(GOLDEN, "cyl_0_05", "cyl_0_05"))
    assert [len(g) for g in bidx] == [40, 40, 60, 60, 26]
    assert all(np.allclose(np.linalg.norm(n, axis=1), 1.0) for n in bnorm)
    # group -2 empty: dropped from the index lists, kept (empty) in the normals
    base = str(tmp_path / "c")
    np.savetxt(base + "_positions.txt", [[0, 0], [1, 0], [0, 1], [1, 1]], delimiter=" , ")
    np.savetxt(base + "_types.txt", [-1, -3, -3, 0], fmt="%d")
    np.savetxt(base + "_boundary.txt", [0, 1, 2], fmt="%d")
    np.savetxt(base + "_interior.txt", [3], fmt="%d")
    np.savetxt(base + "_normals.txt", [[0, -1], [1, 0], [0, 1]], delimiter=" , ")
    pos, interior, bidx, bnorm, *_ = read_medusa_file(base)
    assert len(bidx) == 2 and len(bnorm) == 3 and len(bnorm[1]) == 0


def test_configs_sizes():
    from mfrbf import configs

    sides, phases = configs.c5_sizes()
    assert len(sides) == 64 and sides.min() >= 128 and sides.max() <= 512
    n = configs.c3_stencils(side=256)
    frac = np.mean(n > 30)
    assert 0.06 < frac < 0.10 and n.min() == 30 and set(np.unique(n[n > 30])) <= set(range(40, 51))


def test_ensemble_lane_packing():
    """Lanes hold consecutive clouds up to the node budget; a cloud above the
    budget is a lane of its own; every cloud lands in exactly one lane, in
    order.  C5's 64 clouds at 131,072 nodes per lane."""
    from mfrbf import configs
    from mfrbf.ensemble import pack_lanes

    assert pack_lanes([5, 5, 5], 10) == [[0, 1], [2]]
    assert pack_lanes([3, 20, 3, 3], 10) == [[0], [1], [2, 3]]
    assert pack_lanes([10, 10], 10) == [[0], [1]]
    assert pack_lanes([], 10) == []
    with pytest.raises(ValueError):
        pack_lanes([1], 0)
    sides, _ = configs.c5_sizes()
    sizes = [int(s) * int(s) for s in sides]
    lanes = pack_lanes(sizes, 131072)
    assert [r for g in lanes for r in g] == list(range(64))
    for g in lanes:
        total = sum(sizes[r] for r in g)
        assert total <= 131072 or len(g) == 1
    for a, b in zip(lanes[:-1], lanes[1:]):  # a lane closes only when the next cloud would overflow it
        assert sum(sizes[r] for r in a) + sizes[b[0]] > 131072


def test_gaussian_noise_moments():
    from mfrbf.clouds import gaussian_noise

    g = gaussian_noise(1, 200000, 1e-3)
    assert abs(g.mean()) < 1e-5 and abs(g.std() - 1e-3) < 1e-5


def test_c_client_builds_and_fails_loudly_without_device(tmp_path):
    """The C ABI from plain C99: tests/c_client/mfrbf_client.c compiles with
    warnings as errors against include/mfrbf.h and links libmfrbf.so; without
    a GPU it reports the missing device (exit 2) instead of computing."""
    import subprocess

    from conftest import build_c_client, gpu_available

    exe = build_c_client(tmp_path)
    if gpu_available():
        pytest.skip("a GPU is present: the GPU test runs the client")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "no device" in r.stderr


def _c_params(decl):
    """Parameter types of a C prototype's argument list, const and names
    dropped: 'const double* const* u' -> 'double**'."""
    out = []
    for p in decl.split(","):
        p = p.replace("const", " ").strip()
        if not p or p == "void":
            continue
        m = re.match(r"^(.*[\s*])(\w+)$", p)  # type, then the parameter name
        out.append(m.group(1).replace(" ", ""))
    return out


_JULIA_C = {"Ptr{Cvoid}": "ptr", "Ref{Ptr{Cvoid}}": "mf_ctx**", "Ptr{Cdouble}": "double*",
            "Ptr{Ptr{Cdouble}}": "double**", "Ptr{Int32}": "int32_t*", "Ptr{Int64}": "int64_t*",
            "Ptr{Int8}": "int8_t*", "Int64": "int64_t", "Cint": "int", "Cdouble": "double", "Cstring": "char*"}


def test_julia_ccalls_match_header():
    """Every ccall of the Julia engine (julia/MeshfreeTrixiHIP.jl, which runs
    only where Julia is installed) names a function of include/mfrbf.h with
    the same number of arguments, each of the C type its Julia type passes,
    and the same return type."""
    with open(os.path.join(ROOT, "include", "mfrbf.h")) as f:
        header = f.read()
    protos = {m.group(2): (m.group(1).replace("const", "").replace(" ", ""), _c_params(m.group(3)))
              for m in re.finditer(r"^(int|const char\*)\s+(mf_\w+)\(([^;]*?)\);", header, re.M | re.S)}
    with open(os.path.join(ROOT, "meshfreetrixi.jl_amd", "julia", "MeshfreeTrixiHIP.jl")) as f:
        jl = f.read()
    calls = re.findall(r"ccall\(\(:(mf_\w+),\s*libmfrbf\),\s*(\w+),\s*\(([^()]*)\)", jl, re.S)
    assert len(calls) >= 20
    for name, ret, args in calls:
        assert name in protos, name
        cret, cparams = protos[name]
        assert _JULIA_C[ret] == cret.replace("const", "") or (ret == "Cstring" and cret == "char*"), (name, ret, cret)
        jtypes = [a.strip() for a in args.split(",") if a.strip()]
        assert len(jtypes) == len(cparams), (name, jtypes, cparams)
        for jt, ct in zip(jtypes, cparams):
            jc = _JULIA_C[jt]
            if jc == "ptr":
                assert ct.endswith("*") and ct.count("*") == 1, (name, jt, ct)
            else:
                assert jc == ct, (name, jt, ct)
