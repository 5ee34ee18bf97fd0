"""Point-cloud domain: PointData (kNN on the GPU), BoundaryData,
SerialPointCloudDomain, PointCloudDomain and the Medusa reader.

Mirrors domains/PointCloudDomain/geometry_primatives.jl:294-364,
SerialPointCloud.jl:12-54, PointCloudDomain.jl:23-48 and
auxiliary/medusa/read_medusa_file.jl:2-64.  Indices are 0-based.
"""
import os

import numpy as np

from .solver import RBFSolver, default_engine


class PointData:
    """points (N, 2), neighbors (N, nv) with self first, dx_min, dx_avg
    (geometry_primatives.jl:297-304)."""

    def __init__(self, points, neighbors, dx_min, dx_avg):
        self.points = np.ascontiguousarray(points, dtype=np.float64)
        self.neighbors = np.ascontiguousarray(neighbors, dtype=np.int32)
        self.num_points = self.points.shape[0]
        self.num_neighbors = self.neighbors.shape[1]
        self.dx_min = float(dx_min)
        self.dx_avg = float(dx_avg)

    @classmethod
    def from_basis(cls, points, basis, engine=None, dx_min=None, dx_avg=None):
        """PointData(points, basis) (geometry_primatives.jl:322-359): kNN with
        nv = basis.nv neighbours, on the GPU."""
        points = np.ascontiguousarray(points, dtype=np.float64)
        ctx = (engine or default_engine()).context()
        ctx.set_points(points[:, 0], points[:, 1])
        idx, _, mn, av = ctx.knn(basis.nv)
        return cls(points, idx, mn if dx_min is None else dx_min, av if dx_avg is None else dx_avg)


class BoundaryData:
    """idx (nb,) and outward normals (nb, 2) of one boundary group
    (geometry_primatives.jl:361-364)."""

    def __init__(self, idx, normals):
        self.idx = np.ascontiguousarray(idx, dtype=np.int32)
        self.normals = np.ascontiguousarray(np.asarray(normals, dtype=np.float64).reshape(-1, 2))


class SerialPointCloudDomain:
    """SerialPointCloudDomain(pd, boundary_tags) (SerialPointCloud.jl:12-29)."""

    def __init__(self, pd, boundary_tags):
        self.pd = pd
        self.boundary_tags = dict(boundary_tags)
        self.unsaved_changes = False

    @property
    def ndims(self):
        return self.pd.points.shape[1]


def read_medusa_file(casename):
    """read_medusa_file (auxiliary/medusa/read_medusa_file.jl:2-64).
    Returns positions (N,2), interior_idx, boundary_idxs (non-empty groups
    -1, -2, ...), boundary_normals (one list per group, empty ones kept, as
    the reference does), normals, boundary_ranges, types; 0-based."""
    positions = np.loadtxt(casename + "_positions.txt", delimiter=",", dtype=np.float64, ndmin=2)
    types = np.loadtxt(casename + "_types.txt", dtype=np.int64, ndmin=1)
    boundary_idx = np.loadtxt(casename + "_boundary.txt", dtype=np.int64, ndmin=1)
    interior_idx = np.loadtxt(casename + "_interior.txt", dtype=np.int64, ndmin=1)
    normals = np.loadtxt(casename + "_normals.txt", delimiter=",", dtype=np.float64, ndmin=2)
    num_bound = int(-types.min())
    idxs = [[] for _ in range(num_bound)]
    nrms = [[] for _ in range(num_bound)]
    for j, b in enumerate(boundary_idx):
        for i in range(1, num_bound + 1):
            if types[b] == -i:
                idxs[i - 1].append(int(b))
                nrms[i - 1].append(normals[j])
    boundary_idxs = [np.asarray(g, dtype=np.int32) for g in idxs if len(g) > 0]
    boundary_normals = [np.asarray(g, dtype=np.float64).reshape(-1, 2) for g in nrms]
    lens = [len(g) for g in boundary_idxs]
    ranges, start = [], 0
    for n in lens:
        ranges.append(np.arange(start, start + n))
        start += n
    return positions, interior_idx, boundary_idxs, boundary_normals, normals, ranges, types


def PointCloudDomain(solver_or_basis, casename, boundary_names):
    """PointCloudDomain(solver, casename, Dict(name => group)) for a Medusa
    point set (types.jl:133-145, PointCloudDomain.jl:23-48,
    SerialPointCloud.jl:45-54).  Group numbers are 1-based like the
    reference's."""
    if isinstance(solver_or_basis, RBFSolver):
        basis, engine = solver_or_basis.basis, solver_or_basis.engine
    else:
        basis, engine = solver_or_basis, None
    if not os.path.exists(casename + "_positions.txt"):
        raise FileNotFoundError(casename + "_positions.txt")
    positions, _, bidx, bnorm, _, _, _ = read_medusa_file(casename)
    pd = PointData.from_basis(positions, basis, engine)
    tags = {name: BoundaryData(bidx[g - 1], bnorm[g - 1]) for name, g in boundary_names.items()}
    return SerialPointCloudDomain(pd, tags)
