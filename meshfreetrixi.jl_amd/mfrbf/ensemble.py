"""Ensembles of independent point clouds in one context.

The reference runs one cloud per semidiscretization; an ensemble of
independent runs there is a loop over semidiscretizations.  Here the clouds
are concatenated into one grouped point set (mf_set_points_grouped), so that
every kernel launch of rhs! and of the captured SSPRK33 step covers all of
them:
  - each cloud's stencils come from its own kNN search (a scratch context
    per cloud size) and are offset into the concatenated numbering;
  - the weights are computed once over the whole set (each stencil's
    weights depend only on its own nodes);
  - each cloud keeps its own boundary tags (Dirichlet values with its own
    parameters) and its own Flyer hyperviscosity coefficient
    gamma_r = c dx_min,r^(2k), folded into the rows of the operator;
  - the clouds advance with one common time step, the smallest of their
    CFL steps (a lockstep ensemble).
"""
import numpy as np

from . import _lib as L
from .clouds import cloud_square
from .configs import euler_dt
from .equations import CompressibleEulerEquations2D


class Ensemble:
    """Builds one grouped context from Euler specs (mfrbf.configs) with
    Dirichlet inflow (the spec's analytic state) and do-nothing outflow, and
    the Flyer hyperviscosity (spec.hv = dict(k, c)) when given."""

    def __init__(self, specs, device=0):
        self.specs = list(specs)
        eq = CompressibleEulerEquations2D(self.specs[0].gamma)
        scratch = L.Context(device)
        xs, ys, idxs, u0s, tags = [], [], [], [], []
        self.offsets = [0]
        self.dx_min, self.dx_avg, dts = [], [], []
        for spec in self.specs:
            x, y, tag, nx, ny = cloud_square(spec.side, spec.seed)
            scratch.set_points(x, y)
            idx, _, dmin, davg = scratch.knn(spec.k)
            u0 = np.asarray(spec.ic(np.c_[x, y], 0.0, eq))
            xs.append(x)
            ys.append(y)
            idxs.append(idx + self.offsets[-1])
            u0s.append(u0)
            tags.append(tag)
            self.dx_min.append(dmin)
            self.dx_avg.append(davg)
            dts.append(euler_dt(spec, u0, dmin))
            self.offsets.append(self.offsets[-1] + len(x))
        del scratch
        self.x, self.y = np.concatenate(xs), np.concatenate(ys)
        self.N = len(self.x)
        self.u0 = np.ascontiguousarray(np.concatenate(u0s, axis=1))
        self.dt = float(min(dts))
        self.cloud_dt = dts
        k = self.specs[0].k
        ctx = L.Context(device)
        self.ctx = ctx
        ctx.set_equations(L.MF_EQ_EULER_2D, self.specs[0].gamma)
        ctx.set_points_grouped(self.x, self.y, self.offsets)
        ctx.set_stencil(np.ascontiguousarray(np.concatenate(idxs)))
        spec0 = self.specs[0]
        ctx.compute_operator(spec0.p, spec0.deg, 1, L.MF_OP_FLUX, want=False)
        if spec0.hv is not None:
            hk, hc = spec0.hv["k"], spec0.hv["c"]
            hv, _ = ctx.compute_operator(spec0.p, spec0.deg, 2 * hk, L.MF_OP_HYPERVISCOSITY, want=True, k=k)
            gam = np.concatenate([np.full(self.offsets[r + 1] - self.offsets[r], hc * self.dx_min[r] ** (2 * hk))
                                  for r in range(len(self.specs))])
            self.gamma = gam
            ctx.set_operator(L.MF_OP_HYPERVISCOSITY, hv * gam[:, None])
            ctx.set_source_hyperviscosity(1.0)
        # boundary: per cloud an inflow tag (Dirichlet, the cloud's own
        # analytic state) and an outflow tag (do nothing)
        ptr, kinds, fns, params, nodes = [0], [], [], [], []
        for r, spec in enumerate(self.specs):
            tag = tags[r]
            for code, kind in ((1, L.MF_BC_DIRICHLET), (2, L.MF_BC_DO_NOTHING)):
                sel = np.nonzero(tag == code)[0] + self.offsets[r]
                nodes.append(sel.astype(np.int32))
                ptr.append(ptr[-1] + len(sel))
                kinds.append(kind)
                fns.append(spec.ic.code)
                params.append(spec.ic.params())
        ctx.set_boundary(ptr, kinds, fns, np.concatenate(params), np.concatenate(nodes),
                         np.zeros((ptr[-1], 2)))

    def cloud(self, r, u):
        """Cloud r's part of a state (nvar, N)."""
        return u[:, self.offsets[r]:self.offsets[r + 1]]


def pack_lanes(sizes, lane_nodes):
    """Consecutive clouds (node counts `sizes`) packed into lanes of at most
    `lane_nodes` nodes; a cloud larger than that is a lane of its own.
    Returns the clouds of each lane as lists of indices, in order."""
    if lane_nodes < 1:
        raise ValueError("pack_lanes: lane_nodes must be positive")
    lanes, cur, n = [], [], 0
    for r, m in enumerate(sizes):
        if cur and n + m > lane_nodes:
            lanes.append(cur)
            cur, n = [], 0
        cur.append(r)
        n += m
    if cur:
        lanes.append(cur)
    return lanes


class EnsembleLanes:
    """The same ensemble as lanes: consecutive clouds packed into grouped
    contexts of at most `lane_nodes` nodes (a larger cloud is a lane of its
    own), each lane on its own stream, all stepped with the ensemble's
    common time step and enqueued before any is synchronized.

    One grouped context streams every cloud's operators once per stage; a
    lane of moderate size instead keeps its clouds' operators in the
    last-level cache from one stage to the next while the concurrent lanes
    fill the GPU.  Each cloud's state equals the one grouped context's bit
    for bit (the sources are node-local and every lane uses the same time
    step).  The default, 131,072 nodes per lane (about 63 MB of operators
    and state per stage in split-FP32 with Flyer hyperviscosity), measured
    best for both the 64-cloud ensemble (16k..262k nodes per cloud) and 256
    clouds of 48 x 48 nodes."""

    def __init__(self, specs, lane_nodes=131072, device=0):
        self.specs = list(specs)
        groups = pack_lanes([spec.side * spec.side for spec in self.specs], lane_nodes)
        self.lanes = [Ensemble([self.specs[r] for r in g], device) for g in groups]
        self.dt = min(lane.dt for lane in self.lanes)
        self.N = sum(lane.N for lane in self.lanes)

    def upload_initial(self):
        for lane in self.lanes:
            lane.ctx.upload(L.MF_BUF_U, lane.u0)

    def ssprk33(self, t0, n_steps, precision=L.MF_PREC_SPLIT32, flags=L.MF_STEP_GRAPH):
        """n_steps of every lane from t0 at the common time step; returns
        after all lanes finished unless flags holds MF_STEP_NOSYNC."""
        for lane in self.lanes:
            lane.ctx.ssprk33(t0, self.dt, n_steps, precision, flags | L.MF_STEP_NOSYNC)
        if not flags & L.MF_STEP_NOSYNC:
            self.synchronize()

    def synchronize(self):
        for lane in self.lanes:
            lane.ctx.synchronize()

    def download(self):
        """The ensemble's state (nvar, N), clouds in the order of `specs`."""
        return np.concatenate([lane.ctx.download(L.MF_BUF_U) for lane in self.lanes], axis=1)
