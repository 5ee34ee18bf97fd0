"""Solver types and the HIP engine.

Mirrors solvers/rbfsolver.jl:8-71 (AbstractRBFEngine, RBFFDEngine,
RBFSolver{Basis, Engine}) and solvers/pointcloudsolver/types.jl:49-74
(PointCloudSolver constructors).  RBFFDEngineHIP is the engine slot the
reference reserves for a GPU backend (the empty RBFFDEngineCUDA,
rbfsolver.jl:60-61): every operator build and rhs! of a solver with this
engine runs in libmfrbf on the GPU.
"""
from . import _lib as L
from .basis import RBF, RefPointData


class AbstractRBFEngine:
    pass


class RBFFDEngineHIP(AbstractRBFEngine):
    """RBF-FD engine on an MI355X (gfx950) device through libmfrbf."""

    def __init__(self, device=0):
        self.device = device
        self._ctx = None

    def context(self):
        """A libmfrbf context for setup work (kNN) on this engine's device."""
        if self._ctx is None:
            self._ctx = L.Context(self.device)
        return self._ctx

    def new_context(self):
        return L.Context(self.device)

    def __repr__(self):
        return "RBFFDEngineHIP(device=%d)" % self.device


class RBFSolver:
    """RBFSolver{Basis, Engine} (rbfsolver.jl:68-71)."""

    def __init__(self, basis, engine=None):
        if not isinstance(basis, RefPointData):
            raise TypeError("basis must be a RefPointData")
        self.basis = basis
        self.engine = engine if engine is not None else RBFFDEngineHIP()

    def __repr__(self):
        return "RBFSolver(%r, %r)" % (self.basis, self.engine)


def PointCloudSolver(basis=None, engine=None, polydeg=None, element_type=None,
                     approximation_type=None):
    """PointCloudSolver(basis; engine) or PointCloudSolver(; polydeg,
    element_type, approximation_type, engine) (types.jl:49-74)."""
    if basis is None:
        if polydeg is None or element_type is None:
            raise TypeError("PointCloudSolver needs a basis or polydeg and element_type")
        basis = RefPointData(element_type, approximation_type or RBF(), polydeg)
    return RBFSolver(basis, engine)


_default_engine = None


def default_engine():
    global _default_engine
    if _default_engine is None:
        _default_engine = RBFFDEngineHIP()
    return _default_engine
