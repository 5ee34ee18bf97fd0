"""Reference element / basis data: RefPointData and its constructors.

Mirrors domains/PointCloudDomain/geometry_primatives.jl:1-292 and
solvers/pointcloudsolver/types.jl:34-86 of MeshfreeTrixi.jl.
"""
from math import comb


class Point1D:
    ndims = 1


class Point2D:
    ndims = 2


class Point3D:
    ndims = 3


class PolyharmonicSpline:
    """Odd-order polyharmonic spline r^Nrbf (geometry_primatives.jl:110-115)."""

    def __init__(self, Nrbf=3):
        self.Nrbf = int(Nrbf)

    def __repr__(self):
        return "PolyharmonicSpline(%d)" % self.Nrbf


class HybridGaussianPHS:
    """alpha*exp(-(eps r)^2) + beta*r^Nrbf (geometry_primatives.jl:118-132).
    Declared for API completeness; the GPU weights support the PHS kind only."""

    def __init__(self, Nrbf=3, alpha=1.0, beta=1.0, epsilon=1.0):
        self.Nrbf, self.alpha, self.beta, self.epsilon = Nrbf, alpha, beta, epsilon


class RBF:
    """RBF approximation type; RBF() is the order-3 polyharmonic spline
    (geometry_primatives.jl:97-107, :173-187)."""

    def __init__(self, rbf_type=None):
        self.rbf_type = rbf_type if rbf_type is not None else PolyharmonicSpline(3)
        self.Nrbf = self.rbf_type.Nrbf

    def __repr__(self):
        return "RBF(%r)" % (self.rbf_type,)


_MIN_NV = {1: 10, 2: 15, 3: 20}


class RefPointData:
    """Degree-N RBF-FD basis on a point element with nv-point stencils
    (geometry_primatives.jl:27-35)."""

    def __init__(self, element_type, approximation_type, N, nv=None):
        self.element_type = element_type
        self.approximation_type = approximation_type
        self.N = int(N)
        d = element_type.ndims
        # stencil size rule, geometry_primatives.jl:183-184 / :197-198
        self.nv = int(nv) if nv is not None else max(2 * comb(self.N + d, d), _MIN_NV[d])

    @property
    def ndims(self):
        return self.element_type.ndims

    @property
    def phs_power(self):
        return self.approximation_type.Nrbf

    @property
    def npoly(self):
        d = self.ndims
        return comb(self.N + d, d)

    def __repr__(self):
        return "RefPointData{N=%d, %r, nv=%d}" % (self.N, self.approximation_type, self.nv)


def setproperties(rd, **patch):
    """ConstructionBase.setproperties(rd, (; nv=30)) (geometry_primatives.jl:38-41)."""
    f = dict(element_type=rd.element_type, approximation_type=rd.approximation_type, N=rd.N, nv=rd.nv)
    f.update(patch)
    return RefPointData(f["element_type"], f["approximation_type"], f["N"], f["nv"])


def PointCloudBasis(element_type, polydeg, approximation_type=None, **kwargs):
    """PointCloudBasis(element_type, polydeg; approximation_type=RBF())
    (solvers/pointcloudsolver/types.jl:83-86)."""
    at = approximation_type if approximation_type is not None else RBF()
    return RefPointData(element_type, at, polydeg, kwargs.get("nv"))
