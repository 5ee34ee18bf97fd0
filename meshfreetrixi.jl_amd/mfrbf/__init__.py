"""mfrbf - MI355X-native RBF-FD semidiscretization (MeshfreeTrixi.jl's hot path).

Python host mirror of the reference's Trixi-style API over libmfrbf.so:
PointCloudBasis / PointCloudSolver / PointCloudDomain /
SemidiscretizationHyperbolic / semidiscretize / rhs_ (Trixi.rhs!) /
solve with HistoryCallback and the Zhang-Shu stage limiter, and trixi2vtk
output.  All numerical work runs in HIP kernels on the GPU.
"""
from . import _lib
from ._lib import Context, MfError
from .basis import (HybridGaussianPHS, Point1D, Point2D, Point3D, PointCloudBasis,
                    PolyharmonicSpline, RBF, RefPointData, setproperties)
from .domain import (BoundaryData, PointCloudDomain, PointData, SerialPointCloudDomain,
                     read_medusa_file)
from .equations import (BoundaryConditionDirichlet, BoundaryConditionDoNothing,
                        CompressibleEulerEquations2D, ConstantState, DensityWave,
                        GaussianPulse, LinearGradient, LinearScalarAdvectionEquation2D,
                        boundary_condition_slip_wall, initial_condition_density_wave)
from .semidiscretization import (CallbackSet, HistoryCallback, ODEProblem, ODESolution,
                                 PositivityPreservingLimiterZhangShu, SemidiscretizationHyperbolic,
                                 SSPRK33, limiter_zhang_shu_, rhs_, rhs_vjp, semidiscretize,
                                 solve, ssprk33_gradient)
from .solver import AbstractRBFEngine, PointCloudSolver, RBFFDEngineHIP, RBFSolver
from .sources import (SourceHyperviscosityFlyer, SourceHyperviscosityTominec, SourceIGR,
                      SourceResidualViscosityTominec, SourceTerms, SourceUpwindViscosityTominec)
from .vtk import trixi2vtk

__all__ = [n for n in dir() if not n.startswith("_")]
