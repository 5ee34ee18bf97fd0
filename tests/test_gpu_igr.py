"""GPU parity of the IGR source (sources/IGR.jl): the conjugate-gradient
solve for sigma and its flux terms, through the C ABI against the CPU
oracle.  CG amplifies rounding differences, so the cases are chosen where
the iteration is stable to rounding: alpha = dx^2 (CG converges) and
5 dx^2 (indefinite, the residual grows), both 1e-10; the split-FP32 stencil
part keeps its 1e-6."""
import numpy as np
import pytest

import cases
import oracle as O
from conftest import rel_inf
from mfrbf import _lib as L
from mfrbf import configs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _gpu_rhs(c, prec, t=0.0):
    u = c.u0.copy()
    du = np.zeros_like(u)
    c.ctx.rhs_host(t, u, du, prec)
    return du


@pytest.mark.parametrize("prec,tol", [(L.MF_PREC_F64, 1e-10), (L.MF_PREC_F64_REF, 1e-10),
                                      (L.MF_PREC_SPLIT32, 1e-6)])
def test_rhs_igr(mf, prec, tol):
    _say("rhs euler + IGR alpha = dx^2 prec=%d" % prec)
    spec = configs.c2()
    c0 = cases.build(spec, side=48, use_gpu=False)
    c = cases.build(spec, side=48, igr=c0.dx_min ** 2)
    du = _gpu_rhs(c, prec, t=0.01)
    duo, _, _ = c.prob.rhs(c.u0.copy(), 0.01)
    sg, so = c.ctx.igr_sigma(), c.prob.igr_sigma
    _say("  du rel %.2e  sigma rel %.2e  |sigma| %.2e" % (rel_inf(du, duo), rel_inf(sg, so), np.abs(so).max()))
    assert np.abs(so).max() > 0.0
    assert rel_inf(sg, so) < 1e-10
    assert rel_inf(du, duo) < tol


def test_igr_tiles_change_no_result(mf):
    """The CG operator passes over the stencil kernel's LDS tiles (halo
    values staged once, 16-bit positions) give the bits of the global
    gathers: sigma and du, in every precision."""
    spec = configs.c2()
    c0 = cases.build(spec, side=64, use_gpu=False)
    c = cases.build(spec, side=64, igr=c0.dx_min ** 2)
    for prec in (L.MF_PREC_F64, L.MF_PREC_F64_REF, L.MF_PREC_SPLIT32):
        out = []
        for tiles in (1, 0):
            _say("igr tiles=%d prec=%d" % (tiles, prec))
            c.ctx.set_tuning(L.MF_TUNE_TILES, tiles)
            du = _gpu_rhs(c, prec)
            assert c.ctx.tile_info()["tiles_a"] == tiles
            out.append((du, c.ctx.igr_sigma()))
        c.ctx.set_tuning(L.MF_TUNE_TILES, 1)
        assert np.array_equal(out[0][1], out[1][1]) and np.abs(out[0][1]).max() > 0.0
        assert np.array_equal(out[0][0], out[1][0])


def test_rhs_igr_indefinite(mf):
    """alpha = 5 dx^2: the operator is indefinite and the CG residual grows
    (15x over 20 iterations), yet the iterates stay stable to rounding, and
    GPU and oracle agree to 1e-10.  At SourceIGR's default alpha = 1 the
    iterates grow by ~1e10 and their last bits are chaotic (a 1e-14 change of
    b moves sigma by ~1e-8 relative at 20 dx^2 already), so that case is run
    for completion and finiteness only."""
    _say("rhs euler + IGR alpha = 5 dx^2, then alpha = 1")
    spec = configs.c2()
    c0 = cases.build(spec, side=48, use_gpu=False)
    c = cases.build(spec, side=48, igr=5 * c0.dx_min ** 2)
    du = _gpu_rhs(c, L.MF_PREC_F64, t=0.01)
    duo, _, _ = c.prob.rhs(c.u0.copy(), 0.01)
    _say("  du rel %.2e  sigma rel %.2e" % (rel_inf(du, duo), rel_inf(c.ctx.igr_sigma(), c.prob.igr_sigma)))
    assert rel_inf(c.ctx.igr_sigma(), c.prob.igr_sigma) < 1e-10
    assert rel_inf(du, duo) < 1e-10
    c.ctx.set_source_igr(1.0, 20)
    du = _gpu_rhs(c, L.MF_PREC_F64, t=0.01)
    assert np.all(np.isfinite(du)) and np.all(np.isfinite(c.ctx.igr_sigma()))


def test_rhs_igr_after_viscosity_and_order(mf):
    """Upwind viscosity then IGR (the IGR source must come last: the other
    order is refused)."""
    _say("rhs euler + upwind viscosity + IGR")
    spec = configs.c2()
    c0 = cases.build(spec, side=40, use_gpu=False)
    c = cases.build(spec, side=40, visc="upwind", igr=c0.dx_min ** 2, order=("visc", "igr"))
    du = _gpu_rhs(c, L.MF_PREC_F64, t=0.001)
    duo, _, _ = c.prob.rhs(c.u0.copy(), 0.001)
    assert rel_inf(du, duo) < 1e-10
    c.ctx.set_source_order([L.MF_SRC_IGR, L.MF_SRC_VISCOSITY])
    with pytest.raises(L.MfError) as e:
        _gpu_rhs(c, L.MF_PREC_F64)
    assert e.value.code == L.MF_ERR_UNSUPPORTED


def test_ssprk33_igr(mf):
    """Four SSPRK33 steps with the IGR source (its last kernel writes the
    stages), eager and as a graph."""
    _say("ssprk33 euler + IGR 4 steps")
    spec = configs.c2()
    c0 = cases.build(spec, side=40, use_gpu=False)
    c = cases.build(spec, side=40, igr=c0.dx_min ** 2)
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, 4)
    for flags in (0, L.MF_STEP_GRAPH):
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, c.dt, 4, flags=flags)
        assert rel_inf(c.ctx.download(L.MF_BUF_U), uo) < 1e-10, flags
