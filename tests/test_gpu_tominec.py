"""GPU parity of the Tominec hyperviscosity source (sources/hyperviscosity.jl:
66-134) and of the order of the sources (generic_sources.jl:7-18):
libmfrbf through the C ABI against the CPU oracle on the same seeded inputs.
Tolerances as for the other rhs! tests: FP64 1e-12 norm-wise relative,
split-FP32 1e-6; the reference-order mode bit-exact where no boundary
function (device sin/exp) is evaluated."""
import numpy as np
import pytest

import cases
import oracle as O
from conftest import rel_inf
from mfrbf import _lib as L
from mfrbf import configs

pytestmark = pytest.mark.gpu

TOL = {L.MF_PREC_F64: 1e-12, L.MF_PREC_F64_REF: 1e-12, L.MF_PREC_SPLIT32: 1e-6}
PRECS = [L.MF_PREC_F64, L.MF_PREC_F64_REF, L.MF_PREC_SPLIT32]


def _say(msg):
    print(msg, flush=True)


def _gpu_rhs(c, prec, t=0.0):
    u = c.u0.copy()
    du = np.zeros_like(u)
    c.ctx.rhs_host(t, u, du, prec)
    return du


def _oracle_rhs(c, t=0.0):
    return c.prob.rhs(c.u0.copy(), t)[0]


def _tominec_part(c, prec, t=0.0):
    """du with minus du without the Tominec source, on both sides."""
    dg1, do1 = _gpu_rhs(c, prec, t), _oracle_rhs(c, t)
    c.ctx.set_source_hyperviscosity_tominec(0.0)
    c.prob.cfg.lap = O.d(None)
    dg0, do0 = _gpu_rhs(c, prec, t), _oracle_rhs(c, t)
    c.ctx.set_source_hyperviscosity_tominec(c.tom_gamma)
    c.prob.cfg.lap = O.d(c.prob.lap)
    return dg1, do1, dg1 - dg0, do1 - do0


@pytest.mark.parametrize("prec", PRECS)
def test_rhs_tominec(mf, prec):
    _say("rhs euler + tominec prec=%d" % prec)
    c = cases.build(configs.c2(), side=64, tom=1.0)
    dg, do, pg, po = _tominec_part(c, prec, t=0.01)
    scale = np.abs(do).max()
    _say("  |du| %.3e  |tominec part| %.3e  rel err %.2e  part err %.2e" %
         (scale, np.abs(po).max(), rel_inf(dg, do), np.abs(pg - po).max() / scale))
    assert np.abs(po).max() > 1e-6 * scale  # the source is visible in du
    assert rel_inf(dg, do) < TOL[prec]
    assert np.abs(pg - po).max() < TOL[prec] * scale


@pytest.mark.parametrize("prec", PRECS)
def test_rhs_flyer_and_tominec(mf, prec):
    """Both hyperviscosities, as sources hv and hv2 of the reference's
    hyperviscosity test (test/hyperviscosity_test.jl:42-45)."""
    _say("rhs euler + flyer + tominec prec=%d" % prec)
    c = cases.build(configs.c2(), side=45, hv=(2, 1.0), tom=1.0)  # N % 64 != 0: a partial last slice
    assert rel_inf(_gpu_rhs(c, prec), _oracle_rhs(c)) < TOL[prec]


@pytest.mark.parametrize("order", [("tominec", "visc"), ("visc", "tominec")])
@pytest.mark.parametrize("prec", PRECS)
def test_rhs_tominec_upwind(mf, order, prec):
    """Tominec with the upwind viscosity, in both keyword orders (the
    reference's upwind viscosity test puts Tominec first,
    test/upwind_viscosity_test.jl:50-52)."""
    _say("rhs euler + tominec + upwind order=%s prec=%d" % (order, prec))
    c = cases.build(configs.c2(), side=48, tom=1.0, visc="upwind", order=order)
    du = _gpu_rhs(c, prec, t=0.001)
    duo, eps, _ = c.prob.rhs(c.u0.copy(), 0.001)
    geps, _ = c.ctx.viscosity()
    assert rel_inf(geps, eps) < 1e-12
    assert rel_inf(du, duo) < TOL[prec]


def test_rhs_tominec_after_residual(mf):
    """Residual viscosity (with a history, so that eps_rv is active) followed
    by Tominec; Tominec before the residual viscosity is refused."""
    _say("rhs euler + residual viscosity + tominec")
    c = cases.build(configs.c2(), side=48, tom=1.0, visc="residual", order=("visc", "tominec"))
    u1 = c.u0 * 1.0001
    th = np.zeros(4)
    sh = np.zeros(4 * c.u0.size)
    ad = np.zeros(c.u0.size)
    O.history_update(3, 0, 0.0, c.u0.ravel().copy(), th, sh, ad)
    O.history_update(3, 1, 0.001, u1.ravel().copy(), th, sh, ad)
    c.prob.approx_du[:] = ad
    c.prob.cfg.success_iter = 1
    c.ctx.upload(L.MF_BUF_U, c.u0.copy())
    c.ctx.history_update(0.0, 0)
    c.ctx.upload(L.MF_BUF_U, np.ascontiguousarray(u1))
    c.ctx.history_update(0.001, 1)
    duo, eps, epsc = c.prob.rhs(c.u0.copy(), 0.001)
    for prec in PRECS:
        du = _gpu_rhs(c, prec, t=0.001)
        _say("  prec=%d rel err %.2e" % (prec, rel_inf(du, duo)))
        assert rel_inf(du, duo) < TOL[prec]
        geps, gcode = c.ctx.viscosity()  # from this precision's du
        assert np.array_equal(gcode, epsc) and rel_inf(geps, eps) < (1e-12 if prec != L.MF_PREC_SPLIT32 else 1e-6)
    c.ctx.set_source_order([L.MF_SRC_TOMINEC, L.MF_SRC_VISCOSITY])
    with pytest.raises(L.MfError) as e:
        _gpu_rhs(c, L.MF_PREC_F64)
    assert e.value.code == L.MF_ERR_UNSUPPORTED


@pytest.mark.parametrize("order", [("tominec", "flyer"), ("flyer", "tominec")])
def test_rhs_reference_order_tominec_bitwise(mf, order):
    """Reference-order FP64: lap u in ascending column order, lap' L over
    ascending rows, then du + tmp * (-gamma), each source in keyword order
    (a Flyer source after Tominec is added after it, from its own gather)."""
    _say("rhs f64_ref tominec order=%s" % (order,))
    c = cases.build(configs.c2(), side=48, bc=("nothing", "nothing"), hv=(2, 1.0), tom=1.0, order=order)
    du = _gpu_rhs(c, L.MF_PREC_F64_REF, t=0.02)
    duo = _oracle_rhs(c, t=0.02)
    assert np.array_equal(du, duo)


@pytest.mark.parametrize("visc", [None, "upwind"])
def test_tominec_tiles_change_no_result(mf, visc):
    """LDS-staged halos (lap u records, and with the viscosity G and lap u
    records) give the same bits as the global-memory gathers."""
    _say("tominec tiles on/off visc=%s" % visc)
    c = cases.build(configs.c2(), side=64, tom=1.0, visc=visc)
    for prec in PRECS:
        a = _gpu_rhs(c, prec, t=0.001)
        info = c.ctx.tile_info()
        _say("  prec=%d tiles %s" % (prec, info))
        assert info["tiles_a"] == 1 and info["tiles_b"] == 1  # both phases tiled at this size
        c.ctx.set_tuning(L.MF_TUNE_TILES, 0)
        b = _gpu_rhs(c, prec, t=0.001)
        c.ctx.set_tuning(L.MF_TUNE_TILES, 1)
        assert np.array_equal(a, b), prec


def test_ssprk33_tominec_upwind(mf):
    """Eight SSPRK33 steps with Tominec and the upwind viscosity (fused stage
    updates in the transposed pass), eager and as a graph."""
    _say("ssprk33 euler + tominec + upwind 8 steps")
    c = cases.build(configs.c2(), side=40, tom=1.0, visc="upwind", order=("tominec", "visc"))
    n = 8
    uo = c.u0.copy()
    c.prob.ssprk33(uo, 0.0, c.dt, n)
    for flags in (0, L.MF_STEP_GRAPH):
        c.ctx.upload(L.MF_BUF_U, c.u0.copy())
        c.ctx.ssprk33(0.0, c.dt, n, flags=flags)
        ug = c.ctx.download(L.MF_BUF_U)
        assert rel_inf(ug, uo) < 1e-12, flags


def test_tominec_operator_on_gpu(mf):
    """lap computed on the device (deriv order 2) against the oracle's."""
    _say("tominec lap weights on the GPU")
    c = cases.build(configs.c2(), side=40, tom=1.0, weights="gpu")
    wx, wy = c.ctx.compute_operator(5, 3, 2, L.MF_OP_LAPLACIAN, want=True, k=c.k)
    assert rel_inf(wx + wy, c.lap_w) < 1e-9
    c.ctx.set_source_hyperviscosity_tominec(c.tom_gamma)
    assert rel_inf(_gpu_rhs(c, L.MF_PREC_F64), _oracle_rhs(c)) < 1e-10
