"""GCRO-DR on the GPU against the oracle's restatement (SVD deflation) and the other solvers."""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

DIMS = (4, 4, 4, 4)
BC = orc.BC_FERMION


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


def up(ctx, kind, data):
    f = clg.Field(ctx, kind, clg.F64, DIMS, bc=BC)
    f.upload(data)
    return f


def test_gcrodr_svd_parity_with_oracle(ctx):
    kappa = 0.15
    U = orc.gauge_random(DIMS, 0.5, 1)
    b1, b2 = orc.wilson_gaussian(DIMS, 5), orc.wilson_gaussian(DIMS, 6)
    acc = 1e-11 / np.linalg.norm(b1)
    kw = dict(m=12, k=6, restart=300, accuracy=acc)
    x1o, rc1, c1o, _, urec, hu = orc.gcrodr(0, orc.D, DIMS, BC, kappa, U, b1, phase=0, **kw)
    x2o, rc2, c2o, _, _, _ = orc.gcrodr(0, orc.D, DIMS, BC, kappa, U, b2, phase=1, urec=urec, have_u=1, **kw)
    g = up(ctx, clg.GAUGE, U)
    fb1, fb2 = up(ctx, clg.WILSON, b1), up(ctx, clg.WILSON, b2)
    x1, x2 = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=BC), clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=BC)
    s = clg.Gcrodr(ctx, m_dim=12, k_dim=6, restart=300, accuracy=acc, deflate=clg.DEFLATE_SVD)
    st1, r1 = s.solve(x1, fb1, g, kappa, clg.EFO_F_D, phase=clg.PHASE_START)
    st2, r2 = s.solve(x2, fb2, g, kappa, clg.EFO_F_D, phase=clg.PHASE_IN)
    print("cycles gpu", r1.cycles, r2.cycles, "oracle", c1o, c2o, "res", r1.true_residual, r2.true_residual)
    assert st1 == clg.OK and st2 == clg.OK
    assert abs(r1.cycles - c1o) <= 2 and abs(r2.cycles - c2o) <= 2
    for x, xo in ((x1, x1o), (x2, x2o)):
        assert np.abs(x.download() - xo).max() < 1e-8 * np.abs(xo).max()
    assert r1.true_residual < 2e-11 and r2.true_residual < 2e-11
    s.close()


@pytest.mark.parametrize("deflate", [clg.DEFLATE_REV, clg.DEFLATE_HEV, clg.DEFLATE_SVD])
@pytest.mark.parametrize("even_odd", [0, 1])
def test_gcrodr_deflations_converge(ctx, deflate, even_odd):
    kappa = 0.15
    U = orc.gauge_random(DIMS, 0.5, 1)
    g = up(ctx, clg.GAUGE, U)
    s = clg.Gcrodr(ctx, m_dim=14, k_dim=6, restart=300, accuracy=1e-22, deflate=deflate, even_odd=even_odd)
    for i, phase in enumerate((clg.PHASE_START, clg.PHASE_IN, clg.PHASE_END)):
        b = orc.wilson_gaussian(DIMS, 20 + i)
        fb = up(ctx, clg.WILSON, b)
        x = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=BC)
        st, rep = s.solve(x, fb, g, kappa, clg.EFO_F_D, phase=phase)
        assert st == clg.OK, (deflate, even_odd, i)
        Dx = orc.wilson_op(DIMS, BC, kappa, U, x.download(), orc.D)
        assert np.linalg.norm(Dx - b) < 1e-9 * np.linalg.norm(b)
    s.close()


def test_gcrodr_staggered_dddagger(ctx):
    U = orc.gauge_random(DIMS, 0.5, 1)
    chi = orc.ks_gaussian(DIMS, 3)
    g, fb = up(ctx, clg.GAUGE, U), up(ctx, clg.KS, chi)
    x = clg.Field(ctx, clg.KS, clg.F64, DIMS, bc=BC)
    s = clg.Gcrodr(ctx, m_dim=12, k_dim=4, restart=300, accuracy=1e-22, kind=clg.KS)
    st, rep = s.solve(x, fb, g, 0.2, clg.EFO_F_DDdagger, phase=clg.PHASE_ONCE)
    assert st == clg.OK
    Ax = orc.ks_op(DIMS, BC, 0.2, U, x.download(), orc.DDDAG)
    assert np.linalg.norm(Ax - chi) < 1e-9 * np.linalg.norm(chi)
    s.close()


def test_gcrodr_bad_params(ctx):
    with pytest.raises(clg.CLGError):
        clg.Gcrodr(ctx, m_dim=4)
    with pytest.raises(clg.CLGError):
        clg.Gcrodr(ctx, m_dim=10, k_dim=9)


def test_context_closed_before_its_solver_and_graph():
    """Closing a context first destroys its GCRO-DR solvers and graphs (they hold device fields of the
    context); closing the Python objects afterwards, or finalizing a context that still owns a solver
    through the C ABI, must not touch freed memory."""
    c = clg.Context(0)
    U = orc.gauge_random(DIMS, 0.5, 1)
    g = up(c, clg.GAUGE, U)
    b = up(c, clg.WILSON, orc.wilson_gaussian(DIMS, 5))
    x = clg.Field(c, clg.WILSON, clg.F64, DIMS, bc=BC)
    s = clg.Gcrodr(c, m_dim=8, k_dim=4, restart=50, accuracy=1e-16)
    st, rep = s.solve(x, b, g, 0.12, clg.EFO_F_D, phase=clg.PHASE_START)   # leaves a recycled subspace
    assert st == clg.OK
    clg.Graph.begin(c)
    clg.scale(c, x, 0.5)
    gr = clg.Graph.end(c)
    gr.launch()
    c.sync()
    c.close()
    s.close()
    gr.close()
    del s, gr
    # C ABI: clg_finalize with a live solver destroys it
    L = clg.lib()
    import ctypes as C
    h = C.c_void_p()
    assert L.clg_init(0, C.byref(h)) == 0
    p = clg.GcrodrParams()
    assert L.clg_gcrodr_params_default(C.byref(p)) == 0
    sv = C.c_void_p()
    assert L.clg_gcrodr_create(h, C.byref(p), C.byref(sv)) == 0
    assert L.clg_finalize(h) == 0


def test_gcrodr_pays_near_critical(ctx):
    """A smooth field near its critical kappa, where restarted GMRES(12) stalls on the low modes: GCRO-DR's
    deflated restarts converge with well under half the operator applications, at the oracle's cycle count."""
    dims, kappa = (8, 8, 8, 8), 0.13
    U = orc.gauge_random(dims, 0.1, 1)
    b = orc.wilson_gaussian(dims, 2)
    acc = 1e-10
    _, rc_g, cyc_g, _ = orc.gmres(0, orc.D, dims, BC, kappa, U, b, max_dim=12, restart=3000, accuracy=acc, stop_rule=1)
    _, rc_o, cyc_o, _, _, _ = orc.gcrodr(0, orc.D, dims, BC, kappa, U, b, phase=0, m=20, k=10, restart=3000,
                                         accuracy=acc * np.linalg.norm(b), absolute=1)
    assert rc_g == 0 and rc_o == 0
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    fb = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
    fb.upload(b)
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
    s = clg.Gcrodr(ctx, m_dim=20, k_dim=10, restart=3000, accuracy=acc * np.linalg.norm(b), absolute_accuracy=1,
                   deflate=clg.DEFLATE_SVD)
    st, rep = s.solve(x, fb, g, kappa, clg.EFO_F_D, phase=clg.PHASE_START)
    s.close()
    pg = clg.solver_params(max_dim=12, restart=3000, accuracy=acc, stop_rule=clg.STOP_TRUE_RESIDUAL)
    x2 = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
    st2, rep2 = clg.solve(ctx, pg, x2, fb, g, kappa, clg.EFO_F_InverseD)
    print(f"GCRO-DR cycles {rep.cycles} (oracle {cyc_o}), applies {rep.operator_applies}; GMRES(12) applies "
          f"{rep2.operator_applies} (oracle cycles {cyc_g})")
    assert st == clg.OK and st2 == clg.OK
    assert abs(rep.cycles - cyc_o) <= 2
    assert rep.operator_applies < 0.5 * rep2.operator_applies
    Dx = orc.wilson_op(dims, BC, kappa, U, x.download(), orc.D)
    assert np.linalg.norm(Dx - b) < 2e-10 * np.linalg.norm(b)
