"""GPU solvers and the staggered operator against the CPU oracle.

Solutions are compared with the oracle's FP64 solution (GMRES run to |r|/|b| < 1e-13);
the bar is the solver tolerance times a conditioning margin, written per test.
"""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

KAPPA = 0.1355
DIMS0 = (8, 8, 8, 8)


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c0():
    U = orc.gauge_random(DIMS0, 0.3, 1)
    b = orc.wilson_gaussian(DIMS0, 2)
    x_ref, rc, _, res = orc.gmres(0, orc.D, DIMS0, orc.BC_FERMION, KAPPA, U, b, max_dim=30,
                                  restart=200, accuracy=1e-13, stop_rule=1)
    assert rc == 0 and res < 1e-13
    return U, b, x_ref


def fields(ctx, U, b, kind=clg.WILSON):
    g = clg.Field(ctx, clg.GAUGE, clg.F64, DIMS0)
    g.upload(U)
    fb = clg.Field(ctx, kind, clg.F64, DIMS0, bc=orc.BC_FERMION)
    fb.upload(b)
    fx = clg.Field(ctx, kind, clg.F64, DIMS0, bc=orc.BC_FERMION)
    return g, fb, fx


def relerr(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("eo", [0, 1, 2])
def test_gmres12_f64_c0(ctx, c0, eo):
    U, b, x_ref = c0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=200, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=eo)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, clg.EFO_F_InverseD)
    print(f"eo={eo} cycles={rep.cycles} applies={rep.operator_applies} res={rep.true_residual:.2e}")
    assert st == clg.OK and rep.converged
    assert rep.true_residual < 1e-10
    x = fx.download()
    r = b - orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-10
    assert relerr(x, x_ref) < 1e-8   # |D^-1| <~ 1/(1-8 kappa) ~ 10 on this field


@pytest.mark.parametrize("inner", [clg.F32, clg.F16, clg.INNER_F16_VECTORS])
def test_mixed_precision_defect_correction(ctx, c0, inner):
    U, b, x_ref = c0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=50, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                          inner_prec=inner, inner_accuracy=1e-4 if inner == clg.F32 else 1e-2, max_outer=40)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, clg.EFO_F_InverseD)
    print(f"inner={inner} outer={rep.outer} cycles={rep.cycles} res={rep.true_residual:.2e}")
    assert st == clg.OK and rep.true_residual < 1e-10
    assert relerr(fx.download(), x_ref) < 1e-8


@pytest.mark.parametrize("eo,op", [(0, clg.EFO_F_InverseD), (1, clg.EFO_F_InverseDdagger), (0, clg.EFO_F_InverseDdagger)])
def test_fp16_vector_inner_solve_variants(ctx, c0, eo, op):
    # FP16-storage inner Krylov vectors (full lattice and Schur, D and D^dagger): the FP64 outer
    # loop must still reach |b - A x| / |b| < 1e-10, checked with the GPU FP64 operator
    U, b, x_ref = c0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=50, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=eo,
                          inner_prec=clg.INNER_F16_VECTORS, inner_accuracy=1e-2, max_outer=40)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, op)
    print(f"eo={eo} op={op} outer={rep.outer} cycles={rep.cycles} res={rep.true_residual:.2e}")
    assert st == clg.OK and rep.true_residual < 1e-10
    ax = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0, bc=orc.BC_FERMION)
    clg.wilson_apply(ctx, ax, fx, g, KAPPA, clg.EFO_F_D if op == clg.EFO_F_InverseD else clg.EFO_F_Ddagger)
    assert np.linalg.norm(ax.download() - b) / np.linalg.norm(b) < 1e-10
    if op == clg.EFO_F_InverseD:
        assert relerr(fx.download(), x_ref) < 1e-8


def test_fp16_vector_inner_rejects_unsupported(ctx, c0):
    U, b, _ = c0
    g, fb, fx = fields(ctx, U, b)
    for kw in (dict(max_dim=20), dict(even_odd=2), dict(solver=clg.SOLVER_BICGSTAB)):
        args = dict(max_dim=12, restart=50, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                    inner_prec=clg.INNER_F16_VECTORS, inner_accuracy=1e-2)
        args.update(kw)
        with pytest.raises(clg.CLGError):
            clg.solve(ctx, clg.solver_params(**args), fx, fb, g, KAPPA, clg.EFO_F_InverseD)


def test_reference_stop_rule_cycles_match_oracle(ctx, c0):
    U, b, _ = c0
    bl = float(np.dot(b, b))
    _, rc, cyc_ref, dev_ref = orc.gmres(0, orc.D, DIMS0, orc.BC_FERMION, KAPPA, U, b, max_dim=12, restart=50,
                                        accuracy=1e-20, blength=bl, stop_rule=0)
    assert rc == 0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=50, accuracy=1e-20, absolute_accuracy=0, stop_rule=clg.STOP_REFERENCE)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, clg.EFO_F_InverseD)
    assert st == clg.OK
    assert abs(rep.cycles - cyc_ref) <= 1
    assert rep.last_deviation < 1e-20 * bl


def test_not_converged_writes_best_iterate(ctx, c0):
    U, b, _ = c0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=2, restart=1, accuracy=1e-30)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, clg.EFO_F_InverseD)
    assert st == clg.NOT_CONVERGED and not rep.converged
    x = fx.download()
    r0 = np.linalg.norm(b - orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, b))
    r1 = np.linalg.norm(b - orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, x))
    assert r1 < r0   # x moved from x0 = b towards the solution


@pytest.mark.parametrize("eo", [0, 1])
def test_gmres_when_the_projections_cancel(ctx, c0, eo):
    # DD^dagger at kappa 0.005 is 1 + O(kappa): each Arnoldi vector A v_j is v_j up to a few 1e-2, so
    # the first Gram-Schmidt pass cancels nearly all of |w|^2 and every step takes the second pass
    # (the DGKS branch, not the single-pass fast form); the solution must still meet 1e-12
    U, b, _ = c0
    kappa = 0.005
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=50, accuracy=1e-13, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=eo)
    st, rep = clg.solve(ctx, p, fx, fb, g, kappa, clg.EFO_F_InverseDDdagger)
    assert st == clg.OK and rep.converged
    x = fx.download()
    Ax = orc.wilson_op(DIMS0, orc.BC_FERMION, kappa, U, x, clg.EFO_F_DDdagger - clg.EFO_F_D)
    assert np.linalg.norm(b - Ax) / np.linalg.norm(b) < 1e-12


@pytest.mark.parametrize("op", [clg.EFO_F_InverseDdagger, clg.EFO_F_InverseDDdagger, clg.EFO_F_InverseDD])
def test_inverse_composites(ctx, c0, op):
    U, b, _ = c0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=200, accuracy=1e-11, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, op)
    assert st == clg.OK
    x = fx.download()
    Ax = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, x, op - clg.EFO_F_InverseD)
    assert np.linalg.norm(b - Ax) / np.linalg.norm(b) < 1e-9


def test_x_aliases_b(ctx, c0):
    U, b, x_ref = c0
    g, fb, _ = fields(ctx, U, b)
    p = clg.solver_params(max_dim=12, restart=200, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
    st, rep = clg.solve(ctx, p, fb, fb, g, KAPPA, clg.EFO_F_InverseD)
    assert st == clg.OK and relerr(fb.download(), x_ref) < 1e-8


def test_bicgstab_reference_rule(ctx, c0):
    # |s|^2/|b|^2 < 1e-16 is reachable; BiCGStab stalls near |r|/|b| ~ 3e-9 on this field
    # (the oracle shows the same rho-restart), so the bar is set above that floor.
    U, b, x_ref = c0
    x_o, rc, it_o, dev_o = orc.bicgstab(0, orc.D, DIMS0, orc.BC_FERMION, KAPPA, U, b, step_count=100,
                                        dev_check=5, retry=3, accuracy=1e-16)
    assert rc == 0
    g, fb, fx = fields(ctx, U, b)
    p = clg.solver_params(solver=clg.SOLVER_BICGSTAB, max_step=100, deviation_step=5, restart=3, accuracy=1e-16)
    st, rep = clg.solve(ctx, p, fx, fb, g, KAPPA, clg.EFO_F_InverseD)
    assert st == clg.OK
    assert abs(rep.cycles - it_o) <= 10         # same algorithm; reduction order differs
    assert rep.true_residual < 1e-7
    assert relerr(fx.download(), x_ref) < 1e-6
    assert relerr(fx.download(), x_o) < 1e-6


# ------------------------------------------------------------------ staggered
@pytest.mark.parametrize("op", [clg.EFO_F_D, clg.EFO_F_Ddagger, clg.EFO_F_DD, clg.EFO_F_DDdagger])
def test_ks_parity(ctx, op):
    dims = (4, 4, 4, 8)
    U = orc.gauge_random(dims, 0.4, 1)
    chi = orc.ks_gaussian(dims, 2)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    fi = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    fi.upload(chi)
    fo = clg.Field(ctx, clg.KS, clg.F64, dims)
    clg.ks_apply(ctx, fo, fi, g, 0.05, op)
    ref = orc.ks_op(dims, orc.BC_FERMION, 0.05, U, chi, op)
    assert np.abs(fo.download() - ref).max() / np.abs(ref).max() < 1e-12
    g32 = clg.Field(ctx, clg.GAUGE12, clg.F32, dims)
    g32.upload(U)
    fi32 = clg.Field(ctx, clg.KS, clg.F32, dims, bc=orc.BC_FERMION)
    fi32.upload(chi)
    fo32 = clg.Field(ctx, clg.KS, clg.F32, dims)
    clg.ks_apply(ctx, fo32, fi32, g32, 0.05, op)
    assert np.abs(fo32.download() - ref).max() / np.abs(ref).max() < 1e-6
    # FP32 spinors with FP16 12-real links: the links carry 11-bit mantissas (relative 2^-11 per element)
    g16 = clg.Field(ctx, clg.GAUGE12, clg.F16, dims)
    g16.upload(U)
    fo16 = clg.Field(ctx, clg.KS, clg.F32, dims)
    clg.ks_apply(ctx, fo16, fi32, g16, 0.05, op)
    err16 = np.abs(fo16.download() - ref).max() / np.abs(ref).max()
    assert 1e-6 < err16 < 3e-3
    # the stored links, read back: within FP16 rounding of U (a misplaced row or element would be O(1));
    # against the oracle's operator on exactly those links the kernel is at FP32 accuracy
    U16 = g16.download()
    assert 1e-5 < np.abs(U16 - U).max() < 2e-3
    ref16 = orc.ks_op(dims, orc.BC_FERMION, 0.05, U16, chi, op)
    assert np.abs(fo16.download() - ref16).max() / np.abs(ref16).max() < 2e-6


def test_ks_solve_ddagger(ctx):
    dims = (4, 4, 4, 8)
    U = orc.gauge_random(dims, 0.4, 1)
    chi = orc.ks_gaussian(dims, 2)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    fb = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    fb.upload(chi)
    fx = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    p = clg.solver_params(max_dim=20, restart=100, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, kind=clg.KS)
    st, rep = clg.solve(ctx, p, fx, fb, g, 0.1, clg.EFO_F_InverseDDdagger)
    assert st == clg.OK
    x = fx.download()
    r = chi - orc.ks_op(dims, orc.BC_FERMION, 0.1, U, x, orc.DDDAG)
    assert np.linalg.norm(r) / np.linalg.norm(chi) < 1e-9


def test_ks_launch_configurations(ctx, monkeypatch):
    """Every staggered launch the tuner may pick gives the oracle's D (8^3x16 is large enough to tune)."""
    dims = (8, 8, 8, 16)
    U = orc.gauge_random(dims, 0.4, 5)
    chi = orc.ks_gaussian(dims, 6)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    fi = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    fi.upload(chi)
    ref = orc.ks_op(dims, orc.BC_FERMION, 0.07, U, chi, clg.EFO_F_D)
    for blk in (64, 128, 256):
        for xcd in (0, 8):
            for pm in (0, 2):
                monkeypatch.setenv("CLG_TUNE_FORCE_KS", f"{blk},{xcd},{pm}")
                fo = clg.Field(ctx, clg.KS, clg.F64, dims)
                clg.ks_apply(ctx, fo, fi, g, 0.07, clg.EFO_F_D)
                assert np.abs(fo.download() - ref).max() / np.abs(ref).max() < 1e-12, (blk, xcd, pm)
    monkeypatch.delenv("CLG_TUNE_FORCE_KS")
    clg.autotune_reset()
    clg.autotune_enable(1)
    fo = clg.Field(ctx, clg.KS, clg.F64, dims)
    clg.ks_apply(ctx, fo, fi, g, 0.07, clg.EFO_F_D)
    assert np.abs(fo.download() - ref).max() / np.abs(ref).max() < 1e-12
    assert any(e.startswith("staggered_d") and "dims=8x8x8x16" in e for e in clg.autotune_entries())
    clg.autotune_enable(-1)


# ------------------------------------------------------------------ staggered even-odd
KS_DIMS = (4, 4, 4, 8)


def test_ks_even_odd_operators(ctx):
    """The parity blocks of D D^dagger and the half-lattice hop against the oracle's full-lattice D:
    D D^dagger of a one-parity field has no component on the other parity (the oracle side pins that),
    and its own-parity part is the Schur block (2am)^2 - H H."""
    dims, m = KS_DIMS, 0.07
    V = orc.volume(dims)
    U = orc.gauge_random(dims, 0.4, 3)
    chi = orc.ks_gaussian(dims, 4).reshape(V, 6)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    sites = {par: orc.eo_sites(dims, par) for par in (0, 1)}
    subset = {0: clg.EVEN, 1: clg.ODD}
    for par in (0, 1):
        only = np.zeros_like(chi)
        only[sites[par]] = chi[sites[par]]
        dd = orc.ks_op(dims, orc.BC_FERMION, m, U, only.ravel(), clg.EFO_F_DDdagger).reshape(V, 6)
        assert np.abs(dd[sites[1 - par]]).max() < 1e-12 * np.abs(dd).max()
        fin = clg.Field(ctx, clg.KS, clg.F64, dims, subset=subset[par], bc=orc.BC_FERMION)
        fin.upload(chi[sites[par]].ravel())
        fout = clg.Field(ctx, clg.KS, clg.F64, dims, subset=subset[par], bc=orc.BC_FERMION)
        clg.ks_schur(ctx, fout, fin, g, m)
        want = dd[sites[par]]
        assert np.abs(fout.download().reshape(-1, 6) - want).max() < 1e-12 * np.abs(want).max()
        # hop from this parity to the other: (D only)_(1-par) = H only, since only vanishes there
        d1 = orc.ks_op(dims, orc.BC_FERMION, m, U, only.ravel(), clg.EFO_F_D).reshape(V, 6)[sites[1 - par]]
        xo = clg.Field(ctx, clg.KS, clg.F64, dims, subset=subset[1 - par], bc=orc.BC_FERMION)
        xo.upload(chi[sites[1 - par]].ravel())
        hout = clg.Field(ctx, clg.KS, clg.F64, dims, subset=subset[1 - par], bc=orc.BC_FERMION)
        clg.ks_hop(ctx, hout, fin, g)
        assert np.abs(hout.download().reshape(-1, 6) - d1).max() < 1e-12 * np.abs(d1).max()
        clg.ks_hop(ctx, hout, fin, g, a=-0.5, x=xo, b=0.3)
        want2 = -0.5 * d1 + 0.3 * chi[sites[1 - par]]
        assert np.abs(hout.download().reshape(-1, 6) - want2).max() < 1e-12 * np.abs(want2).max()
    with pytest.raises(clg.CLGError):
        clg.ks_schur(ctx, fout, fout, g, m)          # out aliases in


@pytest.mark.parametrize("op", [clg.EFO_F_InverseDDdagger, clg.EFO_F_InverseD, clg.EFO_F_InverseDdagger])
@pytest.mark.parametrize("inner", [clg.F64, clg.F32])
def test_ks_even_odd_solves(ctx, op, inner):
    dims, m = KS_DIMS, 0.1
    U = orc.gauge_random(dims, 0.4, 1)
    chi = orc.ks_gaussian(dims, 2)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    fb = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    fb.upload(chi)
    fx = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    p = clg.solver_params(max_dim=20, restart=100, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, kind=clg.KS,
                          even_odd=1, inner_prec=inner, inner_accuracy=1e-5)
    st, rep = clg.solve(ctx, p, fx, fb, g, m, op)
    assert st == clg.OK and rep.true_residual < 1e-10
    base = op - clg.EFO_F_InverseD
    r = chi - orc.ks_op(dims, orc.BC_FERMION, m, U, fx.download(), base)
    assert np.linalg.norm(r) / np.linalg.norm(chi) < 1e-9


@pytest.mark.parametrize("even_odd", [0, 1])
@pytest.mark.parametrize("op", [clg.EFO_F_InverseDDdagger, clg.EFO_F_InverseD])
def test_ks_fp16_link_inner_solves(ctx, even_odd, op):
    """FP64 defect correction around inner solves with FP32 vectors and FP16 12-real links: the FP64 true
    residual reaches 1e-10 and maps back to b through the oracle's operator."""
    dims, m = KS_DIMS, 0.1
    U = orc.gauge_random(dims, 0.4, 1)
    chi = orc.ks_gaussian(dims, 2)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.upload(U)
    fb = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    fb.upload(chi)
    fx = clg.Field(ctx, clg.KS, clg.F64, dims, bc=orc.BC_FERMION)
    p = clg.solver_params(max_dim=20, restart=100, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, kind=clg.KS,
                          even_odd=even_odd, inner_prec=clg.F16, inner_accuracy=1e-3, max_outer=60)
    st, rep = clg.solve(ctx, p, fx, fb, g, m, op)
    assert st == clg.OK and rep.true_residual < 1e-10 and rep.outer > 1
    base = op - clg.EFO_F_InverseD
    r = chi - orc.ks_op(dims, orc.BC_FERMION, m, U, fx.download(), base)
    assert np.linalg.norm(r) / np.linalg.norm(chi) < 1e-9
