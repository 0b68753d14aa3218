"""GPU multi-shift solvers and rational approximation against the CPU oracle."""
import json
import os

import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REMEZ = json.load(open(os.path.join(HERE, "golden", "remez_ks.json")))
DIMS = (4, 4, 4, 8)


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


def setup(ctx, kind):
    U = orc.gauge_random(DIMS, 0.4, 1)
    b = orc.ks_gaussian(DIMS, 2) if kind == clg.KS else orc.wilson_gaussian(DIMS, 2)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, DIMS)
    g.upload(U)
    fb = clg.Field(ctx, kind, clg.F64, DIMS, bc=orc.BC_FERMION)
    fb.upload(b)
    return U, b, g, fb


def apply_shifted(kind, op, U, x, s, m):
    if kind == clg.KS:
        Ax = orc.ks_op(DIMS, orc.BC_FERMION, m, U, x, op)
    else:
        Ax = orc.wilson_op(DIMS, orc.BC_FERMION, m, U, x, op)
    return orc.as_complex(Ax) + s * orc.as_complex(x)


@pytest.mark.parametrize("solver", [clg.MS_FOM, clg.MS_BICGSTAB, clg.MS_GMRES])
def test_wilson_complex_shifts(ctx, solver):
    # TestMultiShiftSolver shift set (TestMultiShiftSolver.cpp:12-111)
    U, b, g, fb = setup(ctx, clg.WILSON)
    shifts = [2.0, 1.0, 0.5, 1j, 0.5 + 0.5j, 1 - 1j]
    xs = [clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=orc.BC_FERMION) for _ in shifts]
    p = clg.solver_params(kind=clg.WILSON, max_dim=20, restart=50, max_step=200, deviation_step=5,
                          accuracy=1e-11, absolute_accuracy=1)
    st, rep = clg.multishift_solve(ctx, p, xs, shifts, fb, g, 0.1355, clg.EFO_F_D, kind=solver)
    assert st == clg.OK
    for s, x in zip(shifts, xs):
        r = apply_shifted(clg.WILSON, orc.D, U, x.download(), s, 0.1355) - orc.as_complex(b)
        assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-9


@pytest.mark.parametrize("solver", [clg.MS_FOM, clg.MS_BICGSTAB, clg.MS_GMRES, clg.MS_NESTED])
@pytest.mark.parametrize("prec", [clg.F64, clg.F32])
def test_ks_ddagger_remez_shifts_vs_oracle(ctx, solver, prec):
    U, b, g, fb = setup(ctx, clg.KS)
    c = REMEZ["MinusOneOver2"]
    shifts = c[4:]
    xs = [clg.Field(ctx, clg.KS, clg.F64, DIMS, bc=orc.BC_FERMION) for _ in shifts]
    acc = 1e-11 if prec == clg.F64 else 1e-5
    p = clg.solver_params(kind=clg.KS, max_dim=30, restart=50, max_step=400, deviation_step=5, accuracy=acc,
                          absolute_accuracy=1)
    if solver == clg.MS_NESTED:
        # the nested solver: relative true-residual GMRES per shifted mass
        p = clg.solver_params(kind=clg.KS, max_dim=30, restart=50, accuracy=acc * 1e-1,
                              stop_rule=clg.STOP_TRUE_RESIDUAL, inner_accuracy=1e-4)
    st, rep = clg.multishift_solve(ctx, p, xs, shifts, fb, g, 0.1, clg.EFO_F_DDdagger, kind=solver, prec=prec)
    assert st == clg.OK
    xo, rc, _, _ = orc.msfom(1, orc.DDDAG, DIMS, orc.BC_FERMION, 0.1, U, b, shifts, max_dim=30, restart=50,
                             accuracy=1e-12)
    bar = 1e-9 if prec == clg.F64 else 1e-4
    for k, s in enumerate(shifts):
        x = xs[k].download()
        r = apply_shifted(clg.KS, orc.DDDAG, U, x, s, 0.1) - orc.as_complex(b)
        assert np.linalg.norm(r) / np.linalg.norm(b) < bar
        assert np.linalg.norm(x - xo[k]) / np.linalg.norm(xo[k]) < bar * 100


def test_rational_apply_matches_oracle_and_inverts(ctx):
    U, b, g, fb = setup(ctx, clg.KS)
    p = clg.solver_params(kind=clg.KS, max_dim=30, restart=50, accuracy=1e-12, absolute_accuracy=1)
    half = clg.Field(ctx, clg.KS, clg.F64, DIMS, bc=orc.BC_FERMION)
    c = REMEZ["OneOver2"]
    clg.rational_apply(ctx, p, half, fb, g, 0.1, clg.EFO_F_DDdagger, c)
    xo, _, _, _ = orc.msfom(1, orc.DDDAG, DIMS, orc.BC_FERMION, 0.1, U, b, c[4:], max_dim=30, restart=50,
                            accuracy=1e-13)
    want = c[0] * b + sum(a * x for a, x in zip(c[1:4], xo))
    got = half.download()
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-9
    # r_{-1/2} r_{1/2} phi ~ phi wherever the spectrum lies in the approximation range; with
    # the coefficients rescaled to A/16 (a_k -> 16^p-scaled form) the whole DD^dagger spectrum
    # (m^2 .. m^2 + 16) maps into [6e-4, 1.0]
    def scaled(cs, p):
        n = (len(cs) - 1) // 2
        s = 16.0
        # r(A/s) = c + sum a/(A/s + b) = c + sum (a s)/(A + b s); the target is (A/s)^p
        return [cs[0]] + [a * s for a in cs[1:1 + n]] + [bb * s for bb in cs[1 + n:]]
    up, down = scaled(REMEZ["OneOver2"], 0.5), scaled(REMEZ["MinusOneOver2"], -0.5)
    t = clg.Field(ctx, clg.KS, clg.F64, DIMS, bc=orc.BC_FERMION)
    clg.rational_apply(ctx, p, t, fb, g, 0.3, clg.EFO_F_DDdagger, up)
    clg.rational_apply(ctx, p, t, t, g, 0.3, clg.EFO_F_DDdagger, down)
    assert np.linalg.norm(t.download() - b) / np.linalg.norm(b) < 5e-3


@pytest.mark.parametrize("solver", [clg.MS_FOM, clg.MS_BICGSTAB])
def test_ks_fp16_link_multishift(ctx, solver):
    """Shifted solves with FP32 vectors and FP16 12-real links solve the shifted systems of the operator with
    FP16-rounded links: against that operator (links rounded on the host the way the device stores them) the
    residuals are at the FP32 solve's level, as in the plain FP32 case; against the FP64 links they are at
    the links' rounding level."""
    U, b, g, fb = setup(ctx, clg.KS)
    c = REMEZ["MinusOneOver2"]
    shifts = c[4:]
    xs = [clg.Field(ctx, clg.KS, clg.F64, DIMS, bc=orc.BC_FERMION) for _ in shifts]
    p = clg.solver_params(kind=clg.KS, max_dim=30, restart=50, max_step=400, deviation_step=5, accuracy=1e-5,
                          absolute_accuracy=1)
    st, rep = clg.multishift_solve(ctx, p, xs, shifts, fb, g, 0.1, clg.EFO_F_DDdagger, kind=solver, prec=clg.F16)
    assert st == clg.OK
    U16 = orc.fp16_links_12(U)
    worst16 = worst64 = 0.0
    for k, s in enumerate(shifts):
        x = xs[k].download()
        r16 = apply_shifted(clg.KS, orc.DDDAG, U16, x, s, 0.1) - orc.as_complex(b)
        r64 = apply_shifted(clg.KS, orc.DDDAG, U, x, s, 0.1) - orc.as_complex(b)
        worst16 = max(worst16, np.linalg.norm(r16) / np.linalg.norm(b))
        worst64 = max(worst64, np.linalg.norm(r64) / np.linalg.norm(b))
    print(f"FP16-link shifted solutions: max residual {worst16:.2e} with the FP16-rounded links, {worst64:.2e} "
          f"with the FP64 links")
    assert worst16 < 1e-3 and worst64 < 3e-2
