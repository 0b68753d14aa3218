"""The benchmark's solver workloads at their full stated sizes, under assertions.

- 32^3x32, kappa 0.137, eps 0.5 gauge (seed 1), Gaussian source (seed 2): EO-GMRES(12) with FP32,
  FP16-link and FP16-vector inner solves inside FP64 defect correction to |r|/|b| < 1e-10; the
  residual is recomputed by the oracle's full-lattice D on the downloaded solution.
- 12 point sources on 24^3x48 plus one Gaussian source each on 8^4, 12^4, 16^3x32 and 20^3x40
  (kappa 0.1355, eps 0.5): one batched solve with FP32 inner / FP64 outer and the same 16 systems
  solved one at a time in plain FP64, both to 1e-14; every batched solution equals its single
  solution to 1e-12 (|D^-1| ~ 10 turns the two 1e-14 residuals into at most ~2e-13).
- Staggered 2+1 RHMC at 24^3x8 (masses 0.01 and 0.05, 12-pole rational functions on shifts
  10^(-4 + 5i/11), nested force gradient 8 x 6, beta 5.5, eps 0.4 gauge): one trajectory with FP32
  multi-shift force solves and one in plain FP64 from the same start, momenta and pseudofermions.
  Both trajectories integrate the same Hamiltonian; the FP32 forces differ from the FP64 ones by the
  FP32 solves' 1e-5 accuracy, which moves Delta H by that relative amount times the O(1) energy
  changes along the trajectory, so |dH_32 - dH_64| < 2e-3 (ten times the difference measured when
  the bound was set) with |dH| itself O(0.1-1).
"""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

BC = orc.BC_FERMION


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def c2(ctx):
    dims, kappa = (32, 32, 32, 32), 0.137
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.5, 1)
    b = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
    b.gaussian(2)
    return dims, kappa, g, b, g.download(), b.download()


@pytest.fixture(scope="module")
def c2_oracle(c2):
    """The oracle's FP64 solution of the c2 system (full-lattice GMRES(12) to a true residual below 1e-12 on
    the host cores, about 20 s)."""
    dims, kappa, g, b, U, bh = c2
    xo, rc, cyc, res = orc.gmres(0, orc.D, dims, BC, kappa, U, bh, max_dim=12, restart=400, accuracy=1e-12,
                                 stop_rule=1)
    assert rc == 0 and res < 1e-12
    return xo


@pytest.mark.parametrize("inner,inner_acc", [(clg.F32, 1e-4), (clg.F16, 1e-3), (clg.INNER_F16_VECTORS, 1e-2)])
def test_mixed_solves_at_32x32x32x32_kappa0137(ctx, c2, inner, inner_acc):
    dims, kappa, g, b, U, bh = c2
    x = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
    p = clg.solver_params(max_dim=12, restart=2000, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                          inner_prec=inner, inner_accuracy=inner_acc, max_outer=80)
    st, rep = clg.solve(ctx, p, x, b, g, kappa, clg.EFO_F_InverseD)
    r = bh - orc.wilson_op(dims, BC, kappa, U, x.download())
    res = np.linalg.norm(r) / np.linalg.norm(bh)
    print(f"inner {inner}: {rep.seconds * 1e3:.1f} ms, outer {rep.outer}, cycles {rep.cycles}, "
          f"reported {rep.true_residual:.2e}, oracle D residual {res:.2e}")
    assert st == clg.OK
    assert res < 1e-10
    assert abs(res - rep.true_residual) < 1e-2 * res


def test_c2_solutions_against_the_oracle_solution(ctx, c2, c2_oracle):
    """c2's error against the FP64 oracle solution: the plain FP64 solve and the FP16-vector mixed solve (the
    fastest, lowest-precision inner form) to |r|/|b| < 1e-10 are both within 1e-9 of it (|D^-1| at kappa
    0.137 on this field keeps the error within a few times 1e-10)."""
    dims, kappa, g, b, U, bh = c2
    for inner, inner_acc in ((clg.F64, 0.0), (clg.INNER_F16_VECTORS, 1e-2)):
        x = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
        p = clg.solver_params(max_dim=12, restart=2000, accuracy=1e-10, stop_rule=clg.STOP_TRUE_RESIDUAL,
                              even_odd=1, inner_prec=inner, inner_accuracy=inner_acc, max_outer=80)
        st, rep = clg.solve(ctx, p, x, b, g, kappa, clg.EFO_F_InverseD)
        err = np.linalg.norm(x.download() - c2_oracle) / np.linalg.norm(c2_oracle)
        print(f"inner {inner}: |x - x_oracle| / |x_oracle| = {err:.2e} (true residual {rep.true_residual:.2e})")
        assert st == clg.OK and err < 1e-9


LATTICES = [((24, 24, 24, 48), 12), ((8, 8, 8, 8), 1), ((12, 12, 12, 12), 1), ((16, 16, 16, 32), 1),
            ((20, 20, 20, 40), 1)]


def test_batched_solve_full_config(ctx):
    kappa = 0.1355
    xs, bs, gl = [], [], []
    for dims, nrhs in LATTICES:
        g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g.gauge_random(0.5, 1)
        for r in range(nrhs):
            b = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC)
            if nrhs > 1:
                b.point_source((0, 0, 0, 0), r // 3, r % 3)
            else:
                b.gaussian(2)
            bs.append(b)
            xs.append(clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=BC))
            gl.append(g)
    assert len(xs) == 16
    p = clg.solver_params(max_dim=12, restart=500, accuracy=1e-14, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1,
                          inner_prec=clg.F32, inner_accuracy=1e-4, max_outer=100)
    st, reps = clg.solve_batch(ctx, p, xs, bs, gl, kappa)
    print("batched residuals", [f"{r.true_residual:.1e}" for r in reps])
    assert st == clg.OK and max(r.true_residual for r in reps) < 1e-14
    p64 = clg.solver_params(max_dim=12, restart=2000, accuracy=1e-14, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
    worst = 0.0
    for i in range(len(xs)):
        x1 = clg.Field(ctx, clg.WILSON, clg.F64, xs[i].dims, bc=BC)
        st1, rep1 = clg.solve(ctx, p64, x1, bs[i], gl[i], kappa, clg.EFO_F_InverseD)
        assert st1 == clg.OK and rep1.true_residual < 1e-14, i
        a, ref = xs[i].download(), x1.download()
        err = np.linalg.norm(a - ref) / np.linalg.norm(ref)
        worst = max(worst, err)
        assert err < 1e-12, (i, err)
    print(f"max |x_batch - x_single| / |x_single| = {worst:.2e}")


def fit_rational(power, shifts, lo, hi):
    """[c, a_1..a_n, b_1..b_n] with the shifts fixed: least-squares fit of x^power on [lo, hi] in relative
    error (the benchmark's synthetic RHMC coefficients)."""
    x = np.logspace(np.log10(lo), np.log10(hi), 600)
    A = np.concatenate([np.ones((x.size, 1)), 1.0 / (x[:, None] + shifts[None, :])], axis=1)
    w = x ** (-power)
    coef, *_ = np.linalg.lstsq(A * w[:, None], np.ones_like(x), rcond=None)
    return np.concatenate([coef, shifts])


@pytest.mark.parametrize("ms_kind", [clg.MS_BICGSTAB, clg.MS_FOM])
def test_rhmc_trajectory_fp32_forces_against_fp64(ctx, ms_kind):
    dims = (24, 24, 24, 8)
    shifts = np.logspace(-4, 1, 12)
    flavours = [(mass, fit_rational(-nf / 4.0, shifts, mass * mass, 20.0), fit_rational(nf / 8.0, shifts, mass * mass,
                                                                                          20.0))
                for mass, nf in ((0.01, 2), (0.05, 1))]
    dh = {}
    for tag, prec in (("fp32", clg.F32), ("fp64", clg.F64)):
        g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g.gauge_random(0.4, 1)
        acts = [dict(kind=clg.ACTION_GAUGE_PLAQUETTE, beta=5.5)]
        phis = []
        for mass, md, mc in flavours:
            phi = clg.Field(ctx, clg.KS, clg.F64, dims, bc=BC)
            phis.append(phi)
            sp = clg.solver_params(kind=clg.KS, max_dim=40, restart=200, absolute_accuracy=1,
                                   accuracy=1e-5 if prec == clg.F32 else 1e-8)
            acts.append(dict(kind=clg.ACTION_FERMION_KS, kappa_or_mass=mass, phi=phi, md=md, mc=mc, solver=sp,
                             ms_kind=ms_kind, ms_prec=prec, energy_accuracy=1e-9))
        hp = clg.hmc_params(integrator=clg.INT_NESTED_FORCE_GRADIENT, step_length=1.0, steps=8, nested_steps=6,
                            exp_precision=8, metropolis=1)
        rep = clg.hmc_trajectory(ctx, g, hp, acts, 1)
        print(f"{tag}: dH {rep.h_diff:.6f} (H {rep.h_before:.4f} -> {rep.h_after:.4f}), {rep.seconds:.2f} s, "
              f"unconverged {rep.unconverged_solves}")
        assert rep.unconverged_solves == 0
        assert np.isfinite(rep.h_diff) and abs(rep.h_diff) < 5.0
        dh[tag] = (rep.h_diff, rep.h_before)
    # same start, momenta and pseudofermions: H before agrees to the energy solves' accuracy
    assert abs(dh["fp32"][1] - dh["fp64"][1]) < 1e-6 * abs(dh["fp64"][1])
    assert abs(dh["fp32"][0] - dh["fp64"][0]) < 2e-3


def test_rhmc_fp16_link_forces(ctx):
    """The c4 trajectory with force solves on FP32 vectors and FP16 12-real links: the heat bath keeps
    full-precision links, so H before equals that of the FP32-force trajectory from the same start; the
    forces come from an operator rounded at 2^-11, which moves the end point (Delta H) more than FP32 solves
    do, but the energies are FP64 and the trajectory stays reversible in form."""
    dims = (24, 24, 24, 8)
    shifts = np.logspace(-4, 1, 12)
    flavours = [(mass, fit_rational(-nf / 4.0, shifts, mass * mass, 20.0), fit_rational(nf / 8.0, shifts, mass * mass,
                                                                                          20.0))
                for mass, nf in ((0.01, 2), (0.05, 1))]
    res = {}
    for prec in (clg.F32, clg.F16):
        g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
        g.gauge_random(0.4, 1)
        acts = [dict(kind=clg.ACTION_GAUGE_PLAQUETTE, beta=5.5)]
        phis = []
        for mass, md, mc in flavours:
            phi = clg.Field(ctx, clg.KS, clg.F64, dims, bc=BC)
            phis.append(phi)
            sp = clg.solver_params(kind=clg.KS, max_dim=40, restart=200, absolute_accuracy=1, accuracy=1e-5)
            acts.append(dict(kind=clg.ACTION_FERMION_KS, kappa_or_mass=mass, phi=phi, md=md, mc=mc, solver=sp,
                             ms_kind=clg.MS_BICGSTAB, ms_prec=prec, energy_accuracy=1e-9))
        hp = clg.hmc_params(integrator=clg.INT_NESTED_FORCE_GRADIENT, step_length=1.0, steps=8, nested_steps=6,
                            exp_precision=8, metropolis=1)
        rep = clg.hmc_trajectory(ctx, g, hp, acts, 1)
        print(f"prec {prec}: dH {rep.h_diff:.6f} (H {rep.h_before:.4f} -> {rep.h_after:.4f}), {rep.seconds:.2f} s")
        assert rep.unconverged_solves == 0 and np.isfinite(rep.h_diff)
        res[prec] = rep
    assert abs(res[clg.F16].h_before - res[clg.F32].h_before) <= 1e-12 * abs(res[clg.F32].h_before)
    assert abs(res[clg.F16].h_diff) < 1.0
