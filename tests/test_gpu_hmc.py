"""GPU gauge-field kernels and the HMC integrator/trajectory vs the CPU oracle."""
import json
import os

import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
REMEZ = json.load(open(os.path.join(HERE, "golden", "remez_ks.json")))
DIMS = (4, 4, 4, 4)
V = orc.volume(DIMS)


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


def up(ctx, kind, data, dims=DIMS):
    f = clg.Field(ctx, kind, clg.F64, dims, bc=orc.BC_FERMION)
    f.upload(data)
    return f


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def momentum(seed, nlinks=V * 4):
    return orc.random_momentum(np.random.default_rng(seed), nlinks)


def test_gauge_force_and_plaquette_parity(ctx):
    U = orc.gauge_random(DIMS, 0.4, 1)
    F0 = momentum(3)  # accumulate into a non-zero force
    g, F = up(ctx, clg.GAUGE, U), up(ctx, clg.GAUGE, F0)
    clg.gauge_force(ctx, F, g, 5.5)
    assert rel(F.download(), orc.gauge_force(DIMS, U, 5.5, F=F0)) < 1e-13
    e = clg.plaquette_energy(ctx, g, 5.5)
    assert abs(e - 5.5 / 3 * orc.plaquette_sum(DIMS, U)) < 1e-12 * e


@pytest.mark.parametrize("prec", [0, 1, 8, 20])
def test_exp_mult_parity(ctx, prec):
    U = orc.gauge_random(DIMS, 0.4, 1)
    P = momentum(5)
    g, fp = up(ctx, clg.GAUGE, U), up(ctx, clg.GAUGE, P)
    clg.exp_mult(ctx, fp, g, 0.3, prec)
    got = g.download()
    want = np.empty_like(U)
    for l in range(V * 4):
        e = orc.as_complex(orc.su3_exp(P[18 * l:18 * l + 18], 0.3, prec)).reshape(3, 3)
        u = orc.as_complex(U[18 * l:18 * l + 18]).reshape(3, 3)
        want[18 * l:18 * l + 18] = orc.from_complex((e @ u).ravel())
    assert np.abs(got - want).max() < 1e-14


def test_normalize_parity(ctx):
    U = orc.gauge_random(DIMS, 0.4, 1) * 1.01 + 0.003
    g = up(ctx, clg.GAUGE, U)
    clg.gauge_normalize(ctx, g)
    assert np.abs(g.download() - orc.su3_norm(U)).max() < 1e-14


def test_momentum_heatbath_statistics(ctx):
    dims = (8, 8, 8, 8)
    vol = orc.volume(dims)
    P = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    clg.momentum_random(P, 11, 0)
    p = orc.as_complex(P.download()).reshape(vol * 4, 3, 3)
    assert np.abs(p + np.conj(np.transpose(p, (0, 2, 1)))).max() < 1e-15
    assert np.abs(np.trace(p, axis1=1, axis2=2)).max() < 1e-14
    kin = clg.kinetic_energy(ctx, P)
    assert abs(kin - np.sum(np.abs(p) ** 2)) < 1e-12 * kin
    # ReTr P^dag P = 2 sum_a r_a^2 per link with r_a ~ N(0, 1/4): each of the n = 32 V terms
    # 2 r_a^2 has mean 1/2 and variance 1/2
    n = vol * 4 * 8
    assert abs(kin - 0.5 * n) < 6 * np.sqrt(0.5 * n)
    # a different counter gives an independent draw, the same (seed, counter) the same draw
    P2 = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    clg.momentum_random(P2, 11, 1)
    assert np.abs(P2.download() - P.download()).max() > 0.1
    clg.momentum_random(P2, 11, 0)
    assert np.array_equal(P2.download(), P.download())


def test_gauge_axpy_and_kinetic(ctx):
    P, F = momentum(1), momentum(2)
    fp, ff = up(ctx, clg.GAUGE, P), up(ctx, clg.GAUGE, F)
    clg.gauge_axpy(ctx, fp, -0.37, ff)
    want = P - 0.37 * F
    assert np.abs(fp.download() - want).max() < 1e-15
    assert abs(clg.kinetic_energy(ctx, fp) - np.dot(want, want)) < 1e-13 * np.dot(want, want)


def test_wilson_heatbath_is_d_eta(ctx):
    U = orc.gauge_random(DIMS, 0.4, 1)
    g = up(ctx, clg.GAUGE, U)
    phi = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=orc.BC_FERMION)
    eta = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=orc.BC_FERMION)
    clg.wilson_prepare_hmc(ctx, phi, g, 0.12, 77)
    eta.gaussian(77)
    want = orc.wilson_op(DIMS, orc.BC_FERMION, 0.12, U, eta.download(), orc.D)
    assert rel(phi.download(), want) < 1e-14


def _wilson_params():
    return clg.solver_params(max_dim=20, restart=300, accuracy=1e-13, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)


def _ks_params(phi):
    return clg.solver_params(kind=clg.KS, max_dim=40, restart=200, accuracy=1e-13 * np.linalg.norm(phi),
                             absolute_accuracy=1)


def _actions(ctx, which):
    """(gpu actions, oracle actions) on DIMS for which in {'wilson', 'ks'}"""
    acts_gpu = [dict(kind=clg.ACTION_GAUGE_PLAQUETTE, beta=5.5)]
    acts_orc = [dict(kind=0, beta=5.5)]
    if which == "wilson":
        phi = orc.wilson_gaussian(DIMS, 7) * 0.8
        acts_gpu.append(dict(kind=clg.ACTION_FERMION_WILSON_NF2, kappa_or_mass=0.12, phi=up(ctx, clg.WILSON, phi),
                             solver=_wilson_params()))
        acts_orc.append(dict(kind=1, kappa_or_mass=0.12, phi=phi, accuracy=1e-13))
    else:
        md = REMEZ["MinusOneOver2"]
        phi = orc.ks_gaussian(DIMS, 8)
        acts_gpu.append(dict(kind=clg.ACTION_FERMION_KS, kappa_or_mass=0.3, phi=up(ctx, clg.KS, phi), md=md,
                             solver=_ks_params(phi)))
        acts_orc.append(dict(kind=2, kappa_or_mass=0.3, phi=phi, md=md, accuracy=1e-14))
    return acts_gpu, acts_orc


@pytest.mark.parametrize("which,integrator,exp_prec", [
    ("wilson", clg.INT_NESTED_FORCE_GRADIENT, 0),
    ("wilson", clg.INT_LEAPFROG, 8),
    ("ks", clg.INT_NESTED_FORCE_GRADIENT, 8),
])
def test_hmc_evolve_parity(ctx, which, integrator, exp_prec):
    U = orc.gauge_random(DIMS, 0.3, 1)
    P = momentum(21)
    acts_gpu, acts_orc = _actions(ctx, which)
    g, fp = up(ctx, clg.GAUGE, U), up(ctx, clg.GAUGE, P)
    hp = clg.hmc_params(integrator=integrator, step_length=0.5, steps=2, nested_steps=2, exp_precision=exp_prec)
    rep = clg.hmc_evolve(ctx, g, fp, hp, acts_gpu)
    oint = orc.INT_NESTED_FG if integrator == clg.INT_NESTED_FORCE_GRADIENT else orc.INT_LEAPFROG
    U1, P1, h0, h1 = orc.hmc_evolve(DIMS, U, P, acts_orc, oint, tau=0.5, steps=2, nested=2, exp_prec=exp_prec)
    print(f"{which}: H {rep.h_before:.12f} -> {rep.h_after:.12f}, oracle {h0:.12f} -> {h1:.12f}; "
          f"{rep.force_evaluations} forces, {rep.solver_iterations} solver cycles")
    assert rep.unconverged_solves == 0
    assert np.abs(g.download() - U1).max() < 1e-9
    assert np.abs(fp.download() - P1).max() < 1e-9
    assert abs(rep.h_before - h0) < 1e-10 * abs(h0)
    assert abs(rep.h_after - h1) < 1e-10 * abs(h1)
    assert abs(rep.h_diff - (h0 - h1)) < 1e-7


def test_hmc_evolve_reversible(ctx):
    U = orc.gauge_random(DIMS, 0.3, 1)
    P = momentum(22)
    acts_gpu, _ = _actions(ctx, "wilson")
    g, fp = up(ctx, clg.GAUGE, U), up(ctx, clg.GAUGE, P)
    hp = clg.hmc_params(step_length=0.5, steps=2, nested_steps=2)
    r1 = clg.hmc_evolve(ctx, g, fp, hp, acts_gpu)
    fp.upload(-fp.download())
    r2 = clg.hmc_evolve(ctx, g, fp, hp, acts_gpu)
    assert np.abs(g.download() - U).max() < 1e-9
    assert np.abs(fp.download() + P).max() < 1e-9
    assert abs(r1.h_diff + r2.h_diff) < 1e-8


def test_hmc_trajectory_metropolis(ctx):
    U = orc.gauge_random(DIMS, 0.3, 1)
    acts_gpu, _ = _actions(ctx, "wilson")
    hp = clg.hmc_params(step_length=0.5, steps=3, nested_steps=2)
    reps = []
    for seed in (1, 1, 2):
        g = up(ctx, clg.GAUGE, U)
        rep = clg.hmc_trajectory(ctx, g, hp, acts_gpu, seed)
        reps.append(rep)
        after = g.download()
        if rep.accepted:
            m = orc.as_complex(after).reshape(-1, 3, 3)
            assert np.abs(m @ np.conj(np.transpose(m, (0, 2, 1))) - np.eye(3)).max() < 1e-13
            assert np.abs(after - U).max() > 1e-3
        else:
            assert np.array_equal(after, U)
        assert abs(rep.h_diff) < 5.0
        assert rep.kinetic_before > 0 and rep.force_evaluations > 0 and rep.unconverged_solves == 0
    assert reps[0].h_diff == reps[1].h_diff and reps[0].h_before == reps[1].h_before
    assert reps[0].h_before != reps[2].h_before
    # the fermion energy right after the heat-bath is |eta|^2, E = 12 V
    assert 0.8 * 12 * V < reps[0].action_before[1] < 1.2 * 12 * V
