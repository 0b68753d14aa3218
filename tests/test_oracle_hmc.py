"""Pins of the oracle's HMC pieces: the two SU(3) exponentials against closed forms and
library routines, and the integrators against reversibility and their order of accuracy."""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import orc

DIMS = (4, 4, 4, 4)


def rand_ah(rng):
    a = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
    m = 0.5 * (a - a.conj().T)
    return m - np.trace(m) / 3 * np.eye(3)


def to_arr(m):
    return orc.from_complex(np.ascontiguousarray(m).ravel())


def to_mat(a):
    return orc.as_complex(a).reshape(3, 3)


def cayley(i, j, y, z):
    x = np.zeros((3, 3), complex)
    x[i, i], x[j, j], x[i, j], x[j, i] = 1j * y, -1j * y, z, -np.conj(z)
    e = np.eye(3)
    return (e + x) @ np.linalg.inv(e - x)


def test_taylor_exp_matches_expm():
    rng = np.random.default_rng(1)
    for _ in range(5):
        m = rand_ah(rng)
        for n in (1, 2, 8):
            ref = sum(np.linalg.matrix_power(0.3 * m, k) / math.factorial(k) for k in range(n + 1))
            assert np.abs(to_mat(orc.su3_exp(to_arr(m), 0.3, n)) - ref).max() < 1e-14
        assert np.abs(to_mat(orc.su3_exp(to_arr(m), 0.3, 30)) - expm(0.3 * m)).max() < 1e-14


def test_quick_exp_is_symmetric_cayley_product():
    rng = np.random.default_rng(2)
    for _ in range(5):
        m = rand_ah(rng)
        a = 0.37
        q = to_mat(orc.su3_exp(to_arr(m), a, 0))
        d = np.imag(np.diag(m))
        c01 = cayley(0, 1, a * (d[0] - d[1]) / 12, a * m[0, 1] / 4)
        c02 = cayley(0, 2, a * (d[0] - d[2]) / 12, a * m[0, 2] / 4)
        c12 = cayley(1, 2, a * (d[1] - d[2]) / 6, a * m[1, 2] / 2)
        assert np.abs(q - c01 @ c02 @ c12 @ c02 @ c01).max() < 1e-14
        assert np.abs(q @ q.conj().T - np.eye(3)).max() < 1e-14
        assert abs(np.linalg.det(q) - 1) < 1e-14
        qm = to_mat(orc.su3_exp(to_arr(m), -a, 0))
        assert np.abs(q @ qm - np.eye(3)).max() < 1e-14


def test_quick_exp_error_is_third_order():
    rng = np.random.default_rng(3)
    m = rand_ah(rng)
    errs = [np.abs(to_mat(orc.su3_exp(to_arr(m), a, 0)) - expm(a * m)).max() for a in (0.02, 0.01)]
    assert 6.0 < errs[0] / errs[1] < 10.0


def _setup(seed=5, wilson=True, ks=False):
    rng = np.random.default_rng(seed)
    U = orc.gauge_random(DIMS, 0.3, 1)
    P = orc.random_momentum(rng, orc.volume(DIMS) * 4)
    acts = [dict(kind=0, beta=5.5)]
    if wilson:
        acts.append(dict(kind=1, kappa_or_mass=0.12, bc=orc.BC_FERMION, phi=orc.wilson_gaussian(DIMS, 7)))
    if ks:
        md = [0.3, 0.2, 0.4, 0.05, 0.5]
        acts.append(dict(kind=2, kappa_or_mass=0.3, bc=orc.BC_FERMION, phi=orc.ks_gaussian(DIMS, 8), md=md))
    return U, P, acts


def test_momentum_layout_is_traceless_antihermitian():
    rng = np.random.default_rng(4)
    P = orc.random_momentum(rng, 10)
    for l in range(10):
        m = orc.as_complex(P[18 * l:18 * l + 18]).reshape(3, 3)
        assert np.abs(m + m.conj().T).max() < 1e-15
        assert abs(np.trace(m)) < 1e-15


@pytest.mark.parametrize("integrator,order", [(orc.INT_NESTED_FG, 4), (orc.INT_LEAPFROG, 2)])
def test_integrator_order(integrator, order):
    U, P, acts = _setup()
    dh = []
    for steps in (4, 8):
        _, _, h0, h1 = orc.hmc_evolve(DIMS, U, P, acts, integrator, tau=0.5, steps=steps, nested=1,
                                      exp_prec=24)
        dh.append(abs(h1 - h0))
    ratio = dh[0] / dh[1]
    assert 0.6 * 2 ** order < ratio < 1.6 * 2 ** order, (dh, ratio)


def test_nested_force_gradient_reversible():
    U, P, acts = _setup(ks=True)
    U1, P1, h0, h1 = orc.hmc_evolve(DIMS, U, P, acts, orc.INT_NESTED_FG, tau=0.4, steps=2, nested=2, exp_prec=0)
    U2, P2, h2, h3 = orc.hmc_evolve(DIMS, U1, -P1, acts, orc.INT_NESTED_FG, tau=0.4, steps=2, nested=2,
                                    exp_prec=0)
    assert np.abs(U2 - U).max() < 1e-9
    assert np.abs(P2 + P).max() < 1e-9
    assert abs(h2 - h1) < 1e-9 * abs(h1) and abs(h3 - h0) < 1e-8 * abs(h0)
    assert abs(h1 - h0) < 1e-2 * abs(h0)
