"""Finite-difference pins of the oracle's fermion force kernels.

With the reference's HMC conventions (traceless anti-Hermitian momenta, kinetic energy
sum ReTr P^dagger P, P <- P + eps F, U <- exp(eps P) U), a force F belongs to an action S when
dS/deps = -2 Re sum_links Tr(F^dagger T) along U -> exp(eps T) U for any traceless anti-Hermitian
T.  A dropped term, a wrong sign or projector, or a misplaced link fails this check.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg

import orc

HERE = os.path.dirname(os.path.abspath(__file__))
REMEZ = json.load(open(os.path.join(HERE, "golden", "remez_ks.json")))
DIMS = (4, 4, 4, 4)


def random_algebra(rng, nlink):
    h = rng.normal(size=(nlink, 3, 3)) + 1j * rng.normal(size=(nlink, 3, 3))
    h = 0.5 * (h - np.conj(np.transpose(h, (0, 2, 1))))          # anti-Hermitian
    tr = np.trace(h, axis1=1, axis2=2) / 3.0
    return h - tr[:, None, None] * np.eye(3)[None]


def move(U, T, eps):
    Uc = orc.as_complex(U).reshape(-1, 3, 3)
    out = np.stack([scipy.linalg.expm(eps * T[k]) @ Uc[k] for k in range(Uc.shape[0])])
    return orc.from_complex(out.ravel())


def richardson(S, U, T):
    """dS/deps at 0 from central differences at h and h/2, extrapolated (error O(h^4))"""
    def cd(h):
        return (S(move(U, T, h)) - S(move(U, T, -h))) / (2 * h)
    h = 2e-4
    return (4 * cd(h / 2) - cd(h)) / 3


def predicted(F, T):
    Fc = orc.as_complex(F).reshape(-1, 3, 3)
    return -2.0 * np.real(np.einsum("kij,kij->", np.conj(Fc), T))


def wilson_action(U, phi, kappa):
    x, rc, _, _ = orc.gmres(0, orc.D, DIMS, orc.BC_FERMION, kappa, U, phi, max_dim=30, restart=300,
                            accuracy=1e-14, stop_rule=1)
    assert rc == 0
    return float(np.dot(x, x))


def test_wilson_force_finite_difference():
    rng = np.random.default_rng(11)
    kappa = 0.12
    U = orc.gauge_random(DIMS, 0.4, 1)
    phi = orc.wilson_gaussian(DIMS, 2)
    X, rc, _, _ = orc.gmres(0, orc.DDDAG, DIMS, orc.BC_FERMION, kappa, U, phi, max_dim=30, restart=300,
                            accuracy=1e-14, stop_rule=1)
    assert rc == 0
    Y = orc.wilson_op(DIMS, orc.BC_FERMION, kappa, U, X, orc.DDAG)
    F = orc.wilson_force(DIMS, orc.BC_FERMION, kappa, U, Y, X)
    T = random_algebra(rng, orc.volume(DIMS) * 4)
    fd = richardson(lambda u: wilson_action(u, phi, kappa), U, T)
    pred = predicted(F, T)
    assert abs(fd - pred) < 1e-8 * abs(pred), (fd, pred)


def ks_action(U, phi, two_am, coeffs):
    n = (len(coeffs) - 1) // 2
    xs, rc, _, _ = orc.msfom(1, orc.DDDAG, DIMS, orc.BC_FERMION, two_am, U, phi, coeffs[1 + n:], max_dim=30,
                             restart=100, accuracy=1e-14)
    assert rc == 0
    val = coeffs[0] * np.dot(phi, phi)
    for a, x in zip(coeffs[1:1 + n], xs):
        val += a * np.real(orc.dot(phi, x))
    return float(val)


def test_ks_force_finite_difference():
    rng = np.random.default_rng(12)
    two_am = 0.2
    coeffs = REMEZ["MinusOneOver2"]
    n = (len(coeffs) - 1) // 2
    U = orc.gauge_random(DIMS, 0.4, 1)
    phi = orc.ks_gaussian(DIMS, 2)
    xs, rc, _, _ = orc.msfom(1, orc.DDDAG, DIMS, orc.BC_FERMION, two_am, U, phi, coeffs[1 + n:], max_dim=30,
                             restart=100, accuracy=1e-14)
    assert rc == 0
    ds = np.stack([orc.ks_op(DIMS, orc.BC_FERMION, 0.0, U, x, orc.D) for x in xs])
    F = orc.ks_force(DIMS, orc.BC_FERMION, U, xs, ds, coeffs[1:1 + n])
    T = random_algebra(rng, orc.volume(DIMS) * 4)
    fd = richardson(lambda u: ks_action(u, phi, two_am, coeffs), U, T)
    pred = predicted(F, T)
    assert abs(fd - pred) < 1e-8 * abs(pred), (fd, pred)


def test_gauge_force_finite_difference():
    rng = np.random.default_rng(13)
    beta = 5.5
    U = orc.gauge_random(DIMS, 0.4, 1)
    F = orc.gauge_force(DIMS, U, beta)
    T = random_algebra(rng, orc.volume(DIMS) * 4)
    fd = richardson(lambda u: beta / 3.0 * orc.plaquette_sum(DIMS, u), U, T)
    pred = predicted(F, T)
    assert abs(fd - pred) < 1e-8 * abs(pred), (fd, pred)


def test_plaquette_of_unit_and_random_fields():
    assert orc.plaquette_sum(DIMS, orc.gauge_unit(DIMS)) == 0.0
    U = orc.gauge_random(DIMS, 0.4, 1)
    s = orc.plaquette_sum(DIMS, U)
    assert 0 < s < 6 * orc.volume(DIMS) * 4.5
