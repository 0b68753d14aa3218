"""CPU pins of the oracle's solvers and of the reference's Remez coefficient sets."""
import json
import os

import numpy as np
import pytest

import orc

HERE = os.path.dirname(os.path.abspath(__file__))
REMEZ = json.load(open(os.path.join(HERE, "golden", "remez_ks.json")))


@pytest.mark.parametrize("name,p", [("OneOver2", 0.5), ("MinusOneOver2", -0.5), ("OneOver4", 0.25),
                                    ("MinusOneOver4", -0.25)])
def test_remez_sets_approximate_powers(name, p):
    c = REMEZ[name]
    n = (len(c) - 1) // 2
    x = np.logspace(np.log10(0.003), 0, 100)
    r = c[0] + sum(a / (x + b) for a, b in zip(c[1:1 + n], c[1 + n:]))
    assert np.abs(r / x ** p - 1).max() < 2e-3


def test_gmres_matches_dense_solve_tiny_lattice():
    dims = (2, 2, 2, 4)
    U = orc.gauge_random(dims, 0.6, 3)
    b = orc.wilson_gaussian(dims, 4)
    N = orc.volume(dims) * 12
    Dm = np.zeros((N, N), complex)
    for j in range(N):
        e = np.zeros(N, complex)
        e[j] = 1
        Dm[:, j] = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, 0.13, U, orc.from_complex(e)))
    xd = np.linalg.solve(Dm, orc.as_complex(b))
    x, rc, _, _ = orc.gmres(0, orc.D, dims, orc.BC_FERMION, 0.13, U, b, max_dim=12, restart=100,
                            accuracy=1e-13, stop_rule=1)
    assert rc == 0
    assert np.abs(orc.as_complex(x) - xd).max() < 1e-11


def test_gmres_reference_rule_returns_best_iterate_on_failure():
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    b = orc.wilson_gaussian(dims, 2)
    x, rc, cyc, dev = orc.gmres(0, orc.D, dims, orc.BC_FERMION, 0.1355, U, b, max_dim=3, restart=2,
                                accuracy=1e-30, blength=1.0, stop_rule=0)
    assert rc == 1 and cyc == 2
    r0 = np.linalg.norm(b - orc.wilson_op(dims, orc.BC_FERMION, 0.1355, U, b))
    r1 = np.linalg.norm(b - orc.wilson_op(dims, orc.BC_FERMION, 0.1355, U, x))
    assert r1 < r0


def test_msfom_shifted_residuals():
    # TestMultiShiftSolver shifts (TestMultiShiftSolver.cpp:12-111) on the Wilson D
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    b = orc.wilson_gaussian(dims, 2)
    shifts = [2.0, 1.0, 0.5, 1j, 0.5 + 0.5j, 1 - 1j]
    xs, rc, cyc, err = orc.msfom(0, orc.D, dims, orc.BC_FERMION, 0.1355, U, b, shifts, max_dim=20,
                                 restart=50, accuracy=1e-11)
    assert rc == 0
    for s, x in zip(shifts, xs):
        Ax = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, 0.1355, U, x)) + s * orc.as_complex(x)
        assert np.linalg.norm(Ax - orc.as_complex(b)) / np.linalg.norm(b) < 1e-9


def test_bicgstab_solution():
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    b = orc.wilson_gaussian(dims, 2)
    x, rc, it, dev = orc.bicgstab(0, orc.D, dims, orc.BC_FERMION, 0.1355, U, b, step_count=100,
                                  dev_check=5, retry=3, accuracy=1e-16)
    assert rc == 0
    r = b - orc.wilson_op(dims, orc.BC_FERMION, 0.1355, U, x)
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-7


def _dense_wilson(dims, kappa, U):
    n = orc.volume(dims) * 24
    A = np.zeros((n // 2, n // 2), complex)
    for j in range(n // 2):
        e = np.zeros(n)
        e[2 * j] = 1.0
        A[:, j] = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, kappa, U, e, orc.D))
    return A


def test_gcrodr_matches_dense_solve():
    dims = (2, 2, 2, 4)
    U = orc.gauge_random(dims, 0.5, 3)
    b = orc.wilson_gaussian(dims, 4)
    A = _dense_wilson(dims, 0.12, U)
    want = np.linalg.solve(A, orc.as_complex(b))
    x, rc, cyc, dev, _, _ = orc.gcrodr(0, orc.D, dims, orc.BC_FERMION, 0.12, U, b, m=10, k=4, restart=200,
                                       accuracy=1e-24)
    assert rc == 0
    assert np.abs(orc.as_complex(x) - want).max() < 1e-9 * np.abs(want).max()
    assert dev < 1e-24 * np.dot(b, b)


def test_gcrodr_recycled_solve_agrees_with_fresh_solve():
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    kappa = 0.15
    b1, b2 = orc.wilson_gaussian(dims, 5), orc.wilson_gaussian(dims, 6)
    kw = dict(m=12, k=6, restart=300, accuracy=1e-22)
    x1, rc1, c1, _, urec, hu = orc.gcrodr(0, orc.D, dims, orc.BC_FERMION, kappa, U, b1, phase=0, **kw)
    assert rc1 == 0 and hu == 1
    x2, rc2, c2, _, _, _ = orc.gcrodr(0, orc.D, dims, orc.BC_FERMION, kappa, U, b2, phase=1, urec=urec, have_u=1, **kw)
    x3, rc3, c3, _, _, _ = orc.gcrodr(0, orc.D, dims, orc.BC_FERMION, kappa, U, b2, phase=3, **kw)
    assert rc2 == 0 and rc3 == 0
    r = orc.wilson_op(dims, orc.BC_FERMION, kappa, U, x2, orc.D) - b2
    assert np.linalg.norm(r) < 1e-10 * np.linalg.norm(b2)
    assert np.abs(x2 - x3).max() < 1e-8 * np.abs(x3).max()
