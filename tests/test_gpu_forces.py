"""GPU fermion force kernels, force drivers and pseudofermion energies vs the CPU oracle."""
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


def up(ctx, kind, data):
    f = clg.Field(ctx, kind, clg.F64, DIMS, bc=orc.BC_FERMION)
    f.upload(data)
    return f


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_wilson_force_kernel_parity(ctx):
    U = orc.gauge_random(DIMS, 0.4, 1)
    X = orc.wilson_gaussian(DIMS, 3)
    Y = orc.wilson_gaussian(DIMS, 4)
    F0 = orc.from_complex(orc.as_complex(orc.gauge_random(DIMS, 0.2, 9)) * 0.1)  # accumulate into a non-zero F
    g, fx, fy = up(ctx, clg.GAUGE, U), up(ctx, clg.WILSON, X), up(ctx, clg.WILSON, Y)
    F = up(ctx, clg.GAUGE, F0)
    clg.wilson_force_kernel(ctx, F, g, fx, fy, 0.1355)
    want = orc.wilson_force(DIMS, orc.BC_FERMION, 0.1355, U, Y, X, F=F0)
    assert rel(F.download(), want) < 1e-12


def test_ks_force_kernel_parity(ctx):
    U = orc.gauge_random(DIMS, 0.4, 1)
    phis = [orc.ks_gaussian(DIMS, 10 + k) for k in range(3)]
    ds = [orc.ks_gaussian(DIMS, 20 + k) for k in range(3)]
    num = [0.3, -0.7, 1.1]
    g = up(ctx, clg.GAUGE, U)
    F = up(ctx, clg.GAUGE, np.zeros(orc.volume(DIMS) * 72))
    clg.ks_force_kernel(ctx, F, g, [up(ctx, clg.KS, p) for p in phis], [up(ctx, clg.KS, d) for d in ds], num)
    want = orc.ks_force(DIMS, orc.BC_FERMION, U, np.stack(phis), np.stack(ds), num)
    assert rel(F.download(), want) < 1e-12


def test_wilson_force_driver_and_energy(ctx):
    kappa = 0.12
    U = orc.gauge_random(DIMS, 0.4, 1)
    phi = orc.wilson_gaussian(DIMS, 2)
    X, rc, _, _ = orc.gmres(0, orc.DDDAG, DIMS, orc.BC_FERMION, kappa, U, phi, max_dim=30, restart=300,
                            accuracy=1e-14, stop_rule=1)
    Y = orc.wilson_op(DIMS, orc.BC_FERMION, kappa, U, X, orc.DDAG)
    want = orc.wilson_force(DIMS, orc.BC_FERMION, kappa, U, Y, X)
    g, fphi = up(ctx, clg.GAUGE, U), up(ctx, clg.WILSON, phi)
    F = up(ctx, clg.GAUGE, np.zeros(orc.volume(DIMS) * 72))
    p = clg.solver_params(max_dim=12, restart=300, accuracy=1e-13, stop_rule=clg.STOP_TRUE_RESIDUAL, even_odd=1)
    st, rep = clg.wilson_force(ctx, p, F, fphi, g, kappa)
    assert st == clg.OK
    assert rel(F.download(), want) < 1e-10
    x1, _, _, _ = orc.gmres(0, orc.D, DIMS, orc.BC_FERMION, kappa, U, phi, max_dim=30, restart=300,
                            accuracy=1e-14, stop_rule=1)
    e = clg.wilson_pf_energy(ctx, p, fphi, g, kappa)
    assert abs(e - np.dot(x1, x1)) < 1e-10 * np.dot(x1, x1)


@pytest.mark.parametrize("prec", [clg.F64, clg.F32])
def test_ks_force_driver_and_energy(ctx, prec):
    two_am = 0.2
    c = REMEZ["MinusOneOver2"]
    U = orc.gauge_random(DIMS, 0.4, 1)
    phi = orc.ks_gaussian(DIMS, 2)
    xs, _, _, _ = orc.msfom(1, orc.DDDAG, DIMS, orc.BC_FERMION, two_am, U, phi, c[4:], max_dim=30, restart=100,
                            accuracy=1e-14)
    ds = np.stack([orc.ks_op(DIMS, orc.BC_FERMION, 0.0, U, x, orc.D) for x in xs])
    want = orc.ks_force(DIMS, orc.BC_FERMION, U, xs, ds, c[1:4])
    g, fphi = up(ctx, clg.GAUGE, U), up(ctx, clg.KS, phi)
    F = up(ctx, clg.GAUGE, np.zeros(orc.volume(DIMS) * 72))
    acc = 1e-12 if prec == clg.F64 else 1e-5
    p = clg.solver_params(kind=clg.KS, max_dim=30, restart=100, accuracy=acc, absolute_accuracy=1)
    st, rep = clg.ks_force(ctx, p, F, fphi, g, two_am, c, prec=prec)
    assert st == clg.OK
    assert rel(F.download(), want) < (1e-9 if prec == clg.F64 else 1e-4)
    e = clg.ks_pf_energy(ctx, p, fphi, g, two_am, c, prec=prec)
    ref = c[0] * np.dot(phi, phi) + sum(a * np.real(orc.dot(phi, x)) for a, x in zip(c[1:4], xs))
    assert abs(e - ref) < (1e-11 if prec == clg.F64 else 1e-5) * abs(ref)
