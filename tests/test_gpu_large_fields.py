"""Fields above 4 GiB: a 64^3x32 lattice (8.4 M sites), whose FP64 18-real gauge field takes 4.8 GB.

The kernels form byte offsets in 64 bits from a 32-bit element index; with 32-bit offsets the links past
4 GiB would be read from wrapped (wrong) places inside the same allocation, so the result would differ
from the oracle's by O(1), not fault.  Wilson D and staggered D against the oracle's operators on the
GPU-generated inputs, downloaded (about 0.5 s per oracle application on the host cores)."""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

DIMS = (64, 64, 64, 32)


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


def maxrel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_operators_on_a_gauge_field_above_4gib(ctx):
    g = clg.Field(ctx, clg.GAUGE, clg.F64, DIMS)
    assert g.nbytes() > 4 * 2 ** 30
    g.gauge_random(0.5, 1)
    U = g.download()
    psi = clg.Field(ctx, clg.WILSON, clg.F64, DIMS, bc=orc.BC_FERMION)
    psi.gaussian(2)
    out = clg.Field(ctx, clg.WILSON, clg.F64, DIMS)
    clg.wilson_apply(ctx, out, psi, g, 0.1355, clg.EFO_F_D)
    ref = orc.wilson_op(DIMS, orc.BC_FERMION, 0.1355, U, psi.download(), orc.D)
    err = maxrel(out.download(), ref)
    print(f"Wilson D at {DIMS}: max rel err {err:.2e}")
    assert err < 1e-12
    del out, psi, ref
    chi = clg.Field(ctx, clg.KS, clg.F64, DIMS, bc=orc.BC_FERMION)
    chi.gaussian(3)
    outk = clg.Field(ctx, clg.KS, clg.F64, DIMS)
    clg.ks_apply(ctx, outk, chi, g, 0.05, clg.EFO_F_D)
    refk = orc.ks_op(DIMS, orc.BC_FERMION, 0.05, U, chi.download(), orc.D)
    errk = maxrel(outk.download(), refk)
    print(f"staggered D at {DIMS}: max rel err {errk:.2e}")
    assert errk < 1e-12
