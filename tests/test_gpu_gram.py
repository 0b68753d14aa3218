"""Gram matrices on the matrix cores (clg_gram_matrix, the kernels GCRO-DR's deflation uses) against
numpy's FP64 inner products of the downloaded fields: every entry of A^H B within 1e-12 of the largest."""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("kind,subset,dims,r,c", [
    (clg.WILSON, clg.FULL, (8, 8, 8, 8), 1, 1),
    (clg.WILSON, clg.FULL, (8, 8, 8, 8), 5, 13),
    (clg.WILSON, clg.EVEN, (6, 4, 4, 6), 17, 9),
    (clg.KS, clg.FULL, (8, 8, 8, 8), 32, 32),
    (clg.KS, clg.ODD, (4, 4, 4, 4), 16, 32),
])
def test_gram_matrix_against_numpy(ctx, kind, subset, dims, r, c):
    print(f"gram kind {kind} subset {subset} dims {dims} r {r} c {c}", flush=True)
    fields = []
    for s in range(r + c):
        full = clg.Field(ctx, kind, clg.F64, dims, bc=orc.BC_FERMION)
        full.gaussian(50 + s)
        if subset == clg.FULL:
            fields.append(full)
        else:
            f = clg.Field(ctx, kind, clg.F64, dims, subset=subset, bc=orc.BC_FERMION)
            f.copy_from(full.parity(subset))
            fields.append(f)
    A, B = fields[:r], fields[r:]
    got = clg.gram_matrix(ctx, A, B)
    ha = [orc.as_complex(f.download()) for f in A]
    hb = [orc.as_complex(f.download()) for f in B]
    ref = np.array([[np.vdot(a, b) for b in hb] for a in ha])
    err = np.abs(got - ref).max() / np.abs(ref).max()
    assert err < 1e-12, err
    # A^H A is Hermitian with the squared norms on its diagonal
    g2 = clg.gram_matrix(ctx, A, A)
    assert np.abs(g2 - g2.conj().T).max() <= 1e-12 * np.abs(g2).max()
    assert np.allclose(np.diag(g2).real, [np.vdot(a, a).real for a in ha], rtol=1e-12, atol=0)
