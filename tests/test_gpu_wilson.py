"""GPU parity of the Wilson operator, layout and generators against the CPU oracle.

Bars: FP64 within 1e-12 relative, FP32 within 1e-6 relative (max-norm over the field
relative to the max of the oracle result), even-odd indexing bit-exact.
"""
import numpy as np
import pytest

import cudalatticegauge_amd as clg
import orc

pytestmark = pytest.mark.gpu

KAPPA = 0.1355
DIMS0 = (8, 8, 8, 8)          # 8^4 lattice, eps 0.3 gauge
ODD_BOX = (5, 3, 4, 4)        # odd extents in x, y (full-lattice operator only)


def rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.fixture(scope="module")
def ctx():
    c = clg.Context(0)
    print("device:", c.info(), flush=True)
    yield c
    c.close()


def upload(ctx, kind, prec, dims, data, subset=clg.FULL, bc=orc.BC_FERMION):
    f = clg.Field(ctx, kind, prec, dims, subset=subset, bc=bc)
    f.upload(data)
    return f


@pytest.fixture(scope="module")
def c0(ctx):
    U = orc.gauge_random(DIMS0, 0.3, 1)
    psi = orc.wilson_gaussian(DIMS0, 2)
    return U, psi


def test_upload_download_roundtrip_bit_exact(ctx, c0):
    U, psi = c0
    g = upload(ctx, clg.GAUGE, clg.F64, DIMS0, U)
    f = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    assert np.array_equal(g.download(), U)
    assert np.array_equal(f.download(), psi)


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (2, 4, 6, 8), (5, 3, 4, 4)])
def test_even_odd_views_bit_exact(ctx, dims):
    psi = orc.wilson_gaussian(dims, 3)
    f = upload(ctx, clg.WILSON, clg.F64, dims, psi)
    for sub, par in ((clg.EVEN, 0), (clg.ODD, 1)):
        half = f.parity(sub).download().reshape(-1, 24)
        sites = orc.eo_sites(dims, par)
        assert np.array_equal(half, psi.reshape(-1, 24)[sites])


def test_gpu_generators_match_oracle(ctx):
    dims = (4, 4, 4, 8)
    g = clg.Field(ctx, clg.GAUGE, clg.F64, dims)
    g.gauge_random(0.5, 1)
    assert np.abs(g.download() - orc.gauge_random(dims, 0.5, 1)).max() < 1e-13
    f = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    f.gaussian(2)
    assert np.abs(f.download() - orc.wilson_gaussian(dims, 2)).max() < 1e-13
    k = clg.Field(ctx, clg.KS, clg.F64, dims)
    k.gaussian(2)
    assert np.abs(k.download() - orc.ks_gaussian(dims, 2)).max() < 1e-13


@pytest.mark.parametrize("op", [clg.EFO_F_D, clg.EFO_F_Ddagger, clg.EFO_F_DD, clg.EFO_F_DDdagger])
def test_wilson_f64_parity_c0(ctx, c0, op):
    U, psi = c0
    g = upload(ctx, clg.GAUGE, clg.F64, DIMS0, U)
    fi = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    fo = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0)
    clg.wilson_apply(ctx, fo, fi, g, KAPPA, op)
    ref = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, psi, op)
    assert rel(fo.download(), ref) < 1e-12


@pytest.mark.parametrize("coeff", [(clg.EOCT_Minus, 0, 0), (clg.EOCT_Real, 0.7, 0), (clg.EOCT_Complex, 0.3, -1.1)])
def test_wilson_coefficients(ctx, c0, coeff):
    U, psi = c0
    g = upload(ctx, clg.GAUGE, clg.F64, DIMS0, U)
    fi = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    fo = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0)
    clg.wilson_apply(ctx, fo, fi, g, KAPPA, clg.EFO_F_DDdagger, *coeff)
    ref = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, psi, orc.DDDAG, *coeff)
    assert rel(fo.download(), ref) < 1e-12


def test_wilson_odd_extents_and_point_source(ctx):
    dims = ODD_BOX
    U = orc.gauge_random(dims, 0.5, 4)
    g = upload(ctx, clg.GAUGE, clg.F64, dims, U)
    fi = clg.Field(ctx, clg.WILSON, clg.F64, dims, bc=orc.BC_FERMION)
    fo = clg.Field(ctx, clg.WILSON, clg.F64, dims)
    for coord, spin, col in [((0, 0, 0, 0), 0, 0), ((4, 2, 3, 3), 3, 2), ((2, 1, 0, 3), 1, 1)]:
        fi.point_source(coord, spin, col)
        src = fi.download()
        e = np.zeros_like(src)
        e[orc.site_index(dims, *coord) * 24 + 2 * (spin * 3 + col)] = 1.0
        assert np.array_equal(src, e)
        clg.wilson_apply(ctx, fo, fi, g, KAPPA, clg.EFO_F_D)
        assert rel(fo.download(), orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, e)) < 1e-12


@pytest.mark.parametrize("recon", [clg.GAUGE, clg.GAUGE12])
def test_wilson_f32_parity(ctx, c0, recon):
    U, psi = c0
    g = upload(ctx, recon, clg.F32, DIMS0, U)
    fi = upload(ctx, clg.WILSON, clg.F32, DIMS0, psi)
    fo = clg.Field(ctx, clg.WILSON, clg.F32, DIMS0)
    for op in (clg.EFO_F_D, clg.EFO_F_Ddagger):
        clg.wilson_apply(ctx, fo, fi, g, KAPPA, op)
        ref = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, psi, op)
        assert rel(fo.download(), ref) < 1e-6


def test_wilson_f64_recon12(ctx, c0):
    U, psi = c0
    g = upload(ctx, clg.GAUGE12, clg.F64, DIMS0, U)
    fi = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    fo = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0)
    clg.wilson_apply(ctx, fo, fi, g, KAPPA, clg.EFO_F_D)
    assert rel(fo.download(), orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, psi)) < 1e-12


@pytest.mark.parametrize("dag", [False, True])
def test_even_odd_hop_matches_full_operator(ctx, c0, dag):
    U, psi = c0
    g = upload(ctx, clg.GAUGE, clg.F64, DIMS0, U)
    f = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    op = orc.DDAG if dag else orc.D
    ref = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, psi, op).reshape(-1, 24)
    for out_sub, par in ((clg.EVEN, 0), (clg.ODD, 1)):
        in_sub = clg.ODD if out_sub == clg.EVEN else clg.EVEN
        out = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0, subset=out_sub, bc=orc.BC_FERMION)
        clg.wilson_hop(ctx, out, f.parity(in_sub), g, dagger=dag, a=-KAPPA, x=f.parity(out_sub), b=1.0)
        want = ref[orc.eo_sites(DIMS0, par)]
        assert rel(out.download().reshape(-1, 24), want) < 1e-12


def test_schur_operator_matches_two_full_applications(ctx, c0):
    U, psi = c0
    V = orc.volume(DIMS0)
    ev, od = orc.eo_sites(DIMS0, 0), orc.eo_sites(DIMS0, 1)
    p_full = psi.reshape(V, 24).copy()
    p_full[od] = 0
    Dp = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, p_full.ravel()).reshape(V, 24)
    phi = np.zeros((V, 24))
    phi[od] = -Dp[od] / KAPPA                 # H_oe psi_e
    Dphi = orc.wilson_op(DIMS0, orc.BC_FERMION, KAPPA, U, phi.ravel()).reshape(V, 24)
    want = psi.reshape(V, 24)[ev] + KAPPA * Dphi[ev]   # psi_e - kappa^2 H_eo H_oe psi_e
    g = upload(ctx, clg.GAUGE, clg.F64, DIMS0, U)
    fe = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi.reshape(V, 24)[ev].ravel(), subset=clg.EVEN)
    out = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0, subset=clg.EVEN, bc=orc.BC_FERMION)
    clg.wilson_schur(ctx, out, fe, g, KAPPA, clg.EFO_F_D)
    assert rel(out.download().reshape(-1, 24), want) < 1e-12
    # Mhat^dagger is the adjoint of Mhat
    chi = orc.wilson_gaussian(DIMS0, 9).reshape(V, 24)[ev].ravel()
    fc = upload(ctx, clg.WILSON, clg.F64, DIMS0, chi, subset=clg.EVEN)
    outd = clg.Field(ctx, clg.WILSON, clg.F64, DIMS0, subset=clg.EVEN, bc=orc.BC_FERMION)
    clg.wilson_schur(ctx, outd, fc, g, KAPPA, clg.EFO_F_Ddagger)
    lhs = orc.dot(chi, out.download())
    rhs = orc.dot(outd.download(), psi.reshape(V, 24)[ev].ravel())
    assert abs(lhs - rhs) < 1e-11 * abs(lhs)


def test_blas_dot_norm_axpy(ctx, c0):
    _, psi = c0
    chi = orc.wilson_gaussian(DIMS0, 5)
    a = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    b = upload(ctx, clg.WILSON, clg.F64, DIMS0, chi)
    d = clg.dot(ctx, a, b)
    want = orc.dot(psi, chi)
    assert abs(d - want) < 1e-12 * abs(want) * 10
    assert abs(clg.norm2(ctx, a) - np.dot(psi, psi)) < 1e-12 * np.dot(psi, psi)
    clg.axpy(ctx, b, 0.5 - 0.25j, a)
    z = orc.as_complex(chi) + (0.5 - 0.25j) * orc.as_complex(psi)
    assert np.abs(orc.as_complex(b.download()) - z).max() < 1e-14
    for prec in (clg.F32,):
        a32 = upload(ctx, clg.WILSON, prec, DIMS0, psi)
        assert abs(clg.norm2(ctx, a32) / np.dot(psi, psi) - 1) < 1e-6


def test_precision_conversion(ctx, c0):
    _, psi = c0
    a = upload(ctx, clg.WILSON, clg.F64, DIMS0, psi)
    b = clg.Field(ctx, clg.WILSON, clg.F32, DIMS0)
    h = clg.Field(ctx, clg.WILSON, clg.F16, DIMS0)
    b.copy_from(a)
    h.copy_from(a)
    assert rel(b.download(), psi) < 1e-6
    assert rel(h.download(), psi) < 2e-3


def tiled_fits(dims, xcd):
    """Whether tiled plane schedule xcd = 10*TY + m fits the lattice: 8/TY tiles along x, TY along y."""
    TY = xcd // 10
    return dims[0] % (8 // TY) == 0 and dims[1] % TY == 0


TILED = (20, 21, 40, 41, 80, 81)


@pytest.mark.parametrize("dims", [(4, 6, 8, 16), (4, 4, 16, 16), (8, 8, 8, 16)])
def test_lds_staged_launches(ctx, dims):
    """Pin modes 5..8 stage the z and t neighbours in LDS (one workgroup per (x,y) plane of one
    parity, block = L2*L3/2: 64 and 128 here), also with several planes per workgroup in turn
    (sites per thread); a forced staged mode whose block does not fit the lattice is refused."""
    plane = dims[2] * dims[3] // 2
    U = orc.gauge_random(dims, 0.3, 11)
    psi = orc.wilson_gaussian(dims, 12)
    refs = {dag: orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.DDAG if dag else orc.D)
            for dag in (False, True)}
    g64 = upload(ctx, clg.GAUGE, clg.F64, dims, U)
    f64 = upload(ctx, clg.WILSON, clg.F64, dims, psi)
    g32 = clg.Field(ctx, clg.GAUGE12, clg.F32, dims)
    g32.copy_from(g64)
    f32 = clg.Field(ctx, clg.WILSON, clg.F32, dims, bc=orc.BC_FERMION)
    f32.copy_from(f64)
    g16 = clg.Field(ctx, clg.GAUGE12, clg.F16, dims)
    g16.copy_from(g64)
    f16 = clg.Field(ctx, clg.WILSON, clg.F16, dims, bc=orc.BC_FERMION)
    f16.copy_from(f64)
    cases = ((f64, g64, 1e-12), (f32, g32, 1e-6), (f16, g16, 1e-2))
    ref_eo = refs[False].reshape(-1, 24)
    try:
        tiled = [(5 + i % 4, plane, xcd, 1) for i, xcd in enumerate(TILED) if tiled_fits(dims, xcd)]
        for pm, blk, xcd, spt in [(1, plane, 8, 1), (5, plane, 8, 1), (6, plane, 8, 1), (6, plane, 0, 1),
                                  (5, plane, 1, 1), (7, plane, 8, 1), (8, plane, 0, 1),
                                  (6, plane, 8, 2), (8, plane, 8, 4), (5, plane, 0, 4)] + tiled:
            clg.autotune_force(pm, blk, xcd, spt)
            for f, g, tol in cases:
                for dag in (False, True):
                    out = clg.Field(ctx, clg.WILSON, f.prec, dims)
                    clg.wilson_apply(ctx, out, f, g, KAPPA, clg.EFO_F_Ddagger if dag else clg.EFO_F_D)
                    assert rel(out.download(), refs[dag]) < tol, (pm, blk, xcd, spt, f.prec, dag)
            # even-odd hops read a half field: out_p = x_p - kappa H psi_(1-p)
            for out_sub, par in ((clg.EVEN, 0), (clg.ODD, 1)):
                in_sub = clg.ODD if out_sub == clg.EVEN else clg.EVEN
                out = clg.Field(ctx, clg.WILSON, clg.F64, dims, subset=out_sub, bc=orc.BC_FERMION)
                clg.wilson_hop(ctx, out, f64.parity(in_sub), g64, dagger=False, a=-KAPPA, x=f64.parity(out_sub), b=1.0)
                assert rel(out.download().reshape(-1, 24), ref_eo[orc.eo_sites(dims, par)]) < 1e-12, (pm, par)
        # a staged mode forced with a block other than the plane size fails loudly, for the full
        # operator and for the half-lattice hop, and leaves the output untouched
        for pm in (5, 8):
            clg.autotune_force(pm, 256, 8, 1)
            out = clg.Field(ctx, clg.WILSON, clg.F64, dims)
            with pytest.raises(clg.CLGError):
                clg.wilson_apply(ctx, out, f64, g64, KAPPA, clg.EFO_F_D)
            assert not np.any(out.download())
            oute = clg.Field(ctx, clg.WILSON, clg.F64, dims, subset=clg.EVEN, bc=orc.BC_FERMION)
            with pytest.raises(clg.CLGError):
                clg.wilson_hop(ctx, oute, f64.parity(clg.ODD), g64, dagger=False, a=-KAPPA)
        # so is a tiled schedule whose tiles do not divide the plane grid
        for xcd in TILED:
            if not tiled_fits(dims, xcd):
                clg.autotune_force(6, plane, xcd, 1)
                out = clg.Field(ctx, clg.WILSON, clg.F64, dims)
                with pytest.raises(clg.CLGError):
                    clg.wilson_apply(ctx, out, f64, g64, KAPPA, clg.EFO_F_D)
                assert not np.any(out.download())
    finally:
        clg.autotune_force(0)


@pytest.mark.parametrize("dims", [(2, 4, 6, 8), (6, 6, 6, 6), (4, 8, 4, 12), (2, 2, 16, 16), (4, 2, 32, 16),
                                  (8, 4, 8, 10)])
def test_operators_on_lattice_shapes(ctx, dims):
    """Tuned launches (whatever the tuner picks per shape: staged planes where L2*L3/2 is 64, 128 or 256,
    XCD slabs, sites per thread) against the oracle, Wilson D / D^dagger and staggered D / D^dagger."""
    U = orc.gauge_random(dims, 0.4, 21)
    psi = orc.wilson_gaussian(dims, 22)
    chi = orc.ks_gaussian(dims, 23)
    g = upload(ctx, clg.GAUGE, clg.F64, dims, U)
    fw = upload(ctx, clg.WILSON, clg.F64, dims, psi)
    fk = upload(ctx, clg.KS, clg.F64, dims, chi)
    clg.autotune_enable(1)
    try:
        for op in (clg.EFO_F_D, clg.EFO_F_Ddagger):
            out = clg.Field(ctx, clg.WILSON, clg.F64, dims)
            clg.wilson_apply(ctx, out, fw, g, KAPPA, op)
            assert rel(out.download(), orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, op)) < 1e-12, op
            outk = clg.Field(ctx, clg.KS, clg.F64, dims)
            clg.ks_apply(ctx, outk, fk, g, 0.07, op)
            assert rel(outk.download(), orc.ks_op(dims, orc.BC_FERMION, 0.07, U, chi, op)) < 1e-12, op
    finally:
        clg.autotune_enable(-1)
