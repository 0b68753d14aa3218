"""Pins of the CPU oracle: index/BC/eta/even-odd conventions, gamma table, RNG,
synthetic gauge generator, Wilson and staggered operators.

Each check is chosen so that a plausible mistake in the oracle (a dropped term, a
wrong sign, a transposed link, a swapped projector) fails it: closed forms (free
plane waves), invariants (gamma5-hermiticity, anti-hermiticity, gauge covariance),
known-answer vectors (Philox) and independent library routines (scipy expm).
"""
import itertools

import numpy as np
import pytest
import scipy.linalg

import orc

DIMS_SMALL = (4, 4, 4, 4)
KAPPA = 0.1355


# ----------------------------------------------------------------------------- index
def test_site_index_lexicographic_t_fastest():
    dims = (3, 4, 5, 6)
    for x, y, z, t in [(0, 0, 0, 1), (0, 0, 1, 0), (0, 1, 0, 0), (1, 0, 0, 0), (2, 3, 4, 5)]:
        assert orc.site_index(dims, x, y, z, t) == ((x * 4 + y) * 5 + z) * 6 + t
    assert orc.site_index(dims, 0, 0, 0, 1) == 1


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (2, 4, 6, 8), (5, 5, 4, 4), (3, 2, 3, 2)])
def test_eo_sites_reference_convention(dims):
    V = orc.volume(dims)
    coords = np.array(list(itertools.product(*[range(d) for d in dims])))
    par = coords.sum(1) & 1
    ev, od = orc.eo_sites(dims, 0), orc.eo_sites(dims, 1)
    e = np.arange(V // 2)
    # each half-index e picks the member of the pair (2e, 2e+1) with the wanted parity
    assert np.all((ev >> 1) == e) and np.all((od >> 1) == e)
    assert np.all(par[ev] == 0) and np.all(par[od] == 1)
    assert sorted(np.concatenate([ev, od]).tolist()) == list(range(V))


def test_eo_sites_rejects_odd_t():
    import ctypes as C
    out = np.zeros(100, np.int32)
    rc = orc.lib().orc_eo_sites(orc._i4((2, 2, 2, 3)), 0, out.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc != 0


@pytest.mark.parametrize("dims", [(4, 4, 4, 4), (3, 4, 5, 6)])
def test_moves_and_signs(dims):
    bc = (1, -1, 1, -1)
    fwd, bwd, fsg, bsg = orc.moves(dims, bc)
    for s, (x, y, z, t) in enumerate(itertools.product(*[range(d) for d in dims])):
        c = [x, y, z, t]
        for mu in range(4):
            cp = list(c)
            cp[mu] = (c[mu] + 1) % dims[mu]
            cm = list(c)
            cm[mu] = (c[mu] - 1) % dims[mu]
            assert fwd[s * 4 + mu] == orc.site_index(dims, *cp)
            assert bwd[s * 4 + mu] == orc.site_index(dims, *cm)
            assert fsg[s * 4 + mu] == (bc[mu] if c[mu] == dims[mu] - 1 else 1)
            assert bsg[s * 4 + mu] == (bc[mu] if c[mu] == 0 else 1)


def test_eta_table():
    dims = (3, 4, 2, 2)
    e = orc.eta(dims)
    for s, (x, y, z, t) in enumerate(itertools.product(*[range(d) for d in dims])):
        assert (e[s] >> 0) & 1 == 0
        assert (e[s] >> 1) & 1 == x & 1
        assert (e[s] >> 2) & 1 == (x + y) & 1
        assert (e[s] >> 3) & 1 == (x + y + z) & 1
        assert (e[s] >> 4) & 1 == (x + y + z + t) & 1


# ----------------------------------------------------------------------------- gamma
def test_gamma_clifford_algebra():
    g = [orc.gamma(i) for i in range(5)]
    I = np.eye(4)
    for mu in range(4):
        assert np.allclose(g[mu], g[mu].conj().T)
        for nu in range(4):
            assert np.allclose(g[mu] @ g[nu] + g[nu] @ g[mu], 2 * I * (mu == nu))
        assert np.allclose(g[4] @ g[mu] + g[mu] @ g[4], 0)
    assert np.allclose(g[4], np.diag([1, 1, -1, -1]))
    # chiral basis: gamma5 = gamma1 gamma2 gamma3 gamma4 (fixes the phases of the table)
    assert np.allclose(g[0] @ g[1] @ g[2] @ g[3], g[4])


# ----------------------------------------------------------------------------- RNG
def test_philox_known_answers():
    # Random123 known-answer vectors for philox4x32 with 10 rounds
    assert orc.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert orc.philox([f, f, f, f], [f, f]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert orc.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def _gauss_pair(idx, draw, stream, seed):
    w = orc.philox([idx, draw, stream, 0], [seed, 0x434C4721])
    u = [((w[0] >> 5) * 67108864.0 + (w[1] >> 6) + 0.5) / 9007199254740992.0,
         ((w[2] >> 5) * 67108864.0 + (w[3] >> 6) + 0.5) / 9007199254740992.0]
    r = np.sqrt(-2 * np.log(u[0]))
    return r * np.cos(2 * np.pi * u[1]), r * np.sin(2 * np.pi * u[1])


def _reunitarise(m):
    m = m.copy()
    m[0] /= np.linalg.norm(m[0])
    m[1] -= np.vdot(m[0], m[1]) * m[0]
    m[1] /= np.linalg.norm(m[1])
    m[2] = np.conj(np.cross(m[0], m[1]))
    return m


def test_gauge_generator_matches_scipy_expm():
    dims = (2, 2, 2, 4)
    eps, seed = 0.5, 1
    U = orc.gauge_random(dims, eps, seed)
    for link in [0, 1, 7, 33, 63]:
        g = np.concatenate([_gauss_pair(link, d, 0, seed) for d in range(5)])
        H = np.zeros((3, 3), complex)
        tr = (g[0] + g[1] + g[2]) / 3
        H[0, 0], H[1, 1], H[2, 2] = g[0] - tr, g[1] - tr, g[2] - tr
        H[0, 1] = (g[3] + 1j * g[4]) / np.sqrt(2)
        H[0, 2] = (g[5] + 1j * g[6]) / np.sqrt(2)
        H[1, 2] = (g[7] + 1j * g[8]) / np.sqrt(2)
        H[1, 0], H[2, 0], H[2, 1] = np.conj(H[0, 1]), np.conj(H[0, 2]), np.conj(H[1, 2])
        ref = _reunitarise(scipy.linalg.expm(1j * eps * H))
        got = orc.as_complex(U[link * 18:(link + 1) * 18]).reshape(3, 3)
        assert np.abs(got - ref).max() < 1e-13


def test_gauge_links_are_su3():
    U = orc.as_complex(orc.gauge_random(DIMS_SMALL, 0.5, 1)).reshape(-1, 3, 3)
    eye = np.einsum("nij,nkj->nik", U, U.conj())
    assert np.abs(eye - np.eye(3)).max() < 1e-13
    assert np.abs(np.linalg.det(U) - 1).max() < 1e-13


def test_su3_norm_is_reference_gram_schmidt():
    rng = np.random.default_rng(5)
    m = rng.normal(size=(4, 3, 3)) + 1j * rng.normal(size=(4, 3, 3))
    got = orc.as_complex(orc.su3_norm(orc.from_complex(m.ravel()))).reshape(4, 3, 3)
    for k in range(4):
        assert np.abs(got[k] - _reunitarise(m[k])).max() < 1e-13


def test_gaussian_source_normalisation():
    # TestCommon.cpp:27-34: |psi|^2 ~ 12 V for a Gaussian Wilson source (complex components of unit
    # variance); on 16^4 the estimator's standard deviation is 1/sqrt(12 V) = 0.0011, held to 5 of them
    dims = (16, 16, 16, 16)
    psi = orc.wilson_gaussian(dims, 2)
    nw = 12 * orc.volume(dims)
    assert abs(np.dot(psi, psi) / nw - 1) < 5 / np.sqrt(nw)
    assert abs(psi.mean()) < 5 / np.sqrt(2 * nw)
    chi = orc.ks_gaussian(dims, 2)
    nk = 3 * orc.volume(dims)
    assert abs(np.dot(chi, chi) / nk - 1) < 5 / np.sqrt(nk)


# ----------------------------------------------------------------------------- Wilson
def _coords(dims):
    return np.array(list(itertools.product(*[range(d) for d in dims])))


def _plane_wave_spinor(dims, p, u):
    X = _coords(dims)
    ph = np.exp(1j * X @ np.asarray(p))
    return orc.from_complex((ph[:, None] * u[None, :]).ravel())


@pytest.mark.parametrize("bc", [(1, 1, 1, -1), (1, 1, 1, 1)])
def test_wilson_free_field_plane_wave(bc):
    dims = (4, 4, 4, 6)
    rng = np.random.default_rng(1)
    n = [1, 3, 2, 1]
    p = [2 * np.pi * n[i] / dims[i] + (np.pi / dims[i] if bc[i] < 0 else 0) for i in range(4)]
    u = rng.normal(size=12) + 1j * rng.normal(size=12)
    psi = _plane_wave_spinor(dims, p, u)
    U = orc.gauge_unit(dims)
    out = orc.as_complex(orc.wilson_op(dims, bc, KAPPA, U, psi)).reshape(-1, 12)
    g = [orc.gamma(i) for i in range(4)]
    M = np.eye(4, dtype=complex)
    for mu in range(4):
        M -= KAPPA * (2 * np.cos(p[mu]) * np.eye(4) - 2j * np.sin(p[mu]) * g[mu])
    Mfull = np.kron(M, np.eye(3))  # spin index major, colour minor
    expect = (orc.as_complex(psi).reshape(-1, 12)) @ Mfull.T
    assert np.abs(out - expect).max() < 1e-12


def _random_su3(rng, n):
    z = rng.normal(size=(n, 3, 3)) + 1j * rng.normal(size=(n, 3, 3))
    q, r = np.linalg.qr(z)
    q = q * (np.diagonal(r, axis1=1, axis2=2) / np.abs(np.diagonal(r, axis1=1, axis2=2)))[:, None, :]
    return q / np.linalg.det(q)[:, None, None] ** (1 / 3)


def _gauge_transform(dims, U, g):
    """U'_mu(x) = g(x) U_mu(x) g(x+mu)^dagger"""
    V = orc.volume(dims)
    Uc = orc.as_complex(U).reshape(V, 4, 3, 3)
    fwd = orc.moves(dims, (1, 1, 1, 1))[0].reshape(V, 4)
    out = np.einsum("xij,xmjk,xmlk->xmil", g, Uc, g[fwd].conj())
    return orc.from_complex(out.ravel())


def test_wilson_gauge_covariance():
    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(3)
    U = orc.gauge_random(dims, 0.5, 1)
    psi = orc.wilson_gaussian(dims, 2)
    g = _random_su3(rng, orc.volume(dims))
    Ug = _gauge_transform(dims, U, g)
    gpsi = orc.from_complex(np.einsum("xij,xsj->xsi", g, orc.as_complex(psi).reshape(-1, 4, 3)).ravel())
    for op in (orc.D, orc.DDAG):
        a = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, Ug, gpsi, op)).reshape(-1, 4, 3)
        b = np.einsum("xij,xsj->xsi", g, orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, op)).reshape(-1, 4, 3))
        assert np.abs(a - b).max() < 1e-12


def test_wilson_adjoint_and_gamma5_hermiticity():
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    psi = orc.wilson_gaussian(dims, 2)
    chi = orc.wilson_gaussian(dims, 7)
    Dpsi = orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.D)
    Ddchi = orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, chi, orc.DDAG)
    assert abs(orc.dot(chi, Dpsi) - orc.dot(Ddchi, psi)) < 1e-10 * abs(orc.dot(chi, Dpsi))
    # D^dagger = gamma5 D gamma5
    g5 = np.kron(np.diag([1, 1, -1, -1]), np.eye(3))
    c = orc.as_complex(chi).reshape(-1, 12)
    g5chi = orc.from_complex((c @ g5).ravel())
    alt = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, g5chi, orc.D)).reshape(-1, 12) @ g5
    assert np.abs(alt - orc.as_complex(Ddchi).reshape(-1, 12)).max() < 1e-12


def test_wilson_dense_gamma5_hermitian_small_lattice():
    # TestGamma5Hermitian (CFieldFermionWilsonSquareSU3.cu:1362-1500) on a 5x5x4x4-style odd box
    dims = (3, 2, 2, 2)
    U = orc.gauge_random(dims, 0.8, 4)
    N = orc.volume(dims) * 12
    Dm = np.zeros((N, N), complex)
    for j in range(N):
        e = np.zeros(N, complex)
        e[j] = 1
        Dm[:, j] = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, orc.from_complex(e)))
    G5 = np.kron(np.eye(orc.volume(dims)), np.kron(np.diag([1, 1, -1, -1]), np.eye(3)))
    A = G5 @ Dm
    assert np.abs(A - A.conj().T).max() < 1e-13


def test_wilson_antiperiodic_sign():
    # a constant field: the only t-dependence comes from the boundary sign
    dims = (2, 2, 2, 4)
    U = orc.gauge_unit(dims)
    psi = orc.from_complex(np.tile(np.arange(1, 13) * (1 + 0.5j), orc.volume(dims)))
    per = orc.as_complex(orc.wilson_op(dims, (1, 1, 1, 1), KAPPA, U, psi)).reshape(2, 2, 2, 4, 12)
    apb = orc.as_complex(orc.wilson_op(dims, (1, 1, 1, -1), KAPPA, U, psi)).reshape(2, 2, 2, 4, 12)
    assert np.abs(per[:, :, :, 1:3] - apb[:, :, :, 1:3]).max() < 1e-14  # interior t unchanged
    assert np.abs(per[:, :, :, 0] - apb[:, :, :, 0]).max() > 0.1  # t=0 sees the flipped backward hop


def test_wilson_composites_and_coefficients():
    dims = (4, 4, 2, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    psi = orc.wilson_gaussian(dims, 2)
    Dp = orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.D)
    Ddp = orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.DDAG)
    assert np.allclose(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.DD),
                       orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, Dp, orc.D), atol=1e-13)
    assert np.allclose(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.DDDAG),
                       orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, Ddp, orc.D), atol=1e-13)
    assert np.allclose(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.D, 1), -Dp, atol=1e-14)
    assert np.allclose(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.D, 2, 0.25), 0.25 * Dp, atol=1e-14)
    z = orc.as_complex(orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi, orc.D, 3, 0.3, -0.7))
    assert np.allclose(z, (0.3 - 0.7j) * orc.as_complex(Dp), atol=1e-14)


def test_wilson_f32_agrees_with_f64():
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.5, 1)
    psi = orc.wilson_gaussian(dims, 2)
    a = orc.wilson_op(dims, orc.BC_FERMION, KAPPA, U, psi)
    b = orc.wilson_op_f32(dims, orc.BC_FERMION, KAPPA, U, psi)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-6


# ----------------------------------------------------------------------------- staggered
def test_ks_massless_antihermitian_dense():
    # TestAntiHermitianS (CFieldFermionKSSU3.cu:1424-1490)
    dims = (2, 2, 2, 4)
    U = orc.gauge_random(dims, 0.7, 3)
    N = orc.volume(dims) * 3
    Dm = np.zeros((N, N), complex)
    for j in range(N):
        e = np.zeros(N, complex)
        e[j] = 1
        Dm[:, j] = orc.as_complex(orc.ks_op(dims, orc.BC_FERMION, 0.0, U, orc.from_complex(e)))
    assert np.abs(Dm + Dm.conj().T).max() < 1e-13
    assert np.abs(Dm).max() > 0.1


def test_ks_dagger_and_mass():
    dims = (4, 4, 4, 4)
    U = orc.gauge_random(dims, 0.4, 1)
    chi = orc.ks_gaussian(dims, 2)
    m = 0.05
    D = orc.ks_op(dims, orc.BC_FERMION, m, U, chi, orc.D)
    Dd = orc.ks_op(dims, orc.BC_FERMION, m, U, chi, orc.DDAG)
    D0 = orc.ks_op(dims, orc.BC_FERMION, 0.0, U, chi, orc.D)
    assert np.allclose(D, m * chi + D0, atol=1e-13)
    assert np.allclose(Dd, m * chi - D0, atol=1e-13)


@pytest.mark.parametrize("two_am", [0.0, 0.1])
def test_ks_free_field_ddagger(two_am):
    dims = (4, 4, 4, 6)
    bc = orc.BC_FERMION
    n = [1, 2, 3, 0]
    p = [2 * np.pi * n[i] / dims[i] + (np.pi / dims[i] if bc[i] < 0 else 0) for i in range(4)]
    X = _coords(dims)
    u = np.array([1.0, 0.5 - 0.2j, -0.3j])
    chi = orc.from_complex((np.exp(1j * X @ np.asarray(p))[:, None] * u[None, :]).ravel())
    U = orc.gauge_unit(dims)
    out = orc.ks_op(dims, bc, two_am, U, chi, orc.DDDAG)
    lam = two_am ** 2 + 4 * sum(np.sin(pp) ** 2 for pp in p)
    assert np.abs(orc.as_complex(out) - lam * orc.as_complex(chi)).max() < 1e-12


def test_ks_gauge_covariance():
    dims = (4, 4, 4, 4)
    rng = np.random.default_rng(9)
    U = orc.gauge_random(dims, 0.4, 1)
    chi = orc.ks_gaussian(dims, 2)
    g = _random_su3(rng, orc.volume(dims))
    Ug = _gauge_transform(dims, U, g)
    gchi = orc.from_complex(np.einsum("xij,xj->xi", g, orc.as_complex(chi).reshape(-1, 3)).ravel())
    a = orc.as_complex(orc.ks_op(dims, orc.BC_FERMION, 0.05, Ug, gchi)).reshape(-1, 3)
    b = np.einsum("xij,xj->xi", g, orc.as_complex(orc.ks_op(dims, orc.BC_FERMION, 0.05, U, chi)).reshape(-1, 3))
    assert np.abs(a - b).max() < 1e-12
