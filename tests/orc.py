"""ctypes wrapper of the CPU oracle (oracle/liborc.so).

Test infrastructure only: the oracle is the checker the HIP path is compared
against.  Host arrays use the reference's AoS order (see oracle/clg_oracle.h).
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
_LIB = None

D, DDAG, DD, DDDAG = 0, 1, 2, 3
BC_FERMION = (1, 1, 1, -1)
BC_PERIODIC = (1, 1, 1, 1)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ORACLE_DIR, "liborc.so")
        if not os.path.exists(path):
            subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])
        _LIB = C.CDLL(path)
    return _LIB


def _i4(v):
    return (C.c_int * 4)(*[int(a) for a in v])


def _dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


def volume(dims):
    return int(np.prod(dims))


class Operator(C.Structure):
    _fields_ = [("kind", C.c_int), ("op", C.c_int), ("dims", C.c_int * 4),
                ("bc", C.c_int * 4), ("mass_or_kappa", C.c_double)]


def make_op(kind, op, dims, bc, m):
    o = Operator()
    o.kind, o.op, o.mass_or_kappa = kind, op, float(m)
    for i in range(4):
        o.dims[i] = dims[i]
        o.bc[i] = bc[i]
    return o


def site_index(dims, x, y, z, t):
    return lib().orc_site_index(_i4(dims), x, y, z, t)


def moves(dims, bc):
    V = volume(dims)
    arrs = [np.zeros(V * 4, np.int32) for _ in range(4)]
    ptrs = [a.ctypes.data_as(C.POINTER(C.c_int)) for a in arrs]
    assert lib().orc_build_moves(_i4(dims), _i4(bc), *ptrs) == 0
    return arrs


def eta(dims):
    e = np.zeros(volume(dims), np.uint8)
    assert lib().orc_build_eta(_i4(dims), e.ctypes.data_as(C.POINTER(C.c_ubyte))) == 0
    return e


def eo_sites(dims, parity):
    out = np.zeros(volume(dims) // 2, np.int32)
    rc = lib().orc_eo_sites(_i4(dims), parity, out.ctypes.data_as(C.POINTER(C.c_int)))
    assert rc == 0
    return out


def gamma(which):
    idx, z4 = (C.c_int * 4)(), (C.c_int * 4)()
    assert lib().orc_gamma(which, idx, z4) == 0
    g = np.zeros((4, 4), complex)
    for r in range(4):
        g[r, idx[r]] = 1j ** z4[r]
    return g


def philox(ctr, key):
    out = (C.c_uint * 4)()
    lib().orc_philox((C.c_uint * 4)(*ctr), (C.c_uint * 2)(*key), out)
    return list(out)


def gauge_random(dims, eps, seed):
    U = np.zeros(volume(dims) * 72)
    assert lib().orc_gauge_random(_i4(dims), C.c_double(eps), C.c_uint(seed), _dp(U)) == 0
    return U


def gauge_unit(dims):
    U = np.zeros(volume(dims) * 72)
    assert lib().orc_gauge_unit(_i4(dims), _dp(U)) == 0
    return U


def wilson_gaussian(dims, seed):
    p = np.zeros(volume(dims) * 24)
    assert lib().orc_wilson_gaussian(_i4(dims), C.c_uint(seed), _dp(p)) == 0
    return p


def ks_gaussian(dims, seed):
    p = np.zeros(volume(dims) * 6)
    assert lib().orc_ks_gaussian(_i4(dims), C.c_uint(seed), _dp(p)) == 0
    return p


def su3_norm(m):
    m = np.ascontiguousarray(m, dtype=np.float64).copy()
    assert lib().orc_su3_norm(_dp(m), C.c_long(m.size // 18)) == 0
    return m


def wilson_op(dims, bc, kappa, U, psi, op=D, coeff_type=0, cre=1.0, cim=0.0):
    out = np.zeros_like(psi)
    rc = lib().orc_wilson_op(_i4(dims), _i4(bc), C.c_double(kappa), _dp(U), _dp(psi), _dp(out),
                             op, coeff_type, C.c_double(cre), C.c_double(cim))
    assert rc == 0
    return out


def wilson_op_f32(dims, bc, kappa, U, psi, op=D):
    U32 = np.ascontiguousarray(U, np.float32)
    p32 = np.ascontiguousarray(psi, np.float32)
    out = np.zeros_like(p32)
    rc = lib().orc_wilson_op_f32(_i4(dims), _i4(bc), C.c_float(kappa), _fp(U32), _fp(p32),
                                 _fp(out), op)
    assert rc == 0
    return out


def ks_op(dims, bc, two_am, U, chi, op=D, coeff_type=0, cre=1.0, cim=0.0):
    out = np.zeros_like(chi)
    rc = lib().orc_ks_op(_i4(dims), _i4(bc), C.c_double(two_am), _dp(U), _dp(chi), _dp(out),
                         op, coeff_type, C.c_double(cre), C.c_double(cim))
    assert rc == 0
    return out


def ks_op_f32(dims, bc, two_am, U, chi, op=D):
    U32 = np.ascontiguousarray(U, np.float32)
    c32 = np.ascontiguousarray(chi, np.float32)
    out = np.zeros_like(c32)
    rc = lib().orc_ks_op_f32(_i4(dims), _i4(bc), C.c_float(two_am), _fp(U32), _fp(c32), _fp(out), op)
    assert rc == 0
    return out


def gmres(kind, op, dims, bc, m, U, b, max_dim=12, restart=100, accuracy=1e-20, blength=1.0,
          stop_rule=0, start=None):
    o = make_op(kind, op, dims, bc, m)
    x = np.zeros_like(b)
    cyc, dev = C.c_int(), C.c_double()
    st = _dp(start) if start is not None else None
    rc = lib().orc_gmres(C.byref(o), _dp(U), _dp(b), _dp(x), st, max_dim, restart,
                         C.c_double(accuracy), C.c_double(blength), stop_rule,
                         C.byref(cyc), C.byref(dev))
    assert rc in (0, 1)
    return x, rc, cyc.value, dev.value


def bicgstab(kind, op, dims, bc, m, U, b, step_count=100, dev_check=10, retry=3,
             accuracy=1e-20, small_rho=1e-20, absolute=0):
    o = make_op(kind, op, dims, bc, m)
    x = np.zeros_like(b)
    it, dev = C.c_int(), C.c_double()
    rc = lib().orc_bicgstab(C.byref(o), _dp(U), _dp(b), _dp(x), None, step_count, dev_check,
                            retry, C.c_double(accuracy), C.c_double(small_rho), absolute,
                            C.byref(it), C.byref(dev))
    assert rc in (0, 1)
    return x, rc, it.value, dev.value


def msfom(kind, op, dims, bc, m, U, b, shifts, max_dim=30, restart=50, accuracy=1e-12,
          blength=1.0):
    o = make_op(kind, op, dims, bc, m)
    sh = np.ascontiguousarray(np.array([[complex(s).real, complex(s).imag] for s in shifts]).ravel())
    xs = np.zeros(len(shifts) * b.size)
    cyc, err = C.c_int(), C.c_double()
    rc = lib().orc_msfom(C.byref(o), _dp(U), _dp(b), _dp(xs), _dp(sh), len(shifts), max_dim,
                         restart, C.c_double(accuracy), C.c_double(blength), C.byref(cyc),
                         C.byref(err))
    assert rc in (0, 1)
    return xs.reshape(len(shifts), b.size), rc, cyc.value, err.value


def wilson_force(dims, bc, kappa, U, inv_d, inv_dddag, F=None):
    F = np.zeros(volume(dims) * 72) if F is None else F.copy()
    rc = lib().orc_wilson_force_kernel(_i4(dims), _i4(bc), C.c_double(kappa), _dp(U), _dp(inv_d),
                                       _dp(inv_dddag), _dp(F))
    assert rc == 0
    return F


def ks_force(dims, bc, U, phis, dphis, numer, F=None):
    F = np.zeros(volume(dims) * 72) if F is None else F.copy()
    phis = np.ascontiguousarray(phis, np.float64)
    dphis = np.ascontiguousarray(dphis, np.float64)
    num = np.ascontiguousarray(numer, np.float64)
    rc = lib().orc_ks_force_kernel(_i4(dims), _i4(bc), _dp(U), _dp(phis), _dp(dphis), _dp(num),
                                   len(numer), _dp(F))
    assert rc == 0
    return F


def time_wilson_d(dims, prec, reps, nthreads=0):
    lib().orc_time_wilson_d.restype = C.c_double
    return lib().orc_time_wilson_d(_i4(dims), prec, reps, nthreads)


def num_threads():
    return lib().orc_num_threads()


def dot(a, b):
    return np.vdot(a[0::2] + 1j * a[1::2], b[0::2] + 1j * b[1::2])


def as_complex(a):
    return a[0::2] + 1j * a[1::2]


def from_complex(z):
    out = np.empty(z.size * 2)
    out[0::2] = z.real
    out[1::2] = z.imag
    return out


def plaquette_sum(dims, U):
    out = C.c_double()
    assert lib().orc_plaquette_sum(_i4(dims), _dp(U), C.byref(out)) == 0
    return out.value


def gauge_force(dims, U, beta, F=None):
    F = np.zeros(volume(dims) * 72) if F is None else F.copy()
    assert lib().orc_gauge_force(_i4(dims), _dp(U), C.c_double(beta), _dp(F)) == 0
    return F


def su3_exp(M, a, prec):
    M = np.ascontiguousarray(M, np.float64)
    out = np.zeros(18)
    assert lib().orc_su3_exp(_dp(M), C.c_double(a), prec, _dp(out)) == 0
    return out


class Action(C.Structure):
    _fields_ = [("kind", C.c_int), ("beta", C.c_double), ("kappa_or_mass", C.c_double),
                ("bc", C.c_int * 4), ("phi", C.POINTER(C.c_double)), ("md", C.POINTER(C.c_double)),
                ("md_degree", C.c_int), ("accuracy", C.c_double)]


INT_LEAPFROG, INT_NESTED_FG, INT_NESTED_FG_INNER_LEAPFROG = 0, 1, 2


def hmc_evolve(dims, U, P, actions, integrator=INT_NESTED_FG, tau=1.0, steps=2, nested=2, exp_prec=0):
    """actions: list of dicts kind/beta/kappa_or_mass/bc/phi/md/accuracy.  Returns U, P, H0, H1."""
    U = np.ascontiguousarray(U, np.float64).copy()
    P = np.ascontiguousarray(P, np.float64).copy()
    keep = []
    arr = (Action * len(actions))()
    for i, a in enumerate(actions):
        arr[i].kind = a["kind"]
        arr[i].beta = a.get("beta", 0.0)
        arr[i].kappa_or_mass = a.get("kappa_or_mass", 0.0)
        for d in range(4):
            arr[i].bc[d] = a.get("bc", BC_FERMION)[d]
        if "phi" in a:
            phi = np.ascontiguousarray(a["phi"], np.float64)
            keep.append(phi)
            arr[i].phi = _dp(phi)
        if "md" in a:
            md = np.ascontiguousarray(a["md"], np.float64)
            keep.append(md)
            arr[i].md = _dp(md)
            arr[i].md_degree = (len(md) - 1) // 2
        arr[i].accuracy = a.get("accuracy", 1e-13)
    h0, h1 = C.c_double(), C.c_double()
    rc = lib().orc_hmc_evolve(_i4(dims), _dp(U), _dp(P), arr, len(actions), integrator,
                              C.c_double(tau), steps, nested, exp_prec, C.byref(h0), C.byref(h1))
    assert rc == 0
    return U, P, h0.value, h1.value


def random_momentum(rng, nlinks, sigma=0.5):
    """traceless anti-Hermitian links in the generator layout of the reference (8 real N(0, sigma^2))"""
    r = rng.normal(0.0, sigma, (nlinks, 8))
    i3, i32 = 1 / np.sqrt(3.0), 2 / np.sqrt(3.0)
    M = np.zeros((nlinks, 9), complex)
    M[:, 0] = 1j * (r[:, 7] * i3 + r[:, 2])
    M[:, 1] = r[:, 1] + 1j * r[:, 0]
    M[:, 2] = r[:, 4] + 1j * r[:, 3]
    M[:, 3] = -r[:, 1] + 1j * r[:, 0]
    M[:, 4] = 1j * (r[:, 7] * i3 - r[:, 2])
    M[:, 5] = r[:, 6] + 1j * r[:, 5]
    M[:, 6] = -r[:, 4] + 1j * r[:, 3]
    M[:, 7] = -r[:, 6] + 1j * r[:, 5]
    M[:, 8] = -1j * r[:, 7] * i32
    return from_complex(M.ravel())


def gcrodr(kind, op, dims, bc, m_or_kappa, U, b, m=10, k=8, restart=100, recalc=0, accuracy=1e-20,
           absolute=0, phase=3, urec=None, have_u=0):
    """returns x, status, cycles, deviation, urec, have_u (pass the last two back to recycle)"""
    o = make_op(kind, op, dims, bc, m_or_kappa)
    x = np.zeros_like(b)
    urec = np.zeros(k * b.size) if urec is None else urec.copy()
    hu = C.c_int(have_u)
    cyc, dev = C.c_int(), C.c_double()
    rc = lib().orc_gcrodr(C.byref(o), _dp(U), _dp(b), _dp(x), _dp(urec), C.byref(hu), m, k, restart, recalc,
                          C.c_double(accuracy), absolute, phase, C.byref(cyc), C.byref(dev))
    assert rc in (0, 1)
    return x, rc, cyc.value, dev.value, urec, hu.value


def fp16_links_12(U):
    """The links as FP16 12-real storage holds them, up to the last bit of the rounding: rows 0 and 1 rounded
    to FP16, row 2 rebuilt as the complex conjugate of their cross product."""
    m = U.reshape(-1, 9, 2)
    r = m[:, :6, :].astype(np.float16).astype(np.float64)
    c = r[..., 0] + 1j * r[..., 1]
    row2 = np.conj(np.stack([c[:, 1] * c[:, 5] - c[:, 2] * c[:, 4], c[:, 2] * c[:, 3] - c[:, 0] * c[:, 5],
                             c[:, 0] * c[:, 4] - c[:, 1] * c[:, 3]], axis=1))
    out = np.empty_like(m)
    out[:, :6, :] = r
    out[:, 6:, 0] = row2.real
    out[:, 6:, 1] = row2.imag
    return out.reshape(U.shape)
