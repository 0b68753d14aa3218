"""MI355X-native fermion-matrix hot path of CudaLatticeGauge.

The compute lives in ``libclg_hip.so`` (hand-written gfx950 HIP kernels behind the C ABI
in ``include/clg_hip.h``).  This module is a thin ctypes binding over that ABI used by the
tests and the benchmark; it never computes anything itself and raises if the shared
library has not been built.
"""
import ctypes as C
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libclg_hip.so")

# enums mirrored from include/clg_hip.h
WILSON, KS, GAUGE, GAUGE12 = 0, 1, 2, 3
F64, F32, F16 = 0, 1, 2
INNER_F16_VECTORS = 3  # inner_prec: FP16 storage of the inner Krylov vectors and links
FULL, EVEN, ODD = 0, 1, 2
EFO_F_D, EFO_F_Ddagger, EFO_F_DD, EFO_F_DDdagger = 0, 1, 2, 3
EFO_F_InverseD, EFO_F_InverseDdagger, EFO_F_InverseDD, EFO_F_InverseDDdagger = 4, 5, 6, 7
EOCT_None, EOCT_Minus, EOCT_Real, EOCT_Complex = 0, 1, 2, 3
SOLVER_GMRES, SOLVER_BICGSTAB = 0, 1
MS_FOM, MS_BICGSTAB, MS_GMRES, MS_NESTED = 0, 1, 2, 3
STOP_REFERENCE, STOP_TRUE_RESIDUAL = 0, 1
OK, NOT_CONVERGED = 0, 1
ACTION_GAUGE_PLAQUETTE, ACTION_FERMION_WILSON_NF2, ACTION_FERMION_KS = 0, 1, 2
DEFLATE_REV, DEFLATE_HEV, DEFLATE_SVD = 0, 1, 2
PHASE_START, PHASE_IN, PHASE_END, PHASE_ONCE = 0, 1, 2, 3
INT_LEAPFROG, INT_NESTED_FORCE_GRADIENT = 0, 1
HMC_MAX_ACTIONS = 8

EXPORTED = [
    "clg_init", "clg_finalize", "clg_synchronize", "clg_status_string", "clg_last_error",
    "clg_device_info", "clg_field_create", "clg_field_destroy", "clg_field_set_bc",
    "clg_field_upload", "clg_field_download", "clg_field_copy", "clg_field_zero",
    "clg_field_parity_view", "clg_field_host_doubles", "clg_field_bytes", "clg_file_read", "clg_file_write",
    "clg_md5", "clg_field_load_file", "clg_field_save_file", "clg_gauge_random",
    "clg_gauge_unit", "clg_field_gaussian", "clg_field_point_source", "clg_wilson_apply",
    "clg_wilson_hop", "clg_wilson_schur", "clg_ks_apply", "clg_ks_hop", "clg_ks_schur", "clg_axpy", "clg_scale", "clg_dot",
    "clg_norm2", "clg_solve", "clg_solver_params_default", "clg_time_wilson", "clg_time_ks",
    "clg_multishift_solve", "clg_rational_apply", "clg_wilson_force_kernel", "clg_ks_force_kernel",
    "clg_wilson_force", "clg_ks_force", "clg_wilson_pf_energy", "clg_ks_pf_energy",
    "clg_gauge_momentum_random", "clg_gauge_kinetic_energy", "clg_gauge_axpy", "clg_gauge_exp_mult",
    "clg_gauge_normalize", "clg_gauge_plaquette_energy", "clg_gauge_force", "clg_wilson_prepare_hmc",
    "clg_ks_prepare_hmc", "clg_hmc_params_default", "clg_hmc_evolve", "clg_hmc_trajectory",
    "clg_wilson_hop_batch", "clg_solve_batch", "clg_gcrodr_params_default", "clg_gcrodr_create",
    "clg_gcrodr_destroy", "clg_gcrodr_reset", "clg_gcrodr_solve", "clg_multidot", "clg_multidot_noalloc", "clg_multiaxpy", "clg_gram_matrix",
    "clg_vector_multiply_matrix", "clg_autotune_enable", "clg_autotune_force", "clg_autotune_reset",
    "clg_autotune_count", "clg_autotune_entry", "clg_graph_begin", "clg_graph_end", "clg_graph_launch",
    "clg_graph_nodes", "clg_graph_destroy", "clg_gauge_transform_random", "clg_gauge_transform_apply",
]


class CLGError(RuntimeError):
    pass


class SolverParams(C.Structure):
    _fields_ = [("solver", C.c_int), ("max_dim", C.c_int), ("restart", C.c_int),
                ("max_step", C.c_int), ("deviation_step", C.c_int), ("accuracy", C.c_double),
                ("absolute_accuracy", C.c_int), ("stop_rule", C.c_int), ("even_odd", C.c_int),
                ("inner_prec", C.c_int), ("inner_accuracy", C.c_double), ("max_outer", C.c_int),
                ("kind", C.c_int)]


class SolveReport(C.Structure):
    _fields_ = [("converged", C.c_int), ("cycles", C.c_int), ("outer", C.c_int),
                ("operator_applies", C.c_int), ("true_residual", C.c_double),
                ("last_deviation", C.c_double), ("seconds", C.c_double)]


class GcrodrParams(C.Structure):
    _fields_ = [("m_dim", C.c_int), ("k_dim", C.c_int), ("restart", C.c_int), ("recalculate_r", C.c_int),
                ("accuracy", C.c_double), ("absolute_accuracy", C.c_int), ("deflate", C.c_int), ("kind", C.c_int),
                ("even_odd", C.c_int)]


class Action(C.Structure):
    _fields_ = [("kind", C.c_int), ("beta", C.c_double), ("kappa_or_mass", C.c_double), ("phi", C.c_void_p),
                ("mc", C.POINTER(C.c_double)), ("mc_degree", C.c_int), ("md", C.POINTER(C.c_double)),
                ("md_degree", C.c_int), ("solver", SolverParams), ("energy_accuracy", C.c_double),
                ("ms_kind", C.c_int), ("ms_prec", C.c_int)]


class HmcParams(C.Structure):
    _fields_ = [("integrator", C.c_int), ("step_length", C.c_double), ("steps", C.c_int),
                ("nested_steps", C.c_int), ("inner_leapfrog", C.c_int), ("exp_precision", C.c_int),
                ("metropolis", C.c_int)]


class HmcReport(C.Structure):
    _fields_ = [("h_before", C.c_double), ("h_after", C.c_double), ("h_diff", C.c_double),
                ("kinetic_before", C.c_double), ("kinetic_after", C.c_double),
                ("action_before", C.c_double * HMC_MAX_ACTIONS), ("action_after", C.c_double * HMC_MAX_ACTIONS),
                ("accepted", C.c_int), ("force_evaluations", C.c_int), ("solver_iterations", C.c_int),
                ("unconverged_solves", C.c_int), ("seconds", C.c_double)]


_lib = None


def lib():
    """Load libclg_hip.so; raises if it is missing (there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CLGError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
        sig = {
            "clg_init": (C.c_int, [C.c_int, C.POINTER(vp)]),
            "clg_finalize": (C.c_int, [vp]),
            "clg_synchronize": (C.c_int, [vp]),
            "clg_status_string": (C.c_char_p, [C.c_int]),
            "clg_last_error": (C.c_char_p, []),
            "clg_device_info": (C.c_int, [vp, C.c_char_p, C.c_int]),
            "clg_field_create": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, ip, C.POINTER(vp)]),
            "clg_field_destroy": (C.c_int, [vp]),
            "clg_field_set_bc": (C.c_int, [vp, ip]),
            "clg_field_upload": (C.c_int, [vp, dp]),
            "clg_field_download": (C.c_int, [vp, dp]),
            "clg_field_copy": (C.c_int, [vp, vp]),
            "clg_field_zero": (C.c_int, [vp]),
            "clg_field_parity_view": (C.c_int, [vp, C.c_int, C.POINTER(vp)]),
            "clg_field_host_doubles": (C.c_long, [vp]),
            "clg_field_bytes": (C.c_long, [vp]),
            "clg_file_read": (C.c_int, [C.c_char_p, C.c_int, C.c_int, ip, dp]),
            "clg_file_write": (C.c_int, [C.c_char_p, C.c_int, C.c_int, ip, dp, C.c_char_p]),
            "clg_md5": (C.c_int, [C.c_void_p, C.c_ulong, C.c_char_p]),
            "clg_field_load_file": (C.c_int, [vp, C.c_char_p, C.c_int]),
            "clg_field_save_file": (C.c_int, [vp, C.c_char_p, C.c_int, C.c_char_p]),
            "clg_gauge_random": (C.c_int, [vp, C.c_double, C.c_uint]),
            "clg_gauge_unit": (C.c_int, [vp]),
            "clg_field_gaussian": (C.c_int, [vp, C.c_uint]),
            "clg_field_point_source": (C.c_int, [vp, ip, C.c_int, C.c_int]),
            "clg_wilson_apply": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double]),
            "clg_wilson_hop": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_double, C.c_double, vp, C.c_double, C.c_double]),
            "clg_wilson_schur": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int]),
            "clg_ks_apply": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double]),
            "clg_ks_hop": (C.c_int, [vp, vp, vp, vp, C.c_double, vp, C.c_double]),
            "clg_ks_schur": (C.c_int, [vp, vp, vp, vp, C.c_double]),
            "clg_axpy": (C.c_int, [vp, vp, C.c_double, C.c_double, vp]),
            "clg_scale": (C.c_int, [vp, vp, C.c_double, C.c_double]),
            "clg_dot": (C.c_int, [vp, vp, vp, dp]),
            "clg_norm2": (C.c_int, [vp, vp, dp]),
            "clg_solve": (C.c_int, [vp, C.POINTER(SolverParams), vp, vp, vp, C.c_double, C.c_int, vp,
                                    C.POINTER(SolveReport)]),
            "clg_solver_params_default": (C.c_int, [C.POINTER(SolverParams)]),
            "clg_time_wilson": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, dp]),
            "clg_time_ks": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, dp]),
            "clg_multishift_solve": (C.c_int, [vp, C.POINTER(SolverParams), C.c_int, C.c_int, C.POINTER(vp), dp,
                                               C.c_int, vp, vp, C.c_double, C.c_int, C.POINTER(SolveReport)]),
            "clg_rational_apply": (C.c_int, [vp, C.POINTER(SolverParams), C.c_int, C.c_int, vp, vp, vp, C.c_double,
                                             C.c_int, dp, C.c_int, C.POINTER(SolveReport)]),
            "clg_wilson_force_kernel": (C.c_int, [vp, vp, vp, vp, vp, C.c_double]),
            "clg_ks_force_kernel": (C.c_int, [vp, vp, vp, C.POINTER(vp), C.POINTER(vp), dp, C.c_int]),
            "clg_wilson_force": (C.c_int, [vp, C.POINTER(SolverParams), vp, vp, vp, C.c_double, vp, vp,
                                           C.POINTER(SolveReport)]),
            "clg_ks_force": (C.c_int, [vp, C.POINTER(SolverParams), C.c_int, C.c_int, vp, vp, vp, C.c_double, dp,
                                       C.c_int, C.POINTER(SolveReport)]),
            "clg_wilson_pf_energy": (C.c_int, [vp, C.POINTER(SolverParams), vp, vp, C.c_double, dp,
                                               C.POINTER(SolveReport)]),
            "clg_ks_pf_energy": (C.c_int, [vp, C.POINTER(SolverParams), C.c_int, C.c_int, vp, vp, C.c_double, dp,
                                           C.c_int, dp, C.POINTER(SolveReport)]),
            "clg_gauge_momentum_random": (C.c_int, [vp, C.c_uint, C.c_uint]),
            "clg_gauge_kinetic_energy": (C.c_int, [vp, vp, dp]),
            "clg_gauge_axpy": (C.c_int, [vp, vp, C.c_double, vp]),
            "clg_gauge_exp_mult": (C.c_int, [vp, vp, vp, C.c_double, C.c_int]),
            "clg_gauge_normalize": (C.c_int, [vp, vp]),
            "clg_gauge_transform_random": (C.c_int, [vp, C.c_double, C.c_uint]),
            "clg_gauge_transform_apply": (C.c_int, [vp, vp]),
            "clg_gauge_plaquette_energy": (C.c_int, [vp, vp, C.c_double, dp]),
            "clg_gauge_force": (C.c_int, [vp, vp, vp, C.c_double]),
            "clg_wilson_prepare_hmc": (C.c_int, [vp, vp, vp, C.c_double, C.c_uint]),
            "clg_ks_prepare_hmc": (C.c_int, [vp, C.POINTER(SolverParams), C.c_int, C.c_int, vp, vp, C.c_double, dp,
                                             C.c_int, C.c_uint, C.POINTER(SolveReport)]),
            "clg_hmc_params_default": (C.c_int, [C.POINTER(HmcParams)]),
            "clg_hmc_evolve": (C.c_int, [vp, vp, vp, C.POINTER(HmcParams), C.POINTER(Action), C.c_int,
                                         C.POINTER(HmcReport)]),
            "clg_hmc_trajectory": (C.c_int, [vp, vp, C.POINTER(HmcParams), C.POINTER(Action), C.c_int, C.c_ulonglong,
                                             C.POINTER(HmcReport)]),
            "clg_wilson_hop_batch": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.c_int, dp,
                                               C.POINTER(vp), C.c_double, C.c_int]),
            "clg_solve_batch": (C.c_int, [vp, C.POINTER(SolverParams), C.c_int, C.POINTER(vp), C.POINTER(vp),
                                          C.POINTER(vp), C.c_double, C.c_int, C.POINTER(SolveReport)]),
            "clg_gcrodr_params_default": (C.c_int, [C.POINTER(GcrodrParams)]),
            "clg_gcrodr_create": (C.c_int, [vp, C.POINTER(GcrodrParams), C.POINTER(vp)]),
            "clg_gcrodr_destroy": (C.c_int, [vp]),
            "clg_gcrodr_reset": (C.c_int, [vp]),
            "clg_gcrodr_solve": (C.c_int, [vp, vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, C.POINTER(SolveReport)]),
            "clg_multidot": (C.c_int, [vp, C.c_int, C.POINTER(vp), vp, dp]),
            "clg_multidot_noalloc": (C.c_int, [vp, C.c_int, C.POINTER(vp), vp, vp]),
            "clg_gram_matrix": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), dp]),
            "clg_multiaxpy": (C.c_int, [vp, vp, C.c_int, C.POINTER(vp), dp]),
            "clg_vector_multiply_matrix": (C.c_int, [vp, C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp), dp]),
            "clg_autotune_enable": (C.c_int, [C.c_int]),
            "clg_autotune_force": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int]),
            "clg_autotune_reset": (C.c_int, []),
            "clg_autotune_count": (C.c_int, []),
            "clg_autotune_entry": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
            "clg_graph_begin": (C.c_int, [vp]),
            "clg_graph_end": (C.c_int, [vp, C.POINTER(vp)]),
            "clg_graph_launch": (C.c_int, [vp]),
            "clg_graph_nodes": (C.c_int, [vp]),
            "clg_graph_destroy": (C.c_int, [vp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


# EFieldFileType (CField.h:26-35)
(FILE_BRIDGEPP_TXT, FILE_BRIDGEPP_BIN, FILE_CLG_BIN, FILE_CLG_BIN_COMPRESSED, FILE_CLG_BIN_FLOAT,
 FILE_CLG_BIN_DOUBLE) = range(6)


def file_read(path, file_type, kind, dims):
    """Host-only read of a full-lattice field file into the host AoS order (no device needed)."""
    n = int(np.prod(dims)) * {WILSON: 24, KS: 6, GAUGE: 72, GAUGE12: 72}[kind]
    a = np.empty(n, dtype=np.float64)
    check(lib().clg_file_read(os.fsencode(path), file_type, kind, _i4(dims), a.ctypes.data_as(C.POINTER(C.c_double))))
    return a


def file_write(path, file_type, kind, dims, host):
    """Host-only write of the host AoS order in a CLG binary format; returns the MD5 string."""
    a = np.ascontiguousarray(host, dtype=np.float64)
    n = int(np.prod(dims)) * {WILSON: 24, KS: 6, GAUGE: 72, GAUGE12: 72}[kind]
    if a.size != n:
        raise CLGError(f"file_write expects {n} doubles, got {a.size}")
    out = C.create_string_buffer(33)
    check(lib().clg_file_write(os.fsencode(path), file_type, kind, _i4(dims), a.ctypes.data_as(C.POINTER(C.c_double)),
                               out))
    return out.value.decode()


def md5(data):
    """CLGMD5Hash: upper-case hex MD5 of a bytes-like object."""
    b = bytes(data)
    out = C.create_string_buffer(33)
    check(lib().clg_md5(b, len(b), out))
    return out.value.decode()


def check(status, allow=(OK,)):
    if status not in allow:
        msg = lib().clg_last_error().decode()
        raise CLGError(f"{lib().clg_status_string(status).decode()}: {msg}")
    return status


def _i4(v):
    return (C.c_int * 4)(*[int(a) for a in v])


class Context:
    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().clg_init(device, C.byref(h)))
        self.h = h
        self._fields = weakref.WeakSet()
        self._graphs = weakref.WeakSet()
        self._solvers = weakref.WeakSet()

    def sync(self):
        check(lib().clg_synchronize(self.h))

    def info(self):
        buf = C.create_string_buffer(128)
        check(lib().clg_device_info(self.h, buf, 128))
        return buf.value.decode()

    def close(self):
        """Frees every field of this context (views first), then the context itself."""
        if self.h:
            for g in list(self._graphs):
                g.close()
            for sv in list(self._solvers):
                sv.close()
            fields = list(self._fields)
            for f in fields:
                if f._parent is not None:
                    f.free()
            for f in fields:
                f.free()
            lib().clg_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Field:
    """A device field (Wilson spinor, staggered vector or gauge links)."""

    def __init__(self, ctx, kind, prec, dims, subset=FULL, bc=None, _handle=None, _parent=None):
        self.ctx, self.kind, self.prec, self.subset = ctx, kind, prec, subset
        self.dims = tuple(int(d) for d in dims)
        self._parent = _parent
        if _handle is None:
            h = C.c_void_p()
            check(lib().clg_field_create(ctx.h, kind, prec, subset, _i4(dims), C.byref(h)))
            self.h = h
        else:
            self.h = _handle
        ctx._fields.add(self)
        if bc is not None:
            check(lib().clg_field_set_bc(self.h, _i4(bc)))

    def set_bc(self, bc):
        check(lib().clg_field_set_bc(self.h, _i4(bc)))

    def host_doubles(self):
        return lib().clg_field_host_doubles(self.h)

    def nbytes(self):
        return lib().clg_field_bytes(self.h)

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.size != self.host_doubles():
            raise CLGError(f"upload expects {self.host_doubles()} doubles, got {a.size}")
        check(lib().clg_field_upload(self.h, a.ctypes.data_as(C.POINTER(C.c_double))))

    def download(self):
        a = np.empty(self.host_doubles(), dtype=np.float64)
        check(lib().clg_field_download(self.h, a.ctypes.data_as(C.POINTER(C.c_double))))
        return a

    def load_file(self, path, file_type=None):
        """CField::InitialFieldWithFile (CFieldGaugeSU3.cu:1080-1185); Bridge++ gauge files are
        re-unitarised after loading."""
        check(lib().clg_field_load_file(self.h, os.fsencode(path), FILE_CLG_BIN if file_type is None else file_type))

    def save_file(self, path, file_type=None):
        """CField::SaveToFile (CField.cpp:32-72); returns the MD5 string of the bytes written."""
        out = C.create_string_buffer(33)
        check(lib().clg_field_save_file(self.h, os.fsencode(path), FILE_CLG_BIN if file_type is None else file_type,
                                        out))
        return out.value.decode()

    def copy_from(self, src):
        check(lib().clg_field_copy(self.h, src.h))

    def zero(self):
        check(lib().clg_field_zero(self.h))

    def gaussian(self, seed):
        check(lib().clg_field_gaussian(self.h, seed))

    def point_source(self, coord, spin, colour):
        check(lib().clg_field_point_source(self.h, _i4(coord), spin, colour))

    def gauge_random(self, eps, seed):
        check(lib().clg_gauge_random(self.h, eps, seed))

    def gauge_unit(self):
        check(lib().clg_gauge_unit(self.h))

    def parity(self, subset):
        h = C.c_void_p()
        check(lib().clg_field_parity_view(self.h, subset, C.byref(h)))
        return Field(self.ctx, self.kind, self.prec, self.dims, subset=subset, _handle=h, _parent=self)

    def free(self):
        if self.h and self.ctx.h:
            lib().clg_field_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def wilson_apply(ctx, out, inp, gauge, kappa, op=EFO_F_D, coeff_type=EOCT_None, cre=1.0, cim=0.0):
    check(lib().clg_wilson_apply(ctx.h, out.h, inp.h, gauge.h, kappa, op, coeff_type, cre, cim))


def wilson_hop(ctx, out, inp, gauge, dagger=False, a=1.0, x=None, b=0.0):
    a, b = complex(a), complex(b)
    check(lib().clg_wilson_hop(ctx.h, out.h, inp.h, gauge.h, int(dagger), a.real, a.imag,
                               x.h if x is not None else None, b.real, b.imag))


def wilson_schur(ctx, out, inp, gauge, kappa, op=EFO_F_D):
    check(lib().clg_wilson_schur(ctx.h, out.h, inp.h, gauge.h, kappa, op))


def ks_apply(ctx, out, inp, gauge, two_am, op=EFO_F_D, coeff_type=EOCT_None, cre=1.0, cim=0.0):
    check(lib().clg_ks_apply(ctx.h, out.h, inp.h, gauge.h, two_am, op, coeff_type, cre, cim))


def ks_hop(ctx, out, inp, gauge, a=1.0, x=None, b=0.0):
    check(lib().clg_ks_hop(ctx.h, out.h, inp.h, gauge.h, a, x.h if x is not None else None, b))


def ks_schur(ctx, out, inp, gauge, two_am):
    check(lib().clg_ks_schur(ctx.h, out.h, inp.h, gauge.h, two_am))


def axpy(ctx, y, a, x):
    a = complex(a)
    check(lib().clg_axpy(ctx.h, y.h, a.real, a.imag, x.h))


def scale(ctx, y, a):
    a = complex(a)
    check(lib().clg_scale(ctx.h, y.h, a.real, a.imag))


def dot(ctx, a, b):
    out = (C.c_double * 2)()
    check(lib().clg_dot(ctx.h, a.h, b.h, out))
    return complex(out[0], out[1])


def norm2(ctx, a):
    out = C.c_double()
    check(lib().clg_norm2(ctx.h, a.h, C.byref(out)))
    return out.value


def _handles(fields):
    return (C.c_void_p * len(fields))(*[f.h for f in fields])


def _cplx(values):
    v = np.asarray(values, dtype=np.complex128).ravel()
    return (C.c_double * (2 * len(v)))(*np.column_stack([v.real, v.imag]).ravel().tolist())


def multidot(ctx, V, w):
    """[<V_k, w>] for the fields V (one pass over w per 16 vectors, FP64 sums)."""
    out = (C.c_double * (2 * len(V)))()
    check(lib().clg_multidot(ctx.h, len(V), _handles(V), w.h, out))
    return np.array(out[0::2]) + 1j * np.array(out[1::2])


def gram_matrix(ctx, A, B):
    """[[<A_i, B_j>]] for FP64 fields (matrix cores, FP64 sums)."""
    out = (C.c_double * (2 * len(A) * len(B)))()
    check(lib().clg_gram_matrix(ctx.h, len(A), _handles(A), len(B), _handles(B), out))
    z = np.array(out[0::2]) + 1j * np.array(out[1::2])
    return z.reshape(len(A), len(B))


def multiaxpy(ctx, w, V, coeffs):
    """w += sum_k coeffs[k] V_k."""
    check(lib().clg_multiaxpy(ctx.h, w.h, len(V), _handles(V), _cplx(coeffs)))


def vector_multiply_matrix(ctx, res, left, matrix):
    """CFieldMatrixOperation::VectorMultiplyMatrix: with W = res + left, res_i <- sum_j W_j M[j, i]."""
    M = np.asarray(matrix, dtype=np.complex128)
    if M.shape != (len(res) + len(left), len(res)):
        raise CLGError(f"matrix must be {(len(res) + len(left), len(res))}, got {M.shape}")
    lh = _handles(left) if left else None
    check(lib().clg_vector_multiply_matrix(ctx.h, len(res), _handles(res), len(left), lh, _cplx(M)))


def autotune_enable(mode):
    check(lib().clg_autotune_enable(int(mode)))


def autotune_force(pin_mode=0, block=256, xcd=8, sites_per_thread=1):
    check(lib().clg_autotune_force(int(pin_mode), int(block), int(xcd), int(sites_per_thread)))


def autotune_reset():
    check(lib().clg_autotune_reset())


def autotune_entries():
    out = []
    buf = C.create_string_buffer(512)
    for i in range(lib().clg_autotune_count()):
        check(lib().clg_autotune_entry(i, buf, 512))
        out.append(buf.value.decode())
    return out


class Graph:
    """Launches recorded between Graph.record(ctx) and end(), replayed by launch()."""

    def __init__(self, ctx, handle):
        self.ctx, self.h = ctx, handle
        ctx._graphs.add(self)

    @staticmethod
    def begin(ctx):
        check(lib().clg_graph_begin(ctx.h))

    @staticmethod
    def end(ctx):
        h = C.c_void_p()
        check(lib().clg_graph_end(ctx.h, C.byref(h)))
        return Graph(ctx, h)

    def launch(self):
        check(lib().clg_graph_launch(self.h))

    @property
    def nodes(self):
        return lib().clg_graph_nodes(self.h)

    def close(self):
        if self.h and self.ctx.h:
            lib().clg_graph_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solver_params(**kw):
    p = SolverParams()
    check(lib().clg_solver_params_default(C.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise CLGError(f"unknown solver parameter {k}")
        setattr(p, k, v)
    return p


def solve(ctx, params, x, b, gauge, kappa_or_mass, op=EFO_F_InverseD, start=None):
    rep = SolveReport()
    st = lib().clg_solve(ctx.h, C.byref(params), x.h, b.h, gauge.h, kappa_or_mass, op,
                         start.h if start is not None else None, C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return st, rep


def time_wilson(ctx, out, inp, gauge, kappa, variant, reps):
    ms = C.c_double()
    check(lib().clg_time_wilson(ctx.h, out.h, inp.h, gauge.h, kappa, variant, reps, C.byref(ms)))
    return ms.value


def time_ks(ctx, out, inp, gauge, two_am, reps, variant=0):
    ms = C.c_double()
    check(lib().clg_time_ks(ctx.h, out.h, inp.h, gauge.h, two_am, variant, reps, C.byref(ms)))
    return ms.value


def multishift_solve(ctx, params, xs, shifts, b, gauge, kappa_or_mass, op, kind=MS_FOM, prec=F64):
    """xs[k] = (A + shifts[k])^-1 b for all k (A = op of the params.kind fermion)."""
    sh = np.ascontiguousarray([[complex(s).real, complex(s).imag] for s in shifts], dtype=np.float64).ravel()
    arr = (C.c_void_p * len(xs))(*[x.h.value for x in xs])
    rep = SolveReport()
    st = lib().clg_multishift_solve(ctx.h, C.byref(params), kind, prec, arr,
                                    sh.ctypes.data_as(C.POINTER(C.c_double)), len(xs), b.h, gauge.h,
                                    kappa_or_mass, op, C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return st, rep


def rational_apply(ctx, params, out, inp, gauge, kappa_or_mass, op, coeffs, kind=MS_FOM, prec=F64):
    """out = c inp + sum_k a_k (A + b_k)^-1 inp with coeffs = [c, a_1..a_n, b_1..b_n]."""
    cf = np.ascontiguousarray(coeffs, dtype=np.float64)
    degree = (cf.size - 1) // 2
    if cf.size != 2 * degree + 1:
        raise CLGError("coefficients must be [c, a_1..a_n, b_1..b_n]")
    rep = SolveReport()
    st = lib().clg_rational_apply(ctx.h, C.byref(params), kind, prec, out.h, inp.h, gauge.h, kappa_or_mass, op,
                                  cf.ctypes.data_as(C.POINTER(C.c_double)), degree, C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return st, rep


def _handles(fields):
    return (C.c_void_p * len(fields))(*[f.h.value for f in fields])


def wilson_force_kernel(ctx, force, gauge, inv_dddag, inv_d, kappa):
    check(lib().clg_wilson_force_kernel(ctx.h, force.h, gauge.h, inv_dddag.h, inv_d.h, kappa))


def ks_force_kernel(ctx, force, gauge, phis, dphis, numerators):
    num = np.ascontiguousarray(numerators, dtype=np.float64)
    check(lib().clg_ks_force_kernel(ctx.h, force.h, gauge.h, _handles(phis), _handles(dphis),
                                    num.ctypes.data_as(C.POINTER(C.c_double)), len(phis)))


def wilson_force(ctx, params, force, phi, gauge, kappa, start=None, x_out=None):
    rep = SolveReport()
    st = lib().clg_wilson_force(ctx.h, C.byref(params), force.h, phi.h, gauge.h, kappa,
                                start.h if start is not None else None, x_out.h if x_out is not None else None,
                                C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return st, rep


def ks_force(ctx, params, force, phi, gauge, two_am, coeffs, kind=MS_FOM, prec=F64):
    cf = np.ascontiguousarray(coeffs, dtype=np.float64)
    rep = SolveReport()
    st = lib().clg_ks_force(ctx.h, C.byref(params), kind, prec, force.h, phi.h, gauge.h, two_am,
                            cf.ctypes.data_as(C.POINTER(C.c_double)), (cf.size - 1) // 2, C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return st, rep


def wilson_pf_energy(ctx, params, phi, gauge, kappa):
    e, rep = C.c_double(), SolveReport()
    st = lib().clg_wilson_pf_energy(ctx.h, C.byref(params), phi.h, gauge.h, kappa, C.byref(e), C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return e.value


def ks_pf_energy(ctx, params, phi, gauge, two_am, coeffs, kind=MS_FOM, prec=F64):
    cf = np.ascontiguousarray(coeffs, dtype=np.float64)
    e, rep = C.c_double(), SolveReport()
    st = lib().clg_ks_pf_energy(ctx.h, C.byref(params), kind, prec, phi.h, gauge.h, two_am,
                                cf.ctypes.data_as(C.POINTER(C.c_double)), (cf.size - 1) // 2, C.byref(e),
                                C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return e.value


# ---------------------------------------------------------------- HMC
def momentum_random(P, seed, counter=0):
    check(lib().clg_gauge_momentum_random(P.h, seed, counter))


def kinetic_energy(ctx, P):
    e = C.c_double()
    check(lib().clg_gauge_kinetic_energy(ctx.h, P.h, C.byref(e)))
    return e.value


def gauge_axpy(ctx, y, a, x):
    check(lib().clg_gauge_axpy(ctx.h, y.h, a, x.h))


def exp_mult(ctx, P, gauge, a, exp_precision=0):
    check(lib().clg_gauge_exp_mult(ctx.h, P.h, gauge.h, a, exp_precision))


def gauge_normalize(ctx, gauge):
    check(lib().clg_gauge_normalize(ctx.h, gauge.h))


def gauge_transform_random(gt, eps, seed):
    """g(x) = reunitarise(exp(i eps H(x))) per site, held in the mu = 0 links of the FP64 gauge field gt"""
    check(lib().clg_gauge_transform_random(gt.h, eps, seed))


def gauge_transform_apply(gt, target):
    """U_mu(x) -> g(x) U_mu(x) g(x+mu)^dagger for a gauge field, psi(x) -> g(x) psi(x) for a fermion field"""
    check(lib().clg_gauge_transform_apply(gt.h, target.h))


def plaquette_energy(ctx, gauge, beta):
    e = C.c_double()
    check(lib().clg_gauge_plaquette_energy(ctx.h, gauge.h, beta, C.byref(e)))
    return e.value


def gauge_force(ctx, force, gauge, beta):
    check(lib().clg_gauge_force(ctx.h, force.h, gauge.h, beta))


def wilson_prepare_hmc(ctx, phi, gauge, kappa, seed):
    check(lib().clg_wilson_prepare_hmc(ctx.h, phi.h, gauge.h, kappa, seed))


def ks_prepare_hmc(ctx, params, phi, gauge, two_am, mc, seed, kind=MS_FOM, prec=F64):
    cf = np.ascontiguousarray(mc, dtype=np.float64)
    rep = SolveReport()
    st = lib().clg_ks_prepare_hmc(ctx.h, C.byref(params), kind, prec, phi.h, gauge.h, two_am,
                                  cf.ctypes.data_as(C.POINTER(C.c_double)), (cf.size - 1) // 2, seed, C.byref(rep))
    check(st, allow=(OK, NOT_CONVERGED))
    return st, rep


def hmc_params(**kw):
    p = HmcParams()
    check(lib().clg_hmc_params_default(C.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise CLGError(f"unknown HMC parameter {k}")
        setattr(p, k, v)
    return p


class _Actions:
    """ctypes array of clg_action built from dicts; keeps the coefficient arrays alive."""

    def __init__(self, actions):
        if not 1 <= len(actions) <= HMC_MAX_ACTIONS:
            raise CLGError(f"need 1..{HMC_MAX_ACTIONS} actions")
        self.keep = []
        self.arr = (Action * len(actions))()
        for i, a in enumerate(actions):
            t = self.arr[i]
            t.kind = a["kind"]
            t.beta = a.get("beta", 0.0)
            t.kappa_or_mass = a.get("kappa_or_mass", 0.0)
            t.phi = a["phi"].h.value if a.get("phi") is not None else None
            for key in ("mc", "md"):
                if a.get(key) is not None:
                    cf = np.ascontiguousarray(a[key], dtype=np.float64)
                    self.keep.append(cf)
                    setattr(t, key, cf.ctypes.data_as(C.POINTER(C.c_double)))
                    setattr(t, key + "_degree", (cf.size - 1) // 2)
            t.solver = a.get("solver", solver_params())
            t.energy_accuracy = a.get("energy_accuracy", 0.0)
            t.ms_kind = a.get("ms_kind", MS_FOM)
            t.ms_prec = a.get("ms_prec", F64)


def hmc_evolve(ctx, gauge, momentum, params, actions):
    acts = _Actions(actions)
    rep = HmcReport()
    check(lib().clg_hmc_evolve(ctx.h, gauge.h, momentum.h, C.byref(params), acts.arr, len(actions), C.byref(rep)))
    return rep


def hmc_trajectory(ctx, gauge, params, actions, seed):
    acts = _Actions(actions)
    rep = HmcReport()
    check(lib().clg_hmc_trajectory(ctx.h, gauge.h, C.byref(params), acts.arr, len(actions), seed, C.byref(rep)))
    return rep


# ---------------------------------------------------------------- batches
def wilson_hop_batch(ctx, outs, ins, gauges, a, dagger=False, xs=None, b=0.0, tile=0):
    n = len(outs)
    av = np.ascontiguousarray(a if np.ndim(a) else [a] * n, dtype=np.float64)
    check(lib().clg_wilson_hop_batch(ctx.h, n, _handles(outs), _handles(ins), _handles(gauges), int(dagger),
                                     av.ctypes.data_as(C.POINTER(C.c_double)),
                                     _handles(xs) if xs is not None else None, b, tile))


def solve_batch(ctx, params, xs, bs, gauges, kappa, op=EFO_F_InverseD):
    n = len(xs)
    reps = (SolveReport * n)()
    st = lib().clg_solve_batch(ctx.h, C.byref(params), n, _handles(xs), _handles(bs), _handles(gauges), kappa, op,
                               reps)
    check(st, allow=(OK, NOT_CONVERGED))
    return st, list(reps)


# ---------------------------------------------------------------- GCRO-DR
class Gcrodr:
    """CSLASolverGCRODR with its recycled subspace kept between solves."""

    def __init__(self, ctx, **kw):
        p = GcrodrParams()
        check(lib().clg_gcrodr_params_default(C.byref(p)))
        for k, v in kw.items():
            if not hasattr(p, k):
                raise CLGError(f"unknown GCRODR parameter {k}")
            setattr(p, k, v)
        self.ctx = ctx
        self.params = p
        self.h = C.c_void_p()
        check(lib().clg_gcrodr_create(ctx.h, C.byref(p), C.byref(self.h)))
        ctx._solvers.add(self)

    def solve(self, x, b, gauge, kappa_or_mass, op=EFO_F_D, phase=PHASE_ONCE, start=None):
        rep = SolveReport()
        st = lib().clg_gcrodr_solve(self.h, x.h, b.h, gauge.h, kappa_or_mass, op, phase,
                                    start.h if start is not None else None, C.byref(rep))
        check(st, allow=(OK, NOT_CONVERGED))
        return st, rep

    def reset(self):
        check(lib().clg_gcrodr_reset(self.h))

    def close(self):
        # after Context.close the library has destroyed the solver with its context
        if self.h and self.ctx.h:
            lib().clg_gcrodr_destroy(self.h)
        self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
