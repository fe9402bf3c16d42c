"""ctypes binding of libssf_gpu.so (the C ABI in include/ssf_gpu.h).

The library is the product: this module only moves numpy arrays across the C
ABI. Importing it raises if the HIP library has not been built.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import mesh

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libssf_gpu.so")

STORAGE_CR = 10
STORAGE_SCR = 11
SOLVER_CG = 20
SOLVER_BICG = 30
PREC_NONE = 0
PREC_DIAG = 1
PREC_ILDL = 10
FP64 = 0
MIXED = 1
KIND_HEX8_HEAT_CAP = 3

_P = ctypes.c_void_p
_L = ctypes.c_int64
_D = ctypes.c_double
_I = ctypes.c_int


class SolveParams(ctypes.Structure):
    _fields_ = [("linsolver", ctypes.c_int32), ("prec", ctypes.c_int32), ("ni", ctypes.c_int64),
                ("res", ctypes.c_double), ("zero", ctypes.c_double), ("iv", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("inner_res", ctypes.c_double),
                ("inner_ni", ctypes.c_int64), ("max_outer", ctypes.c_int64)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("ani", ctypes.c_int64), ("ares", ctypes.c_double), ("status", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("iterations", ctypes.c_int64), ("outer", ctypes.c_int64),
                ("rel_residual", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "pad"}


class NrParams(ctypes.Structure):
    _fields_ = [("ninr", ctypes.c_int64), ("niilnr", ctypes.c_int64), ("li", ctypes.c_int64),
                ("errnr", ctypes.c_double), ("incrnr", ctypes.c_double),
                ("minincrnr", ctypes.c_double), ("maxincrnr", ctypes.c_double),
                ("ilambda", ctypes.c_double), ("lambdar", ctypes.c_double),
                ("rnormtnr", ctypes.c_int32), ("check_lambdar", ctypes.c_int32),
                ("check_div", ctypes.c_int32), ("pad", ctypes.c_int32)]


class NrInfo(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double), ("norf", ctypes.c_double),
                ("increments", ctypes.c_int64), ("solves", ctypes.c_int64),
                ("rejected", ctypes.c_int64), ("inner", ctypes.c_int64),
                ("status", ctypes.c_int32), ("pad", ctypes.c_int32)]


SIGNATURES = {
    "ssf_last_error": (ctypes.c_char_p, []),
    "ssf_ctx_create": (_I, [_I, ctypes.POINTER(_P)]),
    "ssf_ctx_destroy": (_I, [_P]),
    "ssf_ctx_sync": (_I, [_P]),
    "ssf_ctx_stream": (_P, [_P]),
    "ssf_matrix_create_elem": (_I, [_P, _I, _L, _L, _L, _P, ctypes.POINTER(_P)]),
    "ssf_matrix_create_elem_device": (_I, [_P, _L, _L, _L, _P, ctypes.POINTER(_P)]),
    "ssf_matrix_create_csr": (_I, [_P, _I, _L, _P, _P, ctypes.POINTER(_P)]),
    "ssf_matrix_destroy": (_I, [_P]),
    "ssf_matrix_sizes": (_I, [_P, _P, _P, _P]),
    "ssf_matrix_pattern": (_I, [_P, _P, _P]),
    "ssf_matrix_zero": (_I, [_P]),
    "ssf_matrix_set_values": (_I, [_P, _P]),
    "ssf_matrix_get_values": (_I, [_P, _P]),
    "ssf_matrix_localize_batch": (_I, [_P, _L, _L, _P, _P]),
    "ssf_matrix_assemble": (_I, [_P, _I, _L, _P, _L, _P, _P, _P, _P, _P, _P, _L, _P]),
    "ssf_gmxv": (_I, [_P, _P, _P]),
    "ssf_gmxv_dev": (_I, [_P, _P, _P]),
    "ssf_solve": (_I, [_P, ctypes.POINTER(SolveParams), _P, _P, ctypes.POINTER(SolveInfo)]),
    "ssf_solve_dev": (_I, [_P, ctypes.POINTER(SolveParams), _P, _P, ctypes.POINTER(SolveInfo)]),
    "ssf_solve_dev_noalloc": (_I, [_P, ctypes.POINTER(SolveParams), _P, _P,
                                   ctypes.POINTER(SolveInfo)]),
    "ssf_spmv_bench": (_I, [_P, _I, _I, ctypes.POINTER(_D)]),
    "ssf_spmv_bytes": (_I, [_P, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "ssf_copy_bench": (_I, [_P, _L, _I, ctypes.POINTER(_D)]),
    "ssf_ctx_set_profiling": (_I, [_P, _I]),
    "ssf_ctx_kernel_time": (_I, [_P, _I, ctypes.POINTER(_D), ctypes.POINTER(_L)]),
    "ssf_matrix_autotune": (_I, [_P, _I, _I, _P]),
    "ssf_matrix_set_tune": (_I, [_P, _I, _P]),
    "ssf_batch_create": (_I, [_P, _L, _P, ctypes.c_int32, ctypes.POINTER(_P)]),
    "ssf_batch_destroy": (_I, [_P]),
    "ssf_batch_sizes": (_I, [_P, _P, _P, _P]),
    "ssf_batch_solve": (_I, [_P, ctypes.POINTER(SolveParams), _P, _P, _P]),
    "ssf_batch_solve_dev": (_I, [_P, ctypes.POINTER(SolveParams), _P, _P, _P]),
    "ssf_batch_autotune": (_I, [_P, _I, _I, _P]),
    "ssf_server_create": (_I, [_I, _L, ctypes.POINTER(SolveParams), ctypes.c_int32, ctypes.c_int32,
                               ctypes.c_int32, ctypes.POINTER(_P)]),
    "ssf_server_destroy": (_I, [_P]),
    "ssf_server_submit": (_I, [_P, _I, _L, _L, _P, _L, _P, _P, _P, _P, _P, _P, _L, _P,
                               ctypes.POINTER(_L)]),
    "ssf_server_wait": (_I, [_P, _L, _P, ctypes.POINTER(SolveInfo), _P]),
    "ssf_server_stats": (_I, [_P, _P]),
    "ssf_matrix_scale": (_I, [_P, _D]),
    "ssf_matrix_add": (_I, [_P, _D, _P]),
    "ssf_np_step_linear": (_I, [_P, _P, _P, _D, _D, ctypes.POINTER(SolveParams), _P, _P, _P,
                                ctypes.POINTER(SolveInfo)]),
    "ssf_cuthill_mckee": (_I, [_P, _L, _L, _L, _P, _I, _P]),
    "ssf_gencodnum": (_I, [_L, _L, _P, _P, ctypes.POINTER(_L)]),
    "ssf_give_code_numbers": (_I, [_L, _L, _P, _L, _L, _P, _P]),
    "ssf_newton_raphson": (_I, [_P, _P, ctypes.POINTER(NrParams), ctypes.POINTER(SolveParams),
                                _P, _L, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.POINTER(NrInfo)]),
    "ssf_matrix_ildl_factor": (_I, [_P, _P]),
    "ssf_ildl_apply_dev": (_I, [_P, _P, _P]),
    "ssf_ildl_bench": (_I, [_P, _I, _I, ctypes.POINTER(_L), ctypes.POINTER(_L),
                            ctypes.POINTER(_D)]),
    "ssf_np_step_dform": (_I, [_P, _P, _P, _D, _D, ctypes.POINTER(SolveParams), _P, _P, _P,
                               ctypes.POINTER(SolveInfo)]),
    "ssf_elements_create": (_I, [_P, _I, _L, _L, _P, _L, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "ssf_elements_destroy": (_I, [_P]),
    "ssf_elements_sizes": (_I, [_P, _P, _P, _P]),
    "ssf_elements_results": (_I, [_P, _P, _P, _L, _P, _P, _P]),
    "ssf_elements_results_dev": (_I, [_P, _P, _P, _L, _P, _P, _P]),
    "ssf_elements_volume_load": (_I, [_P, _P, _P]),
    "ssf_elements_surface_load": (_I, [_P, _L, _P, _P, _P, _P]),
    "ssf_elements_surface_load_dev": (_I, [_P, _L, _P, _P, _P, _P]),
    "ssf_elements_volume_load_dev": (_I, [_P, _P, _P]),
    "ssf_elements_bench": (_I, [_P, _P, _P, _L, _P, _P, _P, _I, _P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python ssf_amd/build.py)")
        # One HIP runtime per process: torch ships its own libamdhip64. Loaded
        # first, it is the runtime this library binds to (same soname), and torch
        # tensors and this library's calls share one device context. Loaded
        # after this library, a second runtime would start and torch would see
        # no GPU.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        l = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


class SsfError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        msg = lib().ssf_last_error().decode(errors="replace")
        raise SsfError(f"ssf error {rc}: {msg}")


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    def __init__(self, device: int = 0):
        self.h = _P()
        _check(lib().ssf_ctx_create(device, ctypes.byref(self.h)))

    def sync(self):
        _check(lib().ssf_ctx_sync(self.h))

    def set_profiling(self, on: bool):
        _check(lib().ssf_ctx_set_profiling(self.h, 1 if on else 0))

    def kernel_time(self, which: int):
        """(mean us, samples) of the FP64 (0) / FP32 (1) SpMV sampled inside solves."""
        us, n = _D(), _L()
        _check(lib().ssf_ctx_kernel_time(self.h, which, ctypes.byref(us), ctypes.byref(n)))
        return us.value, n.value

    def copy_bench(self, nbytes: int, reps: int = 20) -> float:
        """Mean time (us) of a device copy of nbytes (2*nbytes of HBM traffic)."""
        us = _D()
        _check(lib().ssf_copy_bench(self.h, nbytes, reps, ctypes.byref(us)))
        return us.value

    def close(self):
        """Destroy the context; refused (and kept) while matrices on it are alive."""
        if self.h and lib().ssf_ctx_destroy(self.h) == 0:
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Matrix:
    """A gmatrix whose symcomprow / comprow storage lives on the GPU."""

    def __init__(self, ctx: Context, handle):
        self.ctx = ctx
        self.h = handle
        n, negm, nnz = _L(), _L(), _L()
        _check(lib().ssf_matrix_sizes(self.h, ctypes.byref(n), ctypes.byref(negm), ctypes.byref(nnz)))
        self.n, self.negm, self.nnz = n.value, negm.value, nnz.value

    @classmethod
    def from_elements(cls, ctx: Context, storage: int, n: int, cn: np.ndarray) -> "Matrix":
        cn = _i64(cn)
        h = _P()
        _check(lib().ssf_matrix_create_elem(ctx.h, storage, n, cn.shape[0], cn.shape[1], _p(cn),
                                            ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_elements_device(cls, ctx: Context, n: int, cn: np.ndarray) -> "Matrix":
        """symcomprow pattern built on the device; SELL-only matrix (FP64 solves)."""
        cn = _i64(cn)
        h = _P()
        _check(lib().ssf_matrix_create_elem_device(ctx.h, n, cn.shape[0], cn.shape[1], _p(cn),
                                                   ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_csr(cls, ctx: Context, storage: int, adr: np.ndarray, ci: np.ndarray) -> "Matrix":
        adr, ci = _i64(adr), _i64(ci)
        h = _P()
        _check(lib().ssf_matrix_create_csr(ctx.h, storage, adr.shape[0] - 1, _p(adr), _p(ci),
                                           ctypes.byref(h)))
        return cls(ctx, h)

    def close(self):
        if self.h:
            lib().ssf_matrix_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pattern(self):
        adr = np.zeros(self.n + 1, np.int64)
        ci = np.zeros(self.negm, np.int64)
        _check(lib().ssf_matrix_pattern(self.h, _p(adr), _p(ci)))
        return adr, ci

    def zero(self):
        _check(lib().ssf_matrix_zero(self.h))

    def set_values(self, a):
        a = _f64(a)
        assert a.shape[0] == self.negm
        _check(lib().ssf_matrix_set_values(self.h, _p(a)))

    def get_values(self):
        a = np.zeros(self.negm)
        _check(lib().ssf_matrix_get_values(self.h, _p(a)))
        return a

    def localize_batch(self, cn, lm):
        cn, lm = _i64(cn), _f64(lm)
        ne, nc = cn.shape
        _check(lib().ssf_matrix_localize_batch(self.h, ne, nc, _p(cn), _p(lm)))

    def assemble(self, prob: mesh.Problem, lift: bool = True):
        """Element kernels + assembly by entry; returns the Dirichlet lift (or None)."""
        cn = _i64(prob.elem_codes())
        conn = _i64(prob.conn)
        xyz = _f64(prob.coords)
        e = _f64(prob.e)
        nu = _f64(prob.nu) if prob.kind != mesh.KIND_HEX8_HEAT else None
        th = _f64(prob.thick) if prob.kind == mesh.KIND_QUAD4_PSTRESS else None
        pv = _f64(prob.pv) if prob.pv.size else None
        out = np.zeros(self.n) if (lift and pv is not None) else None
        _check(lib().ssf_matrix_assemble(self.h, prob.kind, xyz.shape[0], _p(xyz), conn.shape[0],
                                         _p(conn), _p(cn), _p(e), _p(nu), _p(th), _p(pv),
                                         0 if pv is None else pv.shape[0], _p(out)))
        return out

    def assemble_capacity(self, prob: mesh.Problem, c, rho=None):
        """Heat capacity matrix of a hex8 heat mesh (specific heat c[ne], density
        rho[ne] or None = 1) added into the values."""
        cn = _i64(prob.elem_codes())
        conn = _i64(prob.conn)
        xyz = _f64(prob.coords)
        c = _f64(c)
        rho = None if rho is None else _f64(rho)
        _check(lib().ssf_matrix_assemble(self.h, KIND_HEX8_HEAT_CAP, xyz.shape[0], _p(xyz),
                                         conn.shape[0], _p(conn), _p(cn), _p(c), _p(rho), None,
                                         None, 0, None))

    def scale(self, c: float):
        """gmatrix::scalgm: values *= c"""
        _check(lib().ssf_matrix_scale(self.h, float(c)))

    def add(self, c: float, other: "Matrix"):
        """gmatrix::addgm: values += c * other.values (same pattern)"""
        _check(lib().ssf_matrix_add(self.h, float(c), other.h))

    def gmxv(self, a):
        a = _f64(a)
        b = np.zeros(self.n)
        _check(lib().ssf_gmxv(self.h, _p(a), _p(b)))
        return b

    def solve(self, rhs, ni=100000, res=1e-20, prec=PREC_DIAG, precision=FP64, zero=1e-30,
              x0=None, inner_res=1e-4, inner_ni=0, max_outer=0, linsolver=SOLVER_CG):
        """gmatrix::solve_system on host vectors: cg (symcomprow / comprow) or bicg
        (comprow, no preconditioner)."""
        rhs = _f64(rhs)
        x = np.zeros(self.n) if x0 is None else _f64(x0).copy()
        p = SolveParams(linsolver, prec, ni, res, zero, 0 if x0 is None else 1, precision,
                        inner_res, inner_ni, max_outer)
        info = SolveInfo(-1, -1.0, 0, 0, 0, 0, 0.0)
        _check(lib().ssf_solve(self.h, ctypes.byref(p), _p(rhs), _p(x), ctypes.byref(info)))
        return x, info

    def ildl_factor(self):
        """symcomprow::incomplete_ldl of the current values (weight 0), to the host."""
        out = np.zeros(self.negm)
        _check(lib().ssf_matrix_ildl_factor(self.h, _p(out)))
        return out

    def ildl_apply_dev(self, y_ptr, x_ptr):
        """prec_ildl_scr on device vectors: x = (L D L^T)^-1 y."""
        _check(lib().ssf_ildl_apply_dev(self.h, ctypes.c_void_p(y_ptr), ctypes.c_void_p(x_ptr)))

    def ildl_bench(self, reps=5, graph=True):
        """(forward levels, backward levels, mean usec per application)."""
        nf, nb, us = _L(), _L(), _D()
        _check(lib().ssf_ildl_bench(self.h, reps, 1 if graph else 0, ctypes.byref(nf),
                                    ctypes.byref(nb), ctypes.byref(us)))
        return nf.value, nb.value, us.value

    def solve_dev(self, rhs_ptr, x_ptr, ni=100000, res=1e-20, prec=PREC_DIAG, precision=FP64,
                  zero=1e-30, iv=0, inner_res=1e-4, inner_ni=0, max_outer=0,
                  linsolver=SOLVER_CG):
        p = SolveParams(linsolver, prec, ni, res, zero, iv, precision, inner_res, inner_ni,
                        max_outer)
        info = SolveInfo(-1, -1.0, 0, 0, 0, 0, 0.0)
        _check(lib().ssf_solve_dev(self.h, ctypes.byref(p), ctypes.c_void_p(rhs_ptr),
                                   ctypes.c_void_p(x_ptr), ctypes.byref(info)))
        return info

    def solve_dev_noalloc(self, rhs_ptr, x_ptr, ni=100000, res=1e-20, prec=PREC_DIAG, zero=1e-30,
                          iv=0):
        """ssf_solve_dev_noalloc: the FP64 cg solve with the library's static scratch
        array instead of any allocation (n <= 2 500 000)."""
        p = SolveParams(SOLVER_CG, prec, ni, res, zero, iv, FP64, 0.0, 0, 0)
        info = SolveInfo(-1, -1.0, 0, 0, 0, 0, 0.0)
        _check(lib().ssf_solve_dev_noalloc(self.h, ctypes.byref(p), ctypes.c_void_p(rhs_ptr),
                                           ctypes.c_void_p(x_ptr), ctypes.byref(info)))
        return info

    def spmv_bench(self, precision=FP64, reps=20):
        us = _D()
        _check(lib().ssf_spmv_bench(self.h, precision, reps, ctypes.byref(us)))
        return us.value

    def spmv_bytes(self, precision=FP64):
        a, b = _D(), _D()
        _check(lib().ssf_spmv_bytes(self.h, precision, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def autotune(self, precision=FP64, reps=5):
        """Returns the chosen SpMV tiling (format, lanes, blocks, idx16, map)."""
        out = np.zeros(5, np.int32)
        _check(lib().ssf_matrix_autotune(self.h, precision, reps, _p(out)))
        return tuple(int(v) for v in out)

    def set_tune(self, precision, tune):
        """Pins the SpMV tiling (format, lanes, blocks, idx16[, map]); map defaults
        to 1 (XCD interleave of the SELL slices)."""
        t = list(tune)
        if len(t) == 4:
            t.append(1)
        t = np.ascontiguousarray(t, dtype=np.int32)
        _check(lib().ssf_matrix_set_tune(self.h, precision, _p(t)))


def np_step_linear(K: Matrix, C: Matrix, A: Matrix, alpha: float, dt: float, rhs_ptr, lhs_ptr,
                   tdlhs_ptr, ni=100000, res=1e-20, precision=FP64, zero=1e-30, inner_ni=0,
                   form="v", inner_res=1e-4):
    """One generalized-trapezoidal step on device vectors: lhs (nodal values) and
    tdlhs (their rates) are updated; returns the solve info. form "v" is TRFEL's
    one_step_linear (npsolvert.cpp:367), "d" the step of linear_nonstat_solv_dform
    (npsolvert.cpp:1071)."""
    fn = {"v": "ssf_np_step_linear", "d": "ssf_np_step_dform"}[form]
    p = SolveParams(SOLVER_CG, PREC_DIAG, ni, res, zero, 1, precision, inner_res, inner_ni, 0)
    info = SolveInfo(-1, -1.0, 0, 0, 0, 0, 0.0)
    _check(getattr(lib(), fn)(K.h, C.h, A.h, float(alpha), float(dt), ctypes.byref(p),
                              rhs_ptr, lhs_ptr, tdlhs_ptr, ctypes.byref(info)))
    return info


def gencodnum(flags: np.ndarray, ordering=None) -> tuple[np.ndarray, int]:
    """gtopology::gencodnum on the host through the C ABI: node flags (nn, ndofn)
    (0 fixed, 1 free, k > 1 coupled group, < 0 prescribed) -> code numbers and n."""
    dof = np.array(flags, dtype=np.int64, order="C", copy=True)
    if dof.ndim == 1:
        dof = dof[:, None]
    nd = _L()
    o = None if ordering is None else _i64(ordering)
    _check(lib().ssf_gencodnum(dof.shape[0], dof.shape[1], _p(dof), _p(o), ctypes.byref(nd)))
    return dof, int(nd.value)


def give_code_numbers(conn: np.ndarray, dof: np.ndarray) -> np.ndarray:
    """gtopology::give_code_numbers for every element: (ne, nne * ndofn)."""
    conn = _i64(conn)
    dof = _i64(dof)
    if dof.ndim == 1:
        dof = dof[:, None]
    ne, nne = conn.shape
    cn = np.empty((ne, nne * dof.shape[1]), dtype=np.int64)
    _check(lib().ssf_give_code_numbers(ne, nne, _p(conn), dof.shape[0], dof.shape[1], _p(dof),
                                       _p(cn)))
    return cn


def cuthill_mckee(ctx: Context, nn: int, conn: np.ndarray, reverse=False) -> np.ndarray:
    """gtopology::cuthill_mckee_renumb on the device: ordering[k] = the node numbered k."""
    conn = _i64(conn)
    out = np.zeros(nn, np.int64)
    _check(lib().ssf_cuthill_mckee(ctx.h, nn, conn.shape[0], conn.shape[1], _p(conn),
                                   1 if reverse else 0, _p(out)))
    return out


def newton_raphson(K: Matrix, el: "Elements", fp_ptr, fa_ptr, ra_ptr, fc_ptr, ninr=50, niilnr=10,
                   errnr=1e-6, incrnr=0.1, minincrnr=1e-3, maxincrnr=1.0, ilambda=0.0, li=0,
                   normt=1, check_lambdar=False, lambdar=1.0, ni=20000, res=1e-24,
                   prec=PREC_DIAG, pv_ptr=None, npv=0):
    """MEFEL's gnewton_raphson on device vectors (fp, fc read; fa, ra updated).
    Returns (info, hist) with hist = dict(lam, dlam, nsol, acc) over the increments run."""
    p = NrParams(ninr, niilnr, li, errnr, incrnr, minincrnr, maxincrnr, ilambda, lambdar, normt,
                 1 if check_lambdar else 0, 0, 0)
    sp = SolveParams(SOLVER_CG, prec, ni, res, 1e-30, 0, FP64, 0.0, 0, 0)
    m = ninr - li
    lam, dlam = np.zeros(m), np.zeros(m)
    nsol, acc = np.zeros(m, np.int64), np.zeros(m, np.int32)
    info = NrInfo()
    _check(lib().ssf_newton_raphson(K.h, el.h, ctypes.byref(p), ctypes.byref(sp),
                                    ctypes.c_void_p(pv_ptr), npv, ctypes.c_void_p(fc_ptr),
                                    ctypes.c_void_p(fp_ptr), ctypes.c_void_p(fa_ptr),
                                    ctypes.c_void_p(ra_ptr), _p(lam), _p(dlam), _p(nsol), _p(acc),
                                    ctypes.byref(info)))
    k = int(info.increments)
    return info, {"lam": lam[:k], "dlam": dlam[:k], "nsol": nsol[:k], "acc": acc[:k]}


class Batch:
    """Independent systems x nrhs right-hand sides solved together on the GPU.

    Vectors are interleaved: entry (system s, row i, rhs r) sits at
    (row0[s] + i) * nrhs + r.
    """

    def __init__(self, ctx: Context, mats, nrhs: int = 1):
        self.ctx = ctx
        self.nsys = len(mats)
        self.nrhs = nrhs
        hs = (_P * self.nsys)(*[m.h for m in mats])
        self.h = _P()
        _check(lib().ssf_batch_create(ctx.h, self.nsys, hs, nrhs, ctypes.byref(self.h)))
        nrows, nnz = _L(), _L()
        self.row0 = np.zeros(self.nsys + 1, np.int64)
        _check(lib().ssf_batch_sizes(self.h, ctypes.byref(nrows), ctypes.byref(nnz),
                                     _p(self.row0)))
        self.nrows, self.nnz = nrows.value, nnz.value

    def close(self):
        if self.h:
            lib().ssf_batch_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pack(self, rhs_list):
        """[nsys] arrays of shape (nrhs, n_s) -> interleaved vector."""
        v = np.zeros((self.nrows, self.nrhs))
        for s, r in enumerate(rhs_list):
            r = np.asarray(r, dtype=np.float64).reshape(self.nrhs, -1)
            v[self.row0[s]:self.row0[s + 1], :] = r.T
        return v.reshape(-1)

    def unpack(self, v):
        v = np.asarray(v).reshape(self.nrows, self.nrhs)
        return [v[self.row0[s]:self.row0[s + 1], :].T.copy() for s in range(self.nsys)]

    def _params(self, ni, res, precision, zero, iv, inner_res, inner_ni):
        return SolveParams(SOLVER_CG, PREC_DIAG, ni, res, zero, iv, precision, inner_res,
                           inner_ni, 0)

    def _infos(self):
        return (SolveInfo * (self.nsys * self.nrhs))()

    def solve(self, rhs_list, ni=100000, res=1e-20, precision=FP64, zero=1e-30, x0=None,
              inner_res=1e-4, inner_ni=0):
        """Returns ([nsys] arrays (nrhs, n_s), infos[nsys * nrhs])."""
        y = self.pack(rhs_list)
        x = np.zeros_like(y) if x0 is None else self.pack(x0)
        p = self._params(ni, res, precision, zero, 0 if x0 is None else 1, inner_res, inner_ni)
        infos = self._infos()
        _check(lib().ssf_batch_solve(self.h, ctypes.byref(p), _p(y), _p(x), infos))
        return self.unpack(x), list(infos)

    def ildl_factor(self):
        """symcomprow::incomplete_ldl of the current values (weight 0), to the host."""
        out = np.zeros(self.negm)
        _check(lib().ssf_matrix_ildl_factor(self.h, _p(out)))
        return out

    def ildl_apply_dev(self, y_ptr, x_ptr):
        """prec_ildl_scr on device vectors: x = (L D L^T)^-1 y."""
        _check(lib().ssf_ildl_apply_dev(self.h, ctypes.c_void_p(y_ptr), ctypes.c_void_p(x_ptr)))

    def ildl_bench(self, reps=5, graph=True):
        """(forward levels, backward levels, mean usec per application)."""
        nf, nb, us = _L(), _L(), _D()
        _check(lib().ssf_ildl_bench(self.h, reps, 1 if graph else 0, ctypes.byref(nf),
                                    ctypes.byref(nb), ctypes.byref(us)))
        return nf.value, nb.value, us.value

    def solve_dev(self, rhs_ptr, x_ptr, ni=100000, res=1e-20, precision=FP64, zero=1e-30, iv=0,
                  inner_res=1e-4, inner_ni=0):
        p = self._params(ni, res, precision, zero, iv, inner_res, inner_ni)
        infos = self._infos()
        _check(lib().ssf_batch_solve_dev(self.h, ctypes.byref(p), ctypes.c_void_p(rhs_ptr),
                                         ctypes.c_void_p(x_ptr), infos))
        return list(infos)

    def autotune(self, precision=FP64, reps=5):
        """Returns the chosen SpMM (format, lanes): 2 = SELL-64, 0 = CSR row-vector."""
        out = np.zeros(2, np.int32)
        _check(lib().ssf_batch_autotune(self.h, precision, reps, _p(out)))
        return int(out[0]), int(out[1])


class Server:
    """Request loop on one GPU: element-mesh solves admitted under an HBM budget
    and solved with continuous batching (Jacobi-PCG per request: the FP64
    recurrence, or precision=MIXED: FP32 recurrences refreshed in FP64 every chunk)."""

    def __init__(self, device=0, budget_bytes=200 << 30, ni=100000, res=1e-20, zero=1e-30,
                 chunk=32, max_slots=1024, prep_threads=2, precision=FP64, inner_res=1e-4):
        p = SolveParams(SOLVER_CG, PREC_DIAG, ni, res, zero, 0, precision, inner_res, 0, 0)
        self.h = _P()
        _check(lib().ssf_server_create(device, int(budget_bytes), ctypes.byref(p), chunk,
                                       max_slots, prep_threads, ctypes.byref(self.h)))
        self._n = {}

    def submit(self, prob: mesh.Problem, rhs=None) -> int:
        """Queue one solve of prob (load case rhs, default prob.rhs[0]); returns a ticket."""
        y = _f64(prob.rhs[0] if rhs is None else rhs)
        cn = _i64(prob.elem_codes())
        conn = _i64(prob.conn)
        xyz = _f64(prob.coords)
        e = _f64(prob.e)
        nu = _f64(prob.nu) if prob.kind != mesh.KIND_HEX8_HEAT else None
        th = _f64(prob.thick) if prob.kind == mesh.KIND_QUAD4_PSTRESS else None
        pv = _f64(prob.pv) if prob.pv.size else None
        t = _L()
        _check(lib().ssf_server_submit(self.h, prob.kind, prob.n, xyz.shape[0], _p(xyz),
                                       conn.shape[0], _p(conn), _p(cn), _p(e), _p(nu), _p(th),
                                       _p(pv), 0 if pv is None else pv.shape[0], _p(y),
                                       ctypes.byref(t)))
        self._n[t.value] = prob.n
        return t.value

    def wait(self, ticket: int):
        """Returns (x, info, times[submitted, admitted, assembled, finished])."""
        n = self._n.pop(ticket)
        x = np.zeros(n)
        info = SolveInfo(-1, -1.0, 0, 0, 0, 0, 0.0)
        times = np.zeros(4)
        _check(lib().ssf_server_wait(self.h, ticket, _p(x), ctypes.byref(info), _p(times)))
        return x, info, times

    def stats(self):
        s = np.zeros(8, np.int64)
        _check(lib().ssf_server_stats(self.h, _p(s)))
        return {"finished": int(s[0]), "steps": int(s[1]), "peak_bytes": int(s[2]),
                "reserved_bytes": int(s[3]), "pattern_s": s[4] * 1e-6, "setup_s": s[5] * 1e-6,
                "assembly_s": s[6] * 1e-6, "half_form_requests": int(s[7])}

    def close(self):
        if self.h:
            lib().ssf_server_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Elements:
    """Device-resident element set of a mesh.Problem for the post-solve element
    pass: integration-point strains / stresses (heat: gradients / fluxes) and the
    internal forces (heat: internal fluxes) of a solution."""

    def __init__(self, ctx: Context, prob: mesh.Problem):
        self.ctx = ctx
        self.n = prob.n
        self.ne = prob.ne
        cn = _i64(prob.elem_codes())
        conn = _i64(prob.conn)
        xyz = _f64(prob.coords)
        e = _f64(prob.e)
        nu = _f64(prob.nu) if prob.kind != mesh.KIND_HEX8_HEAT else None
        th = _f64(prob.thick) if prob.kind == mesh.KIND_QUAD4_PSTRESS else None
        self.pv = _f64(prob.pv) if prob.pv.size else None
        self.h = _P()
        _check(lib().ssf_elements_create(ctx.h, prob.kind, prob.n, xyz.shape[0], _p(xyz),
                                         conn.shape[0], _p(conn), _p(cn), _p(e), _p(nu), _p(th),
                                         ctypes.byref(self.h)))
        a = np.zeros(3, np.int32)
        _check(lib().ssf_elements_sizes(self.h, _p(a[0:1]), _p(a[1:2]), _p(a[2:3])))
        self.ngp, self.ncomp, self.ndofe = (int(v) for v in a)

    def results(self, lhs, pv=None, strain=True, stress=True, forces=True):
        """(strain [ne, ngp, ncomp], stress [ne, ngp, ncomp], ifor [n]); outputs
        not asked for are None. pv defaults to the problem's prescribed values."""
        pv = self.pv if pv is None else _f64(pv)
        shp = (self.ne, self.ngp, self.ncomp)
        eps = np.zeros(shp) if strain else None
        sig = np.zeros(shp) if stress else None
        f = np.zeros(self.n) if forces else None
        _check(lib().ssf_elements_results(self.h, _p(_f64(lhs)), _p(pv),
                                          0 if pv is None else pv.shape[0], _p(eps), _p(sig),
                                          _p(f)))
        return eps, sig, f

    def surface_load(self, eid, ls, nodvals, rhs=None):
        """Surface loads of records (element eid[r], face flags ls[r, 6], nodal values
        nodvals[r, 72]) added into a copy of rhs (zeros if None) in record order."""
        out = np.zeros(self.n) if rhs is None else _f64(rhs).copy()
        eid, ls, nodvals = _i64(eid), _i64(ls), _f64(nodvals)
        _check(lib().ssf_elements_surface_load(self.h, eid.shape[0], _p(eid), _p(ls),
                                               _p(nodvals), _p(out)))
        return out

    def surface_load_dev(self, nrec, eid_ptr, ls_ptr, nodvals_ptr, rhs_ptr):
        _check(lib().ssf_elements_surface_load_dev(self.h, nrec, ctypes.c_void_p(eid_ptr),
                                                   ctypes.c_void_p(ls_ptr),
                                                   ctypes.c_void_p(nodvals_ptr),
                                                   ctypes.c_void_p(rhs_ptr)))

    def volume_load(self, nodval, rhs=None):
        """Element volume loads (heat: sources) with nodal values nodval (ne, ndofe)
        added into a copy of rhs (zeros if None) in element order."""
        out = np.zeros(self.n) if rhs is None else _f64(rhs).copy()
        _check(lib().ssf_elements_volume_load(self.h, _p(_f64(nodval)), _p(out)))
        return out

    def results_dev(self, lhs_ptr, pv_ptr, npv, strain_ptr, stress_ptr, ifor_ptr):
        _check(lib().ssf_elements_results_dev(self.h, lhs_ptr, pv_ptr, npv, strain_ptr,
                                              stress_ptr, ifor_ptr))

    def bench(self, lhs_ptr, pv_ptr, npv, strain_ptr, stress_ptr, ifor_ptr, reps=10):
        """(element kernel us, whole call us), means over reps calls"""
        us = np.zeros(2)
        _check(lib().ssf_elements_bench(self.h, lhs_ptr, pv_ptr, npv, strain_ptr, stress_ptr,
                                        ifor_ptr, reps, _p(us)))
        return float(us[0]), float(us[1])

    def close(self):
        if self.h:
            lib().ssf_elements_destroy(self.h)
            self.h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_problem_matrix(ctx: Context, prob: mesh.Problem, storage=STORAGE_SCR):
    """Pattern + device assembly of a mesh.Problem; returns (matrix, rhs[nrhs, n])."""
    m = Matrix.from_elements(ctx, storage, prob.n, prob.elem_codes())
    lift = m.assemble(prob)
    rhs = prob.rhs.copy()
    if lift is not None:
        rhs = rhs + lift[None, :]
    return m, rhs
