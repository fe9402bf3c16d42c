"""ctypes bindings of the CPU oracle (oracle/libcpu_ref.so) and, where it was built,
of the reference's own GEFEL code (oracle/_ref/libssfref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpu_ref.so")
REF_PATH = os.path.join(HERE, "_ref", "libssfref.so")

_P = ctypes.c_void_p
_L = ctypes.c_long
_D = ctypes.c_double
_lib = None
_ref = None


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run make -C oracle)")
        l = ctypes.CDLL(LIB_PATH)
        sig = {
            "cr_codenum": (_L, [_L, _L, _P, _P]),
            "cr_elem_codes": (None, [_L, _L, _P, _L, _P, _P]),
            "cr_scr_pattern": (_L, [_L, _L, _L, _P, _P, _P, _L]),
            "cr_cr_pattern": (_L, [_L, _L, _L, _P, _P, _P, _L]),
            "cr_scr_localized": (None, [_L, _P, _P, _P, _P, _P, _L]),
            "cr_cr_localized": (None, [_L, _P, _P, _P, _P, _P, _L]),
            "cr_scr_assemble": (None, [_L, _P, _P, _P, _L, _L, _P, _P]),
            "cr_cr_assemble": (None, [_L, _P, _P, _P, _L, _L, _P, _P]),
            "cr_mxv_scr": (None, [_L, _P, _P, _P, _P, _P]),
            "cr_mxv_cr": (None, [_L, _P, _P, _P, _P, _P]),
            "cr_mtxv_cr": (None, [_L, _P, _P, _P, _P, _P]),
            "cr_edge_adjacency": (None, [_L, _L, _L, _P, _P, _P]),
            "cr_cuthill_mckee": (ctypes.c_int, [_L, _P, _P, ctypes.c_int, _P]),
            "cr_hex8_surface_load": (None, [_P, _P, _P, _P, _P, _P]),
            "cr_surface_load_mesh": (None, [_L, _P, _P, _P, _P, _P, _P, _P]),
            "cr_newton_raphson": (ctypes.c_double, [_L, _P, _P, _P, ctypes.c_int, _L, _P, _P, _P,
                                                    _P, _P, _P, _P, _P, _P, _P, _P, _L, _D, _L,
                                                    _L, _D, _D, _D, _D, ctypes.c_int, ctypes.c_int,
                                                    _D, _L, _D, _D, _P, _P, _P, _P, _P]),
            "cr_scr_incomplete_ldl": (None, [_L, _P, _P, _P, _D, _P]),
            "cr_scr_prec_ildl": (None, [_L, _P, _P, _P, _P, _P]),
            "cr_scr_cg_prec_ildl": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _P, _L, _D, _P, _P,
                                                   _D, _L]),
            "cr_cr_bicg": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "cr_scr_cg_prec": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "cr_scr_cg": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "cr_cr_cg_prec_noprec": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "cr_hex8_elast_ke": (None, [_P, _P, _P, _D, _D, _P]),
            "cr_hex8_heat_ke": (None, [_P, _P, _P, _D, _P]),
            "cr_quad4_pstress_ke": (None, [_P, _P, _D, _D, _P, _P]),
            "cr_hex8_elast_ke_mesh": (None, [_L, _P, _P, _P, _P, _P]),
            "cr_hex8_heat_ke_mesh": (None, [_L, _P, _P, _P, _P]),
            "cr_quad4_pstress_ke_mesh": (None, [_L, _P, _P, _P, _P, _P, _P]),
            "cr_dirichlet_lift": (None, [_L, _L, _P, _P, _P, _P]),
            "cr_eldispl": (None, [_L, _P, _P, _P, _P]),
            "cr_hex8_elast_ip": (None, [_P, _P, _P, _D, _D, _P, _P, _P, _P]),
            "cr_quad4_pstress_ip": (None, [_P, _P, _D, _D, _P, _P, _P, _P, _P]),
            "cr_hex8_heat_ip": (None, [_P, _P, _P, _D, _P, _P, _P, _P]),
            "cr_elem_results_mesh": (None, [ctypes.c_int, _L, _P, _P, _P, _P, _P, _P, _L, _P,
                                            _P, _P, _P, _P]),
            "cr_hex8_heat_cap": (None, [_P, _P, _P, _D, _P, _P]),
            "cr_hex8_heat_cap_mesh": (None, [_L, _P, _P, _P, _P, _P]),
            "cr_hex8_volume_load": (None, [ctypes.c_int, _P, _P, _P, _P, _P]),
            "cr_volume_load_mesh": (None, [ctypes.c_int, _L, _P, _P, _P, _P, _P]),
            "cr_np_step_linear": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P, _P,
                                                 _P, _P, _L, _D, _D, _P, _P]),
            "cr_np_step_dform": (ctypes.c_int, [_L, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P, _P,
                                                _P, _P, _L, _D, _D, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """The reference's own GEFEL code behind ref_harness.cpp (built only where
    /root/reference exists)."""
    global _ref
    if _ref is None:
        l = ctypes.CDLL(REF_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
        sig = {
            "ref_scr_create": (_P, [_L, _L, _L, _P]),
            "ref_scr_negm": (_L, [_P]),
            "ref_scr_get": (None, [_P, _P, _P, _P]),
            "ref_scr_set_a": (None, [_P, _P]),
            "ref_scr_localized": (None, [_P, _P, _P, _L]),
            "ref_scr_mxv": (None, [_P, _P, _P]),
            "ref_scr_cg_prec": (None, [_P, _P, _P, _L, _D, _P, _P, _D, _L, _L]),
            "ref_scr_cg": (None, [_P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "ref_scr_free": (None, [_P]),
            "ref_cr_create": (_P, [_L, _L, _L, _P]),
            "ref_cr_negm": (_L, [_P]),
            "ref_cr_get": (None, [_P, _P, _P, _P]),
            "ref_cr_localized": (None, [_P, _P, _P, _L]),
            "ref_cr_mxv": (None, [_P, _P, _P]),
            "ref_cr_mtxv": (None, [_P, _P, _P]),
            "ref_cuthill_mckee": (None, [_L, _L, _L, _P, ctypes.c_int, _P]),
            "ref_hex8_surface_load": (None, [_P, _P, _P, _P, _P, _P]),
            "ref_scr_incomplete_ldl": (None, [_P, _D, _P]),
            "ref_scr_prec_ildl": (None, [_P, _P, _P]),
            "ref_cr_set_a": (None, [_P, _P]),
            "ref_cr_bicg": (None, [_P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "ref_cr_cg_prec_noprec": (None, [_P, _P, _P, _L, _D, _P, _P, _D, _L]),
            "ref_cr_free": (None, [_P]),
            "ref_hex8_elast_ke": (None, [_P, _P, _P, _D, _D, _P]),
            "ref_quad4_pstress_ke": (None, [_P, _P, _D, _D, _P, _P]),
            "ref_hex8_heat_ke": (None, [_P, _P, _P, _D, _P]),
            "ref_hex8_elast_ip": (None, [_P, _P, _P, _D, _D, _P, _P, _P, _P]),
            "ref_quad4_pstress_ip": (None, [_P, _P, _D, _D, _P, _P, _P, _P, _P]),
            "ref_hex8_heat_ip": (None, [_P, _P, _P, _D, _P, _P, _P, _P]),
            "ref_locglob": (None, [_P, _P, _P, _L]),
            "ref_hex8_volume_load": (None, [ctypes.c_int, _P, _P, _P, _P, _P]),
            "ref_hex8_heat_cap": (None, [_P, _P, _P, _D, _P, _P]),
            "ref_np_step_linear": (None, [_P, _P, _P, _L, _D, _D, _P, _P, _P, _L, _D, _D, _P, _P]),
            "ref_np_step_dform": (None, [_P, _P, _P, _L, _D, _D, _P, _P, _P, _L, _D, _D, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _ref = l
    return _ref


# ------------------------------------------------------------------ helpers

def i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def scr_pattern(n, cn):
    cn = i64(cn)
    ne, ndofe = cn.shape
    adr = np.zeros(n + 1, np.int64)
    negm = lib().cr_scr_pattern(n, ne, ndofe, _ptr(cn), _ptr(adr), None, -1)
    ci = np.zeros(negm, np.int64)
    lib().cr_scr_pattern(n, ne, ndofe, _ptr(cn), _ptr(adr), _ptr(ci), negm)
    return adr, ci


def cr_pattern(n, cn):
    cn = i64(cn)
    ne, ndofe = cn.shape
    adr = np.zeros(n + 1, np.int64)
    negm = lib().cr_cr_pattern(n, ne, ndofe, _ptr(cn), _ptr(adr), None, -1)
    ci = np.zeros(negm, np.int64)
    lib().cr_cr_pattern(n, ne, ndofe, _ptr(cn), _ptr(adr), _ptr(ci), negm)
    return adr, ci


def element_matrices(prob):
    """K_e for every element of a mesh.Problem, (ne, ndofe, ndofe)."""
    from ssf_amd import mesh
    ne, nd = prob.ne, prob.ndofe
    ke = np.zeros((ne, nd, nd))
    conn = i64(prob.conn)
    if prob.kind == mesh.KIND_HEX8_ELAST:
        lib().cr_hex8_elast_ke_mesh(ne, _ptr(conn), _ptr(f64(prob.coords)), _ptr(f64(prob.e)),
                                    _ptr(f64(prob.nu)), _ptr(ke))
    elif prob.kind == mesh.KIND_HEX8_HEAT:
        lib().cr_hex8_heat_ke_mesh(ne, _ptr(conn), _ptr(f64(prob.coords)), _ptr(f64(prob.e)),
                                   _ptr(ke))
    else:
        xy = f64(prob.coords[:, :2])
        lib().cr_quad4_pstress_ke_mesh(ne, _ptr(conn), _ptr(xy), _ptr(f64(prob.e)),
                                       _ptr(f64(prob.nu)), _ptr(f64(prob.thick)), _ptr(ke))
    return ke


def scr_assemble(n, adr, ci, cn, ke):
    a = np.zeros(ci.shape[0])
    cn = i64(cn)
    ke = f64(ke)
    lib().cr_scr_assemble(n, _ptr(adr), _ptr(ci), _ptr(a), cn.shape[0], cn.shape[1],
                          _ptr(cn), _ptr(ke))
    return a


def cr_assemble(n, adr, ci, cn, ke):
    a = np.zeros(ci.shape[0])
    cn = i64(cn)
    ke = f64(ke)
    lib().cr_cr_assemble(n, _ptr(adr), _ptr(ci), _ptr(a), cn.shape[0], cn.shape[1],
                         _ptr(cn), _ptr(ke))
    return a


def dirichlet_lift(prob, cn, ke):
    rhs = np.zeros(prob.n)
    cn = i64(cn)
    lib().cr_dirichlet_lift(cn.shape[0], cn.shape[1], _ptr(cn), _ptr(f64(ke)),
                            _ptr(f64(prob.pv)), _ptr(rhs))
    return rhs


def mxv_scr(adr, ci, a, b):
    n = adr.shape[0] - 1
    c = np.zeros(n)
    lib().cr_mxv_scr(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(f64(b)), _ptr(c))
    return c


def mxv_cr(adr, ci, a, b):
    n = adr.shape[0] - 1
    c = np.zeros(n)
    lib().cr_mxv_cr(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(f64(b)), _ptr(c))
    return c


def scr_cg_prec(adr, ci, a, y, ni, res, zero=1e-30, x0=None):
    n = adr.shape[0] - 1
    x = np.zeros(n) if x0 is None else f64(x0).copy()
    ani = ctypes.c_long(-1)
    ares = ctypes.c_double(-1.0)
    st = lib().cr_scr_cg_prec(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(x), _ptr(f64(y)),
                              ni, res, ctypes.byref(ani), ctypes.byref(ares), zero,
                              0 if x0 is None else 1)
    return x, ani.value, ares.value, st


def scr_cg(adr, ci, a, y, ni, res, zero=1e-30):
    n = adr.shape[0] - 1
    x = np.zeros(n)
    ani = ctypes.c_long(-1)
    ares = ctypes.c_double(-1.0)
    st = lib().cr_scr_cg(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(x), _ptr(f64(y)), ni, res,
                         ctypes.byref(ani), ctypes.byref(ares), zero, 0)
    return x, ani.value, ares.value, st


def cr_cg_prec_noprec(adr, ci, a, y, ni, res, zero=1e-30):
    n = adr.shape[0] - 1
    x = np.zeros(n)
    ani = ctypes.c_long(-1)
    ares = ctypes.c_double(-1.0)
    st = lib().cr_cr_cg_prec_noprec(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(x),
                                    _ptr(f64(y)), ni, res, ctypes.byref(ani),
                                    ctypes.byref(ares), zero, 0)
    return x, ani.value, ares.value, st


def incomplete_ldl(adr, ci, a, gamma=0.0):
    """symcomprow::incomplete_ldl: the factor incdec[negm] (L off the diagonal, D on it)."""
    n = adr.shape[0] - 1
    out = np.zeros(a.shape[0])
    lib().cr_scr_incomplete_ldl(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), gamma, _ptr(out))
    return out


def prec_ildl(adr, ci, incdec, y):
    """symcomprow::prec_ildl_scr: (L D L^T)^-1 y."""
    n = adr.shape[0] - 1
    x = np.zeros(n)
    lib().cr_scr_prec_ildl(n, _ptr(adr), _ptr(ci), _ptr(f64(incdec)), _ptr(x), _ptr(f64(y)))
    return x


def scr_cg_prec_ildl(adr, ci, a, y, ni, res, zero=1e-30, incdec=None, x0=None):
    """cg_prec with tprec = 10 (incomplete LDL^T, factored here when incdec is None).
    Returns (x, ani, ares, status)."""
    n = adr.shape[0] - 1
    inc = incomplete_ldl(adr, ci, a) if incdec is None else f64(incdec)
    x = np.zeros(n) if x0 is None else f64(x0).copy()
    ani = ctypes.c_long(-1)
    ares = ctypes.c_double(-1.0)
    st = lib().cr_scr_cg_prec_ildl(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(inc), _ptr(x),
                                   _ptr(f64(y)), ni, res, ctypes.byref(ani), ctypes.byref(ares),
                                   zero, 0 if x0 is None else 1)
    return x, ani.value, ares.value, st


def newton_raphson(prob, adr, ci, a, fp, fc=None, ilambda=0.0, li=0, ninr=50, niilnr=10,
                   errnr=1e-6, incrnr=0.1, minincrnr=1e-3, maxincrnr=1.0, normt=1,
                   check_lambdar=False, lambdar=1.0, sni=20000, sres=1e-24, szero=1e-30,
                   ra=None, fa=None):
    """gnewton_raphson with linear elastic internal forces and Jacobi cg_prec solves.
    Returns (lambda, ra, fa, hist) with hist = dict(lam, dlam, nsol, acc) over the
    increments run."""
    n = prob.n
    fc = np.zeros(n) if fc is None else f64(fc)
    ra = np.zeros(n) if ra is None else f64(ra).copy()
    fa = np.zeros(n) if fa is None else f64(fa).copy()
    m = ninr - li
    lam, dlam = np.zeros(m), np.zeros(m)
    nsol, acc = np.zeros(m, np.int64), np.zeros(m, np.int64)
    counts = np.zeros(2, np.int64)
    pv = f64(prob.pv) if prob.pv.size else np.zeros(1)
    lam_end = lib().cr_newton_raphson(
        n, _ptr(adr), _ptr(ci), _ptr(f64(a)), prob.kind, prob.ne, _ptr(i64(prob.conn)),
        _ptr(f64(prob.coords)), _ptr(i64(prob.elem_codes())), _ptr(f64(prob.e)),
        _ptr(f64(prob.nu)), _ptr(f64(prob.thick)), _ptr(pv), _ptr(fc), _ptr(f64(fp)), _ptr(fa),
        _ptr(ra), li, ilambda, ninr, niilnr, errnr, incrnr, minincrnr, maxincrnr, normt,
        1 if check_lambdar else 0, lambdar, sni, sres, szero, _ptr(lam), _ptr(dlam), _ptr(nsol),
        _ptr(acc), _ptr(counts))
    k = int(counts[0])
    hist = {"lam": lam[:k], "dlam": dlam[:k], "nsol": nsol[:k], "acc": acc[:k],
            "solves": int(counts[1])}
    return lam_end, ra, fa, hist


def edge_adjacency(nn, conn):
    """adjacnodes_edge: CSR (ptr, adj) of the element-edge graph, neighbours ascending."""
    conn = i64(conn)
    ne, nne = conn.shape
    ptr = np.zeros(nn + 1, np.int64)
    adj = np.zeros(max(1, 2 * ne * (12 if nne == 8 else 4)), np.int64)
    lib().cr_edge_adjacency(nn, ne, nne, _ptr(conn), _ptr(ptr), _ptr(adj))
    return ptr, adj[:ptr[-1]].copy()


def cuthill_mckee(nn, conn, reverse=False):
    """cuthill_mckee_renumb: ordering[k] = the node numbered k; None if the graph is
    not connected."""
    ptr, adj = edge_adjacency(nn, conn)
    out = np.zeros(nn, np.int64)
    rc = lib().cr_cuthill_mckee(nn, _ptr(ptr), _ptr(adj), 1 if reverse else 0, _ptr(out))
    return out if rc == 0 else None


def cr_bicg(adr, ci, a, y, ni, res, zero=1e-30, x0=None):
    """comprow::bicg; x0 given: start from it (iv = 1). Returns (x, ani, ares, status)."""
    n = adr.shape[0] - 1
    x = np.zeros(n) if x0 is None else f64(x0).copy()
    ani = ctypes.c_long(-1)
    ares = ctypes.c_double(-1.0)
    st = lib().cr_cr_bicg(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(x), _ptr(f64(y)), ni, res,
                          ctypes.byref(ani), ctypes.byref(ares), zero,
                          0 if x0 is None else 1)
    return x, ani.value, ares.value, st


def mtxv_cr(adr, ci, a, b):
    n = adr.shape[0] - 1
    c = np.zeros(n)
    lib().cr_mtxv_cr(n, _ptr(adr), _ptr(ci), _ptr(f64(a)), _ptr(f64(b)), _ptr(c))
    return c


def nonsymmetric_values(adr, ci, a, gamma, seed):
    """Values of a nonsymmetric system on a comprow pattern: every off-diagonal
    a_ij scaled by 1 + gamma t_ij with t_ij uniform in [-1, 1] drawn per entry,
    so a_ij and a_ji are scaled independently (test data, not a reference
    routine)."""
    rng = np.random.default_rng(seed)
    n = adr.shape[0] - 1
    rows = np.repeat(np.arange(n), np.diff(adr))
    t = rng.uniform(-1.0, 1.0, a.shape[0])
    return np.where(ci == rows, a, a * (1.0 + gamma * t))


def assemble_problem(prob, storage="scr"):
    """Oracle assembly of a mesh.Problem: (adr, ci, a, rhs[nrhs, n]) with the
    Dirichlet lift added to every right-hand side."""
    cn = prob.elem_codes()
    ke = element_matrices(prob)
    if storage == "scr":
        adr, ci = scr_pattern(prob.n, cn)
        a = scr_assemble(prob.n, adr, ci, cn, ke)
    else:
        adr, ci = cr_pattern(prob.n, cn)
        a = cr_assemble(prob.n, adr, ci, cn, ke)
    rhs = prob.rhs.copy()
    if prob.pv.size:
        rhs = rhs + dirichlet_lift(prob, cn, ke)[None, :]
    return adr, ci, a, rhs


# integration points and components per point of the post-solve element pass:
# hex8 elasticity (eps_x, eps_y, eps_z, gamma_yz, gamma_xz, gamma_xy), hex8 heat
# (gradient / flux x, y, z), quad4 plane stress (eps_x, eps_y, gamma_xy, eps_z)
POST_SHAPE = {0: (8, 6), 1: (8, 3), 2: (4, 4)}


def elem_results(prob, lhs):
    """Oracle post-solve pass of a mesh.Problem for the solution lhs[n]:
    (ip strains or gradients [ne, ngp, nc], ip stresses or fluxes, internal forces [n])."""
    ngp, nc = POST_SHAPE[prob.kind]
    eps = np.zeros((prob.ne, ngp, nc))
    sig = np.zeros((prob.ne, ngp, nc))
    ifor = np.zeros(prob.n)
    pv = f64(prob.pv) if prob.pv.size else np.zeros(1)
    lib().cr_elem_results_mesh(prob.kind, prob.ne, _ptr(i64(prob.conn)), _ptr(f64(prob.coords)),
                               _ptr(i64(prob.elem_codes())), _ptr(f64(prob.e)),
                               _ptr(f64(prob.nu)), _ptr(f64(prob.thick)), prob.n,
                               _ptr(f64(lhs)), _ptr(pv), _ptr(eps), _ptr(sig), _ptr(ifor))
    return eps, sig, ifor


def elem_ip(kind, x, y, z, ma, mb, th, r):
    """One element of the post-solve pass: (ip values, ip values, element forces)."""
    ngp, nc = POST_SHAPE[kind]
    nd = r.shape[0]
    eps = np.zeros((ngp, nc))
    sig = np.zeros((ngp, nc))
    f = np.zeros(nd)
    x, y, z, r = f64(x), f64(y), f64(z), f64(r)
    if kind == 0:
        lib().cr_hex8_elast_ip(_ptr(x), _ptr(y), _ptr(z), ma, mb, _ptr(r), _ptr(eps), _ptr(sig),
                               _ptr(f))
    elif kind == 1:
        lib().cr_hex8_heat_ip(_ptr(x), _ptr(y), _ptr(z), ma, _ptr(r), _ptr(eps), _ptr(sig), _ptr(f))
    else:
        t = f64(np.full(4, th))
        lib().cr_quad4_pstress_ip(_ptr(x), _ptr(y), ma, mb, _ptr(t), _ptr(r), _ptr(eps), _ptr(sig),
                                  _ptr(f))
    return eps, sig, f


def capacity_matrices(prob, c, rho=None):
    """C_e of every element of a hex8 heat mesh, (ne, 8, 8): specific heat c[ne],
    density rho[ne] (None = 1), equal at the element's nodes."""
    ne = prob.ne
    cm = np.zeros((ne, 8, 8))
    lib().cr_hex8_heat_cap_mesh(ne, _ptr(i64(prob.conn)), _ptr(f64(prob.coords)), _ptr(f64(c)),
                                None if rho is None else _ptr(f64(rho)), _ptr(cm))
    return cm


def np_step_linear(adr, ci, aK, aC, alpha, dt, rhs, lhs, tdlhs, ni=20000, res=1e-20,
                   zero=1e-30, form="v"):
    """One generalized-trapezoidal step; lhs and tdlhs are updated in place.
    form "v": one_step_linear, "d": the d-form step. Returns (status, ani, ares, aA)."""
    n = adr.shape[0] - 1
    aA = np.zeros(aK.shape[0])
    d, p, f = np.zeros(n), np.zeros(n), np.zeros(n)
    ani = ctypes.c_long(-1)
    ares = ctypes.c_double(-1.0)
    fn = {"v": lib().cr_np_step_linear, "d": lib().cr_np_step_dform}[form]
    st = fn(n, _ptr(adr), _ptr(ci), _ptr(f64(aK)), _ptr(f64(aC)), _ptr(aA), alpha, dt,
            _ptr(f64(rhs)), _ptr(lhs), _ptr(tdlhs), _ptr(d), _ptr(p), _ptr(f), ni, res, zero,
            ctypes.byref(ani), ctypes.byref(ares))
    return st, ani.value, ares.value, aA


def surface_load(prob, eid, ls, nodvals, rhs=None):
    """Surface loads of hex8 elasticity records (element eid[r], face flags
    ls[r, 6], nodal values nodvals[r, 72]) added into a copy of rhs (zeros if
    None) by locglob in record order."""
    out = np.zeros(prob.n) if rhs is None else f64(rhs).copy()
    eid, ls, nodvals = i64(eid), i64(ls), f64(nodvals)
    lib().cr_surface_load_mesh(eid.shape[0], _ptr(eid), _ptr(ls), _ptr(nodvals),
                               _ptr(i64(prob.conn)), _ptr(f64(prob.coords)),
                               _ptr(i64(prob.elem_codes())), _ptr(out))
    return out


def hex8_surface_load(x, y, z, ls, nodvals):
    """nf[24] of one hex8 element (node_forces_surf)."""
    nf = np.zeros(24)
    lib().cr_hex8_surface_load(_ptr(f64(x)), _ptr(f64(y)), _ptr(f64(z)), _ptr(i64(ls)),
                               _ptr(f64(nodvals)), _ptr(nf))
    return nf


def volume_load(prob, nodval, rhs=None):
    """Element volume loads (elasticity) / heat sources of a hex8 mesh added into
    rhs (a copy of rhs, or zeros) in element order; nodval (ne, ndofe)."""
    out = np.zeros(prob.n) if rhs is None else f64(rhs).copy()
    lib().cr_volume_load_mesh(prob.kind, prob.ne, _ptr(i64(prob.conn)), _ptr(f64(prob.coords)),
                              _ptr(i64(prob.elem_codes())), _ptr(f64(nodval)), _ptr(out))
    return out
