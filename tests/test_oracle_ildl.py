"""CPU checks that pin the oracle's incomplete LDL^T (cr_scr_incomplete_ldl,
cr_scr_prec_ildl, cr_scr_cg_prec_ildl: symcomprow::incomplete_ldl,
prec_ildl_scr and cg_prec with tprec = 10, GEFEL/scr.cpp:958-1203):

* bitwise against the reference: the committed fixture (factor of the heat and
  elasticity matrices, the preconditioner applied to a seeded vector, the
  preconditioned CG's x, ani, ares) and, where oracle/_ref was built, live on
  further seeds with gamma 0 and gamma > 0;
* mathematics: where the pattern admits no fill (tridiagonal, dense) the factor
  is the exact LDL^T, so applying it solves A x = y and cg_prec stops after one
  iteration; on any pattern L D L^T equals A at the pattern's positions
  (gamma = 0), and with gamma > 0 its diagonal is a_ii - gamma times the
  dropped fill of row i.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, f64, i64
from ssf_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
CASES = {"heat": lambda: mesh.heat_cube(4, seed=22),
         "elast": lambda: mesh.elastic_column(2, 2, 3, seed=21)}


def _dense(adr, ci, v):
    n = adr.shape[0] - 1
    M = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(adr))
    M[rows, ci] = v
    return M


def _ldl(adr, ci, inc):
    """L (unit lower) and D from incdec."""
    n = adr.shape[0] - 1
    diag = adr[1:] - 1
    D = inc[diag]
    L = _dense(adr, ci, inc)
    L[np.arange(n), np.arange(n)] = 1.0
    return L, D


@pytest.mark.parametrize("name", ["heat", "elast"])
def test_ildl_bitwise_fixture(name):
    g = np.load(GOLD)
    prob = CASES[name]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    inc = O.incomplete_ldl(adr, ci, a)
    assert np.array_equal(inc, g[f"{name}_ildl_inc"])
    assert np.array_equal(O.prec_ildl(adr, ci, inc, g[f"{name}_ildl_yv"]), g[f"{name}_ildl_pv"])
    x, ani, ares, st = O.scr_cg_prec_ildl(adr, ci, a, rhs[0], 2000, 1e-24)
    assert st == 1 and np.array_equal(x, g[f"{name}_ildl_x"])
    assert ani == g[f"{name}_ildl_ani"] and ares == g[f"{name}_ildl_ares"]


def test_ildl_live_reference():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = O.ref()
    for seed, make in enumerate((lambda s: mesh.heat_cube(5, seed=s),
                                 lambda s: mesh.elastic_column(2, 3, 3, seed=s),
                                 lambda s: mesh.plate(7, 4, seed=s, nrhs=1))):
        prob = make(80 + seed)
        adr, ci, a, rhs = O.assemble_problem(prob, "scr")
        cn = i64(prob.elem_codes())
        h = R.ref_scr_create(prob.n, prob.ne, prob.ndofe, _ptr(cn))
        R.ref_scr_set_a(h, _ptr(f64(a)))
        y = np.random.default_rng(seed).standard_normal(prob.n)
        for gamma in (0.7, 0.0):  # the harness keeps the last factor (gamma 0) for cg
            inc = O.incomplete_ldl(adr, ci, a, gamma)
            inc2 = np.zeros(a.shape[0])
            R.ref_scr_incomplete_ldl(h, gamma, _ptr(inc2))
            assert np.array_equal(inc, inc2)
            p2 = np.zeros(prob.n)
            R.ref_scr_prec_ildl(h, _ptr(p2), _ptr(y.copy()))
            assert np.array_equal(O.prec_ildl(adr, ci, inc, y), p2)
        x, ani, ares, _ = O.scr_cg_prec_ildl(adr, ci, a, rhs[0], 2000, 1e-22)
        x2 = np.zeros(prob.n)
        a2, r2 = ctypes.c_long(-1), ctypes.c_double(-1.0)
        R.ref_scr_cg_prec(h, _ptr(x2), _ptr(f64(rhs[0]).copy()), 2000, 1e-22, ctypes.byref(a2),
                          ctypes.byref(r2), 1e-30, 0, 10)
        assert np.array_equal(x, x2) and ani == a2.value and ares == r2.value
        R.ref_scr_free(h)


def _scr_of_dense(M, keep):
    """symcomprow arrays of the lower triangle of M at the positions keep[i, j]."""
    n = M.shape[0]
    adr, ci, a = [0], [], []
    for i in range(n):
        for j in range(i + 1):
            if j == i or keep[i, j]:
                ci.append(j)
                a.append(M[i, j])
        adr.append(len(ci))
    return i64(adr), i64(ci), f64(a)


@pytest.mark.parametrize("shape", ["tridiagonal", "dense"])
def test_ildl_is_exact_without_fill(shape):
    n = 9
    rng = np.random.default_rng(81)
    if shape == "tridiagonal":
        M = np.diag(rng.uniform(3, 4, n)) + np.diag(rng.uniform(-1, 1, n - 1), -1)
        M = np.tril(M) + np.tril(M, -1).T
        keep = np.abs(np.subtract.outer(np.arange(n), np.arange(n))) == 1
    else:
        B = rng.standard_normal((n, n))
        M = B @ B.T + n * np.eye(n)
        keep = np.ones((n, n), bool)
    adr, ci, a = _scr_of_dense(M, keep)
    inc = O.incomplete_ldl(adr, ci, a)
    L, D = _ldl(adr, ci, inc)
    assert np.abs(L @ np.diag(D) @ L.T - M).max() <= 1e-13 * np.abs(M).max()
    y = rng.standard_normal(n)
    assert np.abs(O.prec_ildl(adr, ci, inc, y) - np.linalg.solve(M, y)).max() <= 1e-12
    x, ani, ares, st = O.scr_cg_prec_ildl(adr, ci, a, y, 50, 1e-24)
    assert st == 1 and ani == 0


@pytest.mark.parametrize("make", [lambda: mesh.heat_cube(5, seed=82),
                                  lambda: mesh.elastic_column(2, 2, 3, seed=83)])
def test_ildl_matches_a_on_the_pattern(make):
    prob = make()
    adr, ci, a, _ = O.assemble_problem(prob, "scr")
    inc = O.incomplete_ldl(adr, ci, a)
    L, D = _ldl(adr, ci, inc)
    P = L @ np.diag(D) @ L.T
    A = _dense(adr, ci, a)
    on = A != 0
    on |= _dense(adr, ci, np.ones_like(a)) != 0
    assert np.abs((P - (A + np.tril(A, -1).T))[on]).max() <= 1e-12 * np.abs(a).max()
    # and it is incomplete: fill outside the pattern is dropped
    assert np.abs(np.tril(P, -1)[~on & np.tril(np.ones_like(on), -1).astype(bool)]).max() > 0


def test_ildl_gamma_lumps_the_dropped_fill():
    prob = mesh.heat_cube(4, seed=84)
    adr, ci, a, _ = O.assemble_problem(prob, "scr")
    gamma = 0.6
    inc = O.incomplete_ldl(adr, ci, a, gamma)
    L, D = _ldl(adr, ci, inc)
    P = L @ np.diag(D) @ L.T
    n = prob.n
    on = _dense(adr, ci, np.ones_like(a)) != 0
    for i in range(n):
        lj = ci[adr[i]]
        dropped = [j for j in range(lj, i) if not on[i, j]]
        want = a[adr[i + 1] - 1] - gamma * sum(P[i, j] for j in dropped)
        assert abs(P[i, i] - want) <= 1e-12 * abs(want)


def test_ildl_checks_catch_planted_mistakes():
    """A factorization that leaves D out of the row products (a plausible slip)
    fails the pattern check that the oracle's factor passes."""
    prob = mesh.heat_cube(3, seed=85)
    adr, ci, a, _ = O.assemble_problem(prob, "scr")
    n = prob.n
    inc = np.zeros_like(a)
    diag = adr[1:] - 1
    for i in range(n):
        for ii in range(adr[i], diag[i]):
            j = ci[ii]
            s = sum(inc[p] * inc[q] for p in range(adr[i], ii) for q in range(adr[j], diag[j])
                    if ci[p] == ci[q])  # D_k missing
            inc[ii] = (a[ii] - s) / inc[diag[j]]
        inc[diag[i]] = a[diag[i]] - sum(inc[p] ** 2 * inc[diag[ci[p]]]
                                        for p in range(adr[i], diag[i]))
    on = _dense(adr, ci, np.ones_like(a)) != 0
    A = _dense(adr, ci, a)
    A = A + np.tril(A, -1).T

    def pattern_error(f):
        L, D = _ldl(adr, ci, f)
        return np.abs((L @ np.diag(D) @ L.T - A)[on]).max()

    assert pattern_error(O.incomplete_ldl(adr, ci, a)) <= 1e-12 * np.abs(a).max()
    assert pattern_error(inc) > 1e-6 * np.abs(a).max()
