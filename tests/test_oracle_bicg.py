"""CPU checks that pin the oracle's biconjugate gradients (cr_cr_bicg, restating
comprow::bicg GEFEL/cr.cpp:1353-1419, and cr_mtxv_cr, restating comprow::mtxv_cr
:843-858) before they are trusted:

* bitwise against the reference: the committed fixture (the reference's own
  comprow::bicg and mtxv_cr on nonsymmetric values of the heat and elasticity
  patterns, from x = 0 and for 5 iterations from a start vector) and, where
  oracle/_ref was built, live on further seeds;
* mathematics: the transpose identity b.(A^T c) = (A b).c; the solution of a
  direct solve on small dense nonsymmetric systems; on a symmetric matrix the
  BiCG iterates are CG's (same solution, iteration count within a few); the
  zero right-hand side and ni-spent conventions of the reference.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, f64, i64
from ssf_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
CASES = {"heat": lambda: mesh.heat_cube(4, seed=22), "elast": lambda: mesh.elastic_column(2, 2, 3, seed=21)}


def _pattern(prob):
    return O.cr_pattern(prob.n, prob.elem_codes())


@pytest.mark.parametrize("name", ["heat", "elast"])
def test_bicg_bitwise_fixture(name):
    g = np.load(GOLD)
    prob = CASES[name]()
    adr, ci = _pattern(prob)
    a, y = g[f"{name}_bicg_a"], g[f"{name}_bicg_y"]
    assert np.array_equal(O.mtxv_cr(adr, ci, a, g[f"{name}_bicg_b"]), g[f"{name}_bicg_mtxv"])
    x, ani, ares, st = O.cr_bicg(adr, ci, a, y, 2000, 1e-24)
    assert st == 1
    assert np.array_equal(x, g[f"{name}_bicg_x"])
    assert ani == g[f"{name}_bicg_ani"] and ares == g[f"{name}_bicg_ares"]
    x5, ani5, ares5, st5 = O.cr_bicg(adr, ci, a, y, 5, 1e-24, x0=g[f"{name}_bicg_x0"])
    assert st5 == 0 and ani5 == 5 == g[f"{name}_bicg_ani5"]
    assert np.array_equal(x5, g[f"{name}_bicg_x5"]) and ares5 == g[f"{name}_bicg_ares5"]


def test_bicg_live_reference():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = O.ref()
    for seed in range(4):
        prob = mesh.heat_cube(5, seed=60 + seed)
        adr, ci, a, rhs = O.assemble_problem(prob, "cr")
        an = O.nonsymmetric_values(adr, ci, a, 0.2 + 0.1 * seed, seed)
        cn = i64(prob.elem_codes())
        h = R.ref_cr_create(prob.n, prob.ne, 8, _ptr(cn))
        R.ref_cr_set_a(h, _ptr(f64(an)))
        for x0 in (None, np.random.default_rng(seed).standard_normal(prob.n)):
            x, ani, ares, _ = O.cr_bicg(adr, ci, an, rhs[0], 500, 1e-22, x0=x0)
            x2 = np.zeros(prob.n) if x0 is None else x0.copy()
            a2, r2 = ctypes.c_long(-1), ctypes.c_double(-1.0)
            R.ref_cr_bicg(h, _ptr(x2), _ptr(f64(rhs[0]).copy()), 500, 1e-22, ctypes.byref(a2),
                          ctypes.byref(r2), 1e-30, 0 if x0 is None else 1)
            assert np.array_equal(x, x2) and ani == a2.value and ares == r2.value
        R.ref_cr_free(h)


def _dense_cr(A):
    n = A.shape[0]
    adr = i64(np.arange(n + 1) * n)
    ci = i64(np.tile(np.arange(n), n))
    return adr, ci, f64(A.ravel())


def test_transpose_identity():
    prob = mesh.heat_cube(4, seed=61)
    adr, ci, a, _ = O.assemble_problem(prob, "cr")
    an = O.nonsymmetric_values(adr, ci, a, 0.5, 3)
    rng = np.random.default_rng(4)
    b, c = rng.standard_normal(prob.n), rng.standard_normal(prob.n)
    Ab = np.zeros(prob.n)
    O.lib().cr_mxv_cr(prob.n, _ptr(adr), _ptr(ci), _ptr(an), _ptr(b), _ptr(Ab))
    lhs = float(b @ O.mtxv_cr(adr, ci, an, c))
    assert abs(lhs - float(Ab @ c)) <= 1e-12 * abs(lhs)
    # and it is the transpose, not the matrix again
    A = np.zeros((prob.n, prob.n))
    rows = np.repeat(np.arange(prob.n), np.diff(adr))
    A[rows, ci] = an
    assert np.allclose(O.mtxv_cr(adr, ci, an, c), A.T @ c, rtol=0, atol=1e-12 * np.abs(A).max())
    assert not np.allclose(A.T @ c, A @ c)


@pytest.mark.parametrize("n", [3, 6, 10])
def test_bicg_small_dense_direct(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n)) + n * np.eye(n)  # nonsymmetric, well conditioned
    y = rng.standard_normal(n)
    adr, ci, a = _dense_cr(A)
    x, ani, ares, st = O.cr_bicg(adr, ci, a, y, 10 * n, 1e-28)
    assert st == 1
    assert np.abs(x - np.linalg.solve(A, y)).max() <= 1e-12 * np.abs(x).max()
    assert ani <= 2 * n  # finite termination in exact arithmetic after n steps


def test_bicg_on_a_symmetric_matrix_is_cg():
    prob = mesh.heat_cube(5, seed=62)
    adr, ci, a, rhs = O.assemble_problem(prob, "cr")
    xb, anib, _, stb = O.cr_bicg(adr, ci, a, rhs[0], 2000, 1e-26)
    # CG with the same stopping rule on the same system
    xc, anic, _, stc = O.cr_cg_prec_noprec(adr, ci, a, rhs[0], 2000, 1e-13)
    assert stb == 1 and stc == 1
    assert np.abs(xb - xc).max() <= 1e-9 * np.abs(xc).max()
    assert abs(anib - anic) <= 3


def test_bicg_conventions():
    adr, ci, a = _dense_cr(np.array([[4.0, 1.0], [-2.0, 3.0]]))
    # zero right-hand side: x untouched, ani 0, ares = y.y
    x0 = np.array([7.0, -1.0])
    x, ani, ares, st = O.cr_bicg(adr, ci, a, np.zeros(2), 10, 1e-20, x0=x0)
    assert st == 3 and ani == 0 and ares == 0.0 and np.array_equal(x, x0)
    # ni spent: ani = ni, ares = the last r.r
    x, ani, ares, st = O.cr_bicg(adr, ci, a, np.array([1.0, 2.0]), 1, 1e-40)
    assert st == 0 and ani == 1 and ares > 0.0


def test_bicg_checks_catch_planted_mistakes():
    """A BiCG that multiplies the shadow direction by A instead of A^T (a
    plausible slip) loses finite termination on a small nonsymmetric system."""
    n = 8
    rng = np.random.default_rng(9)
    A = rng.standard_normal((n, n)) + 2.5 * np.eye(n)
    y = rng.standard_normal(n)

    def bicg(transposed, iters):
        x = np.zeros(n)
        r = y.copy()
        rr, d, dd = r.copy(), r.copy(), r.copy()
        nom = r @ rr
        for _ in range(iters):
            p = A @ d
            pp = (A.T if transposed else A) @ dd
            alpha = nom / (dd @ p)
            x += alpha * d
            r -= alpha * p
            rr -= alpha * pp
            nomn = r @ rr
            beta = nomn / nom
            nom = nomn
            d = r + beta * d
            dd = rr + beta * dd
        return x

    xs = np.linalg.solve(A, y)
    assert np.abs(bicg(True, n) - xs).max() <= 1e-8 * np.abs(xs).max()
    assert np.abs(bicg(False, n) - xs).max() > 1e-4 * np.abs(xs).max()
    adr, ci, a = _dense_cr(A)
    x, _, _, _ = O.cr_bicg(adr, ci, a, y, n, 1e-40)
    assert np.abs(x - xs).max() <= 1e-8 * np.abs(xs).max()
