"""CPU checks that pin the oracle's Newton-Raphson load control (cr_newton_raphson,
restating gnewton_raphson MEFEL/SRC/newtonraph.cpp:51-268 for linear elastic
elements). The reference function runs on MEFEL's global problem objects, which
the GEFEL-only harness does not build, so it is pinned by what its control flow
and linear elasticity fix:

* superposition: at the attained lambda, ra solves K ra = fc + lambda fp;
* the increment sequence: dlambda doubles after every second increment accepted
  without inner iterations, is capped at maxincrnr, and is cut to reach lambdar;
* rejection: with errnr below the solver's reach every increment is rejected,
  ra and lambda are restored and dlambda halves down to minincrnr, where the
  loop stops;
* inner iterations: a loose solver needs several corrections per increment, and
  each one reduces the unbalanced forces fa - fi (internal forces from the
  element pass), so ra still converges to the direct solution;
* the two supported residual norms (relative to the load, absolute).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sl

import oracle as O
from ssf_amd import mesh


def _case(make=lambda: mesh.cantilever(4, 2, 2)):
    prob = make()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    A = sp.csr_matrix((a, ci, adr), shape=(prob.n, prob.n))
    A = (sp.tril(A) + sp.tril(A, -1).T).tocsc()
    return prob, adr, ci, a, rhs[0], A


def test_nr_superposition_and_increment_sequence():
    prob, adr, ci, a, fp, A = _case()
    fc = 0.3 * np.roll(fp, 7)
    lam, ra, fa, h = O.newton_raphson(prob, adr, ci, a, fp, fc=fc, incrnr=0.1, maxincrnr=0.35,
                                      check_lambdar=True, lambdar=1.3, errnr=1e-6)
    assert lam == pytest.approx(1.3, abs=1e-15)
    assert np.abs(ra - sl.spsolve(A, fc + lam * fp)).max() <= 1e-10 * np.abs(ra).max()
    assert np.abs(fa - (fc + lam * fp)).max() <= 1e-12 * np.abs(fa).max()
    assert (h["acc"] == 1).all() and (h["nsol"] == 1).all()
    # 0.1, 0.1, then doubled after every second accepted increment, capped at 0.35
    d = h["dlam"]
    assert d[0] == d[1] == 0.1 and d[2] == d[3] == 0.2 and d[4] == 0.35
    assert np.all(d[4:-1] == 0.35) or d[-1] < 0.35
    assert np.cumsum(d)[-1] == pytest.approx(1.3, abs=1e-14)


def test_nr_rejection_halves_down_to_the_minimum():
    prob, adr, ci, a, fp, A = _case()
    ra0 = np.zeros(prob.n)
    lam, ra, fa, h = O.newton_raphson(prob, adr, ci, a, fp, incrnr=0.5, minincrnr=0.07,
                                      niilnr=2, errnr=1e-30, ra=ra0)
    assert lam == 0.0 and not ra.any() and np.abs(fa).max() <= 1e-9 * np.abs(fp).max()
    assert (h["acc"] == 0).all() and (h["nsol"] == 3).all()  # 1 + niilnr solves each
    assert list(h["dlam"]) == [0.5, 0.25, 0.125, 0.07]  # 0.0625 < min -> 0.07, then stop


def test_nr_inner_iterations_with_a_loose_solver():
    prob, adr, ci, a, fp, A = _case(lambda: mesh.elastic_column(2, 2, 4, seed=50))
    fp = fp if np.abs(fp).max() > 0 else np.ones(prob.n)
    lam, ra, fa, h = O.newton_raphson(prob, adr, ci, a, fp, incrnr=0.5, check_lambdar=True,
                                      lambdar=1.0, errnr=1e-9, sres=1e-6, niilnr=20)
    assert lam == pytest.approx(1.0, abs=1e-15) and (h["acc"] == 1).all()
    # every increment that moved the load needed corrections (the last one can be
    # the zero-size increment that closes at lambdar)
    assert (h["nsol"][h["dlam"] > 0] > 1).all()
    assert np.abs(ra - sl.spsolve(A, fa)).max() <= 1e-7 * np.abs(ra).max()


def test_nr_norms():
    prob, adr, ci, a, fp, A = _case()
    kw = dict(incrnr=1.0, niilnr=0, sres=1e-8, ninr=1, minincrnr=1.0)
    # relative to the load: ~1e-4 residual of a 1e-8 squared-residual solve passes 1e-2
    lam, _, _, h = O.newton_raphson(prob, adr, ci, a, fp, errnr=1e-2, normt=1, **kw)
    assert lam == 1.0 and h["acc"][0] == 1
    # absolute: the same residual measured in force units (loads ~1e5) fails it
    lam, _, _, h = O.newton_raphson(prob, adr, ci, a, fp, errnr=1e-2, normt=4, **kw)
    assert lam == 0.0 and h["acc"][0] == 0


def test_nr_checks_catch_planted_mistakes():
    """Increments that subtract the internal forces with the wrong sign (fb = fa +
    fi), a plausible slip, do not reach the direct solution."""
    prob, adr, ci, a, fp, A = _case()
    ra = np.zeros(prob.n)
    for _ in range(5):
        _, _, fi = O.elem_results(prob, ra)
        dr, _, _, _ = O.scr_cg_prec(adr, ci, a, fp + fi, 5000, 1e-24)
        ra = ra + dr
    xs = sl.spsolve(A, fp)
    assert np.abs(ra - xs).max() > 1e-3 * np.abs(xs).max()
    lam, ra_good, _, _ = O.newton_raphson(prob, adr, ci, a, fp, incrnr=1.0, ninr=1, errnr=1e-6)
    assert np.abs(ra_good - xs).max() <= 1e-10 * np.abs(xs).max()
