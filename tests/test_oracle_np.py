"""CPU checks that pin the oracle's nonstationary heat steps (cr_hex8_heat_cap,
cr_np_step_linear for the v-form, cr_np_step_dform for the d-form) before they
are trusted:

* bitwise against the reference: the capacity matrices of the harness (GEFEL's
  jac_3d, bf_lin_hex_3d, scprd, nnj in the loop order of
  linhext::capacity_matrix) and three steps of one_step_linear and of the
  linear_nonstat_solv_dform loop run on the reference's own symcomprow objects
  (mxv_scr, scalmat_scr, addmat_scr, cg_prec), from the committed fixture and,
  where oracle/_ref was built, live;
* mathematics: the exact unit-cube mass matrix (2x2x2 points integrate N_i N_j
  exactly), sum of entries = rho c V on a distorted element, the amplification
  factor (1 - (1-alpha) dt lambda) / (1 + alpha dt lambda) of the generalized
  trapezoidal rule on one unknown (both forms), the two forms agreeing to the
  solver tolerance, and the stationary solution reached by backward Euler with a
  long step.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, f64, i64
from ssf_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
HEX = np.array([[1, 1, 1], [0, 1, 1], [0, 0, 1], [1, 0, 1],
                [1, 1, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0]], float)


def _cap(X, c=1.0, dens=None):
    out = np.zeros(64)
    dens = np.ones(8) if dens is None else dens
    O.lib().cr_hex8_heat_cap(_ptr(f64(X[:, 0])), _ptr(f64(X[:, 1])), _ptr(f64(X[:, 2])), c,
                             _ptr(f64(dens)), _ptr(out))
    return out.reshape(8, 8)


def _heat_case():
    return mesh.heat_cube(4, seed=22)


def test_capacity_bitwise_fixture():
    g = np.load(GOLD)
    prob = _heat_case()
    cm = O.capacity_matrices(prob, g["heat_np_c"], g["heat_np_rho"])
    assert np.array_equal(cm, g["heat_np_cap"])


def test_np_steps_bitwise_fixture():
    g = np.load(GOLD)
    prob = _heat_case()
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    cm = O.capacity_matrices(prob, g["heat_np_c"], g["heat_np_rho"])
    aC = O.scr_assemble(prob.n, adr, ci, prob.elem_codes(), cm)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    for step in range(3):
        st, ani, ares, _ = O.np_step_linear(adr, ci, aK, aC, 0.5, 0.01, rhs[0], T, v, 5000, 1e-20)
        assert st == 1
        assert np.array_equal(T, g["heat_np_T"][step]) and np.array_equal(v, g["heat_np_v"][step])
        assert ani == g["heat_np_ani"][step] and ares == g["heat_np_ares"][step]


def test_np_dform_steps_bitwise_fixture():
    g = np.load(GOLD)
    prob = _heat_case()
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    cm = O.capacity_matrices(prob, g["heat_np_c"], g["heat_np_rho"])
    aC = O.scr_assemble(prob.n, adr, ci, prob.elem_codes(), cm)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    for step, (alpha, dt) in enumerate(((0.5, 0.01), (0.5, 0.01), (1.0, 0.02))):
        st, ani, ares, _ = O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], T, v, 5000,
                                            1e-20, form="d")
        assert st == 1
        assert np.array_equal(T, g["heat_npd_T"][step]) and np.array_equal(v, g["heat_npd_v"][step])
        assert ani == g["heat_npd_ani"][step] and ares == g["heat_npd_ares"][step]


def test_np_live_reference():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = O.ref()
    rng = np.random.default_rng(30)
    for _ in range(10):
        X = HEX + rng.uniform(-0.25, 0.25, HEX.shape)
        dens = rng.uniform(0.5, 2.0, 8)
        c = rng.uniform(0.1, 5.0)
        a = _cap(X, c, dens)
        b = np.zeros(64)
        R.ref_hex8_heat_cap(_ptr(f64(X[:, 0])), _ptr(f64(X[:, 1])), _ptr(f64(X[:, 2])), c,
                            _ptr(f64(dens)), _ptr(b))
        assert np.array_equal(a.ravel(), b)
    prob = mesh.heat_cube(5, seed=31)
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    c = rng.uniform(0.5, 2.0, prob.ne)
    cm = O.capacity_matrices(prob, c)
    cn = i64(prob.elem_codes())
    aC = O.scr_assemble(prob.n, adr, ci, cn, cm)
    hs = [R.ref_scr_create(prob.n, prob.ne, 8, _ptr(cn)) for _ in range(3)]
    R.ref_scr_set_a(hs[0], _ptr(f64(aK)))
    R.ref_scr_set_a(hs[1], _ptr(f64(aC)))
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    T2, v2 = T.copy(), v.copy()
    for alpha, dt in ((1.0, 0.05), (0.5, 0.02), (0.75, 0.1)):
        st, ani, ares, _ = O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], T, v, 5000, 1e-22)
        a2, r2 = ctypes.c_long(-1), ctypes.c_double(-1.0)
        R.ref_np_step_linear(hs[0], hs[1], hs[2], prob.n, alpha, dt, _ptr(f64(rhs[0]).copy()),
                             _ptr(T2), _ptr(v2), 5000, 1e-22, 1e-30, ctypes.byref(a2),
                             ctypes.byref(r2))
        assert np.array_equal(T, T2) and np.array_equal(v, v2)
        assert ani == a2.value and ares == r2.value
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    T2, v2 = T.copy(), v.copy()
    for alpha, dt in ((1.0, 0.05), (0.5, 0.02), (0.75, 0.1)):
        st, ani, ares, _ = O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], T, v, 5000, 1e-22,
                                            form="d")
        a2, r2 = ctypes.c_long(-1), ctypes.c_double(-1.0)
        R.ref_np_step_dform(hs[0], hs[1], hs[2], prob.n, alpha, dt, _ptr(f64(rhs[0]).copy()),
                            _ptr(T2), _ptr(v2), 5000, 1e-22, 1e-30, ctypes.byref(a2),
                            ctypes.byref(r2))
        assert np.array_equal(T, T2) and np.array_equal(v, v2)
        assert ani == a2.value and ares == r2.value
    for h in hs:
        R.ref_scr_free(h)


def test_capacity_unit_cube_closed_form():
    # int_0^1 N_i N_j dx = 1/3 (same end) or 1/6, per direction
    M = _cap(HEX, 1.0)
    s = 2 * HEX - 1
    ref = np.prod(np.where(s[:, None, :] == s[None, :, :], 1 / 3, 1 / 6), axis=2)
    assert np.abs(M - ref).max() <= 1e-15
    rng = np.random.default_rng(32)
    X = HEX * np.array([2.0, 0.5, 1.5]) + rng.uniform(-0.1, 0.1, HEX.shape)
    c, rho = 3.0, 0.7
    M = _cap(X, c, np.full(8, rho))
    # partition of unity: sum_ij C_ij = rho c V, V by a 4-point rule on the map
    g, w = np.polynomial.legendre.leggauss(4)
    V = 0.0
    for a, wa in zip(g, w):
        for b, wb in zip(g, w):
            for cc, wc in zip(g, w):
                dN = np.stack([s[:, 0] * (1 + s[:, 1] * b) * (1 + s[:, 2] * cc),
                               s[:, 1] * (1 + s[:, 0] * a) * (1 + s[:, 2] * cc),
                               s[:, 2] * (1 + s[:, 0] * a) * (1 + s[:, 1] * b)]) / 8
                V += np.linalg.det(dN @ X) * wa * wb * wc
    assert abs(M.sum() - rho * c * V) <= 1e-13 * rho * c * V
    assert np.allclose(M, M.T, rtol=0, atol=1e-16) and np.linalg.eigvalsh(M).min() > 0


def _one_unknown(k, m, alpha, dt, T0, steps, form="v", src=0.0):
    adr = i64([0, 1])
    ci = i64([0])
    T = np.array([T0])
    v = np.array([(src - k * T0) / m])  # consistent initial rate: m v = f - k T
    for _ in range(steps):
        O.np_step_linear(adr, ci, f64([k]), f64([m]), alpha, dt, f64([src]), T, v, 10, 1e-30,
                         form=form)
    return T[0], v[0]


@pytest.mark.parametrize("form", ["v", "d"])
@pytest.mark.parametrize("alpha", [0.5, 0.75, 1.0])
def test_np_amplification_factor(alpha, form):
    k, m, dt, T0, N = 3.0, 2.0, 0.4, 5.0, 12
    lam = k / m
    r = (1 - (1 - alpha) * dt * lam) / (1 + alpha * dt * lam)
    T, v = _one_unknown(k, m, alpha, dt, T0, N, form)
    assert abs(T - T0 * r ** N) <= 1e-13 * T0
    assert abs(v + lam * T) <= 1e-12 * T0  # the rate stays consistent


@pytest.mark.parametrize("alpha", [0.5, 1.0])
def test_np_forms_agree_with_a_source(alpha):
    """With a constant source f the exact discrete solution relaxes to f / k by
    the same factor in both forms: T_N - f/k = (T_0 - f/k) r^N."""
    k, m, dt, T0, N, src = 3.0, 2.0, 0.4, 5.0, 9, 4.5
    lam = k / m
    r = (1 - (1 - alpha) * dt * lam) / (1 + alpha * dt * lam)
    for form in ("v", "d"):
        T, _ = _one_unknown(k, m, alpha, dt, T0, N, form, src)
        assert abs(T - (src / k + (T0 - src / k) * r ** N)) <= 1e-13 * T0


def test_np_forms_agree_on_a_mesh():
    prob = _heat_case()
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    aC = O.scr_assemble(prob.n, adr, ci, prob.elem_codes(),
                        O.capacity_matrices(prob, np.full(prob.ne, 1.3)))
    Tv, vv = np.zeros(prob.n), np.zeros(prob.n)
    Td, vd = np.zeros(prob.n), np.zeros(prob.n)
    for alpha, dt in ((0.5, 0.01), (0.5, 0.01), (1.0, 0.05), (0.75, 0.02)):
        O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], Tv, vv, 5000, 1e-28)
        O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], Td, vd, 5000, 1e-28, form="d")
        assert np.abs(Td - Tv).max() <= 1e-11 * np.abs(Tv).max()
        assert np.abs(vd - vv).max() <= 1e-9 * np.abs(vv).max()


def test_np_backward_euler_reaches_the_stationary_solution():
    prob = _heat_case()
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    aC = O.scr_assemble(prob.n, adr, ci, prob.elem_codes(),
                        O.capacity_matrices(prob, np.ones(prob.ne)))
    xs, _, _, _ = O.scr_cg_prec(adr, ci, aK, rhs[0], 5000, 1e-26)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    for _ in range(3):
        O.np_step_linear(adr, ci, aK, aC, 1.0, 1e9, rhs[0], T, v, 5000, 1e-26)
    assert np.abs(T - xs).max() <= 1e-7 * np.abs(xs).max()


def test_np_dform_backward_euler_reaches_the_stationary_solution():
    prob = _heat_case()
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    aC = O.scr_assemble(prob.n, adr, ci, prob.elem_codes(),
                        O.capacity_matrices(prob, np.ones(prob.ne)))
    xs, _, _, _ = O.scr_cg_prec(adr, ci, aK, rhs[0], 5000, 1e-26)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    O.np_step_linear(adr, ci, aK, aC, 1.0, 1e9, rhs[0], T, v, 5000, 1e-26, form="d")
    assert np.abs(T - xs).max() <= 1e-7 * np.abs(xs).max()


def test_np_checks_catch_planted_mistakes():
    k, m, dt, T0, N, alpha = 3.0, 2.0, 0.4, 5.0, 12, 0.75
    lam = k / m
    good, _ = _one_unknown(k, m, alpha, dt, T0, N)
    # predictor with alpha and 1 - alpha swapped
    r_swapped = (1 - alpha * dt * lam) / (1 + alpha * dt * lam)
    assert abs(good - T0 * r_swapped ** N) > 1e-3 * T0
    # capacity without the density
    M1 = _cap(HEX, 2.0, np.full(8, 3.0))
    assert abs(M1.sum() - 2.0 * 3.0) <= 1e-14 and abs(_cap(HEX, 2.0).sum() - 6.0) > 1.0
    # d-form right-hand side without its alpha dt factor: the source enters with
    # weight 1 / (alpha dt) too much and the relaxation limit moves off f / k
    src, T, N = 4.5, 5.0, 40
    d, v = T, (src - k * T) / m
    for _ in range(N):
        dd = d + (1 - alpha) * dt * v
        T = (src + m * dd) / (k * dt * alpha + m)
        v = (T - dd) / dt / alpha
        d = T
    assert abs(T - src / k) > 1e-2
    assert abs(_one_unknown(k, m, alpha, dt, 5.0, N, "d", src)[0] - src / k) <= 1e-6
