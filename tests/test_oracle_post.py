"""CPU checks that pin the oracle's post-solve element pass (cr_eldispl, cr_*_ip,
cr_elem_results_mesh) before it is trusted:

* bitwise against the reference's own GEFEL routines composed in the element
  loops' order (oracle/ref_harness.cpp): the committed fixture
  tests/golden/ref_vectors.npz (post_*) and, where oracle/_ref was built, live
  on fresh elements and a fresh mesh;
* the code-number rule of eldispl on a hand-built table;
* mathematics: linear fields give the exact constant strain / gradient at every
  integration point (patch test), with the closed-form D eps, plane-stress eps_z
  and q = -k grad T; rigid-body motions give zero strain and zero force; the
  element forces equal K_e r with the independently pinned K_e (and -K_e r for
  the internal fluxes of heat); the assembled forces equal A x - lift.
"""
import os

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, f64, i64
from ssf_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
CASES = {
    "elast": lambda: mesh.elastic_column(2, 2, 3, seed=21),
    "cant": lambda: mesh.cantilever(4, 2, 2),
    "heat": lambda: mesh.heat_cube(4, seed=22),
    "quad": lambda: mesh.plate(6, 3, seed=23, nrhs=1),
}
HEX = np.array([[1, 1, 1], [0, 1, 1], [0, 0, 1], [1, 0, 1],
                [1, 1, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0]], float)
QUAD = np.array([[1, 1], [0, 1], [0, 0], [1, 0]], float)


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def _hex(rng, amp=0.15):
    return HEX + rng.uniform(-amp, amp, HEX.shape)


def _quad(rng, amp=0.1):
    return QUAD + rng.uniform(-amp, amp, QUAD.shape)


def _ip(kind, X, ma, mb, th, r):
    if kind == 2:
        return O.elem_ip(2, X[:, 0], X[:, 1], np.zeros(4), ma, mb, th, r)
    return O.elem_ip(kind, X[:, 0], X[:, 1], X[:, 2], ma, mb, th, r)


def _ke(kind, X, ma, mb, th):
    k = np.zeros(576 if kind == 0 else 64)
    if kind == 0:
        O.lib().cr_hex8_elast_ke(_ptr(f64(X[:, 0])), _ptr(f64(X[:, 1])), _ptr(f64(X[:, 2])), ma,
                                 mb, _ptr(k))
    elif kind == 1:
        O.lib().cr_hex8_heat_ke(_ptr(f64(X[:, 0])), _ptr(f64(X[:, 1])), _ptr(f64(X[:, 2])), ma,
                                _ptr(k))
    else:
        O.lib().cr_quad4_pstress_ke(_ptr(f64(X[:, 0])), _ptr(f64(X[:, 1])), ma, mb,
                                    _ptr(f64(np.full(4, th))), _ptr(k))
    nd = 24 if kind == 0 else 8
    return k.reshape(nd, nd)


def _d3(E, nu):
    """isotropic 6x6 D in Voigt order x, y, z, yz, xz, xy with engineering shears"""
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[[0, 1, 2], [0, 1, 2]] += 2 * mu
    D[3:, 3:] = np.eye(3) * mu
    return D


# ------------------------------------------------------- bitwise vs reference

@pytest.mark.parametrize("name", list(CASES))
def test_post_pass_bitwise(gold, name):
    prob = CASES[name]()
    eps, sig, ifor = O.elem_results(prob, gold[f"{name}_post_x"])
    assert np.array_equal(eps, gold[f"{name}_post_eps"])
    assert np.array_equal(sig, gold[f"{name}_post_sig"])
    assert np.array_equal(ifor, gold[f"{name}_post_ifor"])


def test_post_pass_live_reference():
    """Oracle vs the reference build on elements and a mesh outside the fixture."""
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = O.ref()
    rng = np.random.default_rng(8)
    for _ in range(20):
        X = _hex(rng, 0.25)
        x, y, z = f64(X[:, 0]), f64(X[:, 1]), f64(X[:, 2])
        e, nu, k = rng.uniform(1e9, 1e11), rng.uniform(0.1, 0.45), rng.uniform(0.1, 10)
        r = rng.standard_normal(24) * 1e-3
        a = [np.zeros(48), np.zeros(48), np.zeros(24)]
        b = [np.zeros(48), np.zeros(48), np.zeros(24)]
        O.lib().cr_hex8_elast_ip(_ptr(x), _ptr(y), _ptr(z), e, nu, _ptr(r), *map(_ptr, a))
        R.ref_hex8_elast_ip(_ptr(x), _ptr(y), _ptr(z), e, nu, _ptr(r), *map(_ptr, b))
        assert all(np.array_equal(p, q) for p, q in zip(a, b))
        t = f64(r[:8] * 100)
        a = [np.zeros(24), np.zeros(24), np.zeros(8)]
        b = [np.zeros(24), np.zeros(24), np.zeros(8)]
        O.lib().cr_hex8_heat_ip(_ptr(x), _ptr(y), _ptr(z), k, _ptr(t), *map(_ptr, a))
        R.ref_hex8_heat_ip(_ptr(x), _ptr(y), _ptr(z), k, _ptr(t), *map(_ptr, b))
        assert all(np.array_equal(p, q) for p, q in zip(a, b))
        th = f64(rng.uniform(0.5, 2, 4))
        a = [np.zeros(16), np.zeros(16), np.zeros(8)]
        b = [np.zeros(16), np.zeros(16), np.zeros(8)]
        O.lib().cr_quad4_pstress_ip(_ptr(x[:4]), _ptr(y[:4]), e, nu, _ptr(th), _ptr(r[:8]),
                                    *map(_ptr, a))
        R.ref_quad4_pstress_ip(_ptr(x[:4]), _ptr(y[:4]), e, nu, _ptr(th), _ptr(f64(r[:8])),
                               *map(_ptr, b))
        assert all(np.array_equal(p, q) for p, q in zip(a, b))
    prob = mesh.elastic_column(3, 2, 4, seed=78)
    xs = rng.standard_normal(prob.n) * 1e-4
    _, _, ifor = O.elem_results(prob, xs)
    cn = i64(prob.elem_codes())
    fr = np.zeros(prob.n)
    for e in range(prob.ne):
        nodes = prob.conn[e]
        r = np.zeros(24)
        O.lib().cr_eldispl(24, _ptr(cn[e].copy()), _ptr(xs), None, _ptr(r))
        out = [np.zeros(48), np.zeros(48), np.zeros(24)]
        R.ref_hex8_elast_ip(*(_ptr(f64(prob.coords[nodes, d])) for d in range(3)), prob.e[e],
                            prob.nu[e], _ptr(r), *map(_ptr, out))
        R.ref_locglob(_ptr(fr), _ptr(out[2]), _ptr(cn[e].copy()), 24)
    assert np.array_equal(ifor, fr)


def test_eldispl_code_number_rule():
    cn = i64([3, 0, -2, 1, -1, 2])
    lhs = f64([10.0, 20.0, 30.0])
    pv = f64([-5.0, -7.0])
    r = np.zeros(6)
    O.lib().cr_eldispl(6, _ptr(cn), _ptr(lhs), _ptr(pv), _ptr(r))
    assert list(r) == [30.0, 0.0, -7.0, 10.0, -5.0, 20.0]


# ----------------------------------------------------------------- mathematics

def _patch_elast(eps, sig, G, E, nu):
    """largest deviation from the exact constant strain of u = G x and its D eps"""
    ev = np.array([G[0, 0], G[1, 1], G[2, 2], G[1, 2] + G[2, 1], G[0, 2] + G[2, 0],
                   G[0, 1] + G[1, 0]])
    sv = _d3(E, nu) @ ev
    return (np.abs(eps - ev).max() / np.abs(ev).max(), np.abs(sig - sv).max() / np.abs(sv).max())


def test_hex8_elast_patch_rigid_and_ke():
    rng = np.random.default_rng(11)
    for _ in range(5):
        X = _hex(rng)
        E, nu = rng.uniform(1e9, 5e10), rng.uniform(0.15, 0.4)
        G = rng.standard_normal((3, 3)) * 1e-3
        c = rng.standard_normal(3)
        r = (X @ G.T + c).ravel()
        eps, sig, f = _ip(0, X, E, nu, 1.0, r)
        de, ds = _patch_elast(eps, sig, G, E, nu)
        assert de <= 1e-12 and ds <= 1e-12, (de, ds)
        # rigid-body motion: translation + small rotation u = c + w x X
        w = rng.standard_normal(3) * 1e-3
        r = (c + np.cross(w, X)).ravel()
        eps, sig, f = _ip(0, X, E, nu, 1.0, r)
        assert np.abs(eps).max() <= 1e-15 * 10 and np.abs(f).max() <= 1e-12 * E * 1e-3
        # element forces = K_e r
        r = rng.standard_normal(24) * 1e-3
        K = _ke(0, X, E, nu, 1.0)
        _, _, f = _ip(0, X, E, nu, 1.0, r)
        assert np.abs(f - K @ r).max() <= 1e-12 * np.abs(K).max() * np.abs(r).max() * 24


def test_hex8_heat_patch_and_ke():
    rng = np.random.default_rng(12)
    for _ in range(5):
        X = _hex(rng)
        k = rng.uniform(0.1, 10.0)
        g = rng.standard_normal(3)
        r = X @ g + 3.0
        grad, flux, f = _ip(1, X, k, 0.0, 1.0, r)
        assert np.abs(grad - g).max() <= 1e-12 * np.abs(g).max()
        assert np.abs(flux + k * g).max() <= 1e-12 * k * np.abs(g).max()
        # internal fluxes of q = -k grad T: f = -K_e r (constants give zero)
        r = rng.standard_normal(8)
        K = _ke(1, X, k, 0.0, 1.0)
        _, _, f = _ip(1, X, k, 0.0, 1.0, r)
        assert np.abs(f + K @ r).max() <= 1e-12 * np.abs(K).max() * 8
        _, _, f = _ip(1, X, k, 0.0, 1.0, np.full(8, 5.0))
        assert np.abs(f).max() <= 1e-13 * k * 5


def test_quad4_patch_rigid_and_ke():
    rng = np.random.default_rng(13)
    for _ in range(5):
        X = _quad(rng)
        E, nu, t = rng.uniform(1e9, 5e10), rng.uniform(0.15, 0.4), rng.uniform(0.5, 2.0)
        G = rng.standard_normal((2, 2)) * 1e-3
        r = (X @ G.T + rng.standard_normal(2)).ravel()
        eps, sig, f = _ip(2, X, E, nu, t, r)
        ex, ey, gxy = G[0, 0], G[1, 1], G[0, 1] + G[1, 0]
        c = E / (1 - nu * nu)
        ev = np.array([ex, ey, gxy, -nu / (1 - nu) * (ex + ey)])
        sv = np.array([c * (ex + nu * ey), c * (ey + nu * ex), E / 2 / (1 + nu) * gxy, 0.0])
        assert np.abs(eps - ev).max() <= 1e-12 * np.abs(ev).max()
        assert np.abs(sig - sv).max() <= 1e-12 * np.abs(sv).max()
        assert np.all(sig[:, 3] == 0.0)
        w = rng.standard_normal() * 1e-3
        r = (np.stack([-w * X[:, 1], w * X[:, 0]], 1) + rng.standard_normal(2)).ravel()
        eps, _, f = _ip(2, X, E, nu, t, r)
        assert np.abs(eps).max() <= 1e-14 and np.abs(f).max() <= 1e-12 * E * 1e-3
        r = rng.standard_normal(8) * 1e-3
        K = _ke(2, X, E, nu, t)
        _, _, f = _ip(2, X, E, nu, t, r)
        assert np.abs(f - K @ r).max() <= 1e-12 * np.abs(K).max() * np.abs(r).max() * 8


def test_negative_jacobian_keeps_the_reference_sign():
    """elem_integration multiplies by the signed Jacobian (MEFEL/SRC/linhex.cpp:2184,
    plelemlq.cpp:2350) while K_e and the heat internal fluxes use |J|: on an element
    numbered the other way round the elastic forces are -K_e r."""
    rng = np.random.default_rng(14)
    X = _hex(rng)[[4, 5, 6, 7, 0, 1, 2, 3]]  # top and bottom swapped: J < 0
    r = rng.standard_normal(24) * 1e-3
    K = _ke(0, X, 1e10, 0.3, 1.0)
    _, _, f = _ip(0, X, 1e10, 0.3, 1.0, r)
    assert np.abs(f + K @ r).max() <= 1e-12 * np.abs(K).max() * 24 * np.abs(r).max()
    K = _ke(1, X, 2.0, 0.0, 1.0)
    _, _, f = _ip(1, X, 2.0, 0.0, 1.0, r[:8])
    assert np.abs(f + K @ r[:8]).max() <= 1e-12 * np.abs(K).max() * 8 * np.abs(r).max()


@pytest.mark.parametrize("name", list(CASES))
def test_assembled_forces_equal_a_x_minus_lift(name):
    prob = CASES[name]()
    adr, ci, a, _ = O.assemble_problem(prob, "scr")
    x = np.random.default_rng(15).standard_normal(prob.n) * 1e-3
    _, _, ifor = O.elem_results(prob, x)
    cn = prob.elem_codes()
    lift = (O.dirichlet_lift(prob, cn, O.element_matrices(prob)) if prob.pv.size
            else np.zeros(prob.n))
    ax = O.mxv_scr(adr, ci, a, x) - lift
    sign = -1.0 if prob.kind == mesh.KIND_HEX8_HEAT else 1.0
    assert np.abs(ifor - sign * ax).max() <= 1e-12 * np.abs(ax).max()


def test_post_checks_catch_planted_mistakes():
    """Mistakes planted in copies of the oracle's outputs fail the checks above."""
    rng = np.random.default_rng(16)
    X = _hex(rng)
    E, nu = 2e10, 0.3
    G = rng.standard_normal((3, 3)) * 1e-3
    eps, sig, _ = _ip(0, X, E, nu, 1.0, (X @ G.T).ravel())
    assert max(_patch_elast(eps, sig, G, E, nu)) <= 1e-12
    swapped = eps[:, [0, 1, 2, 4, 3, 5]]  # gamma_yz <-> gamma_xz
    assert _patch_elast(swapped, sig, G, E, nu)[0] > 1e-3
    half = eps.copy()
    half[:, 3:] *= 0.5  # tensor instead of engineering shears
    assert _patch_elast(half, sig, G, E, nu)[0] > 1e-3
    # flux without the minus sign
    g = rng.standard_normal(3)
    _, flux, _ = _ip(1, X, 2.0, 0.0, 1.0, X @ g)
    assert np.abs(-flux + 2.0 * g).max() > 1.0
    # quad4 forces without the thickness
    Xq = _quad(rng)
    r = rng.standard_normal(8) * 1e-3
    K = _ke(2, Xq, 1e10, 0.25, 0.6)
    _, _, f1 = _ip(2, Xq, 1e10, 0.25, 1.0, r)
    assert np.abs(f1 - K @ r).max() > 0.1 * np.abs(K @ r).max()
    # |J| instead of the signed Jacobian in the elastic forces of a reversed
    # element would give +K_e r
    Xr = X[[4, 5, 6, 7, 0, 1, 2, 3]]
    Kr = _ke(0, Xr, E, nu, 1.0)
    rr = rng.standard_normal(24) * 1e-3
    _, _, fr = _ip(0, Xr, E, nu, 1.0, rr)
    tol = 1e-12 * np.abs(Kr).max() * 24 * np.abs(rr).max()
    assert np.abs(fr + Kr @ rr).max() <= tol
    assert np.abs(-fr + Kr @ rr).max() > 1e6 * tol
