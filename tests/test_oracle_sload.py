"""CPU checks that pin the oracle's surface loads (cr_hex8_surface_load,
cr_surface_load_mesh: linhex::node_forces_surf MEFEL/SRC/linhex.cpp:2543 and the
record loop of loadcase::assemble MEFEL/SRC/loadcase.cpp:572-597):

* bitwise against the reference's GEFEL routines composed in node_forces_surf's
  loop order (fixture of records on the elasticity cases, and live on random
  distorted elements with every face flag);
* mathematics: a uniform traction t on a rectangular face gives t A / 4 at each
  of its four nodes and nothing elsewhere; the nodal forces of any face sum to
  the integral of the bilinear traction over the face (2x2 points are exact on
  a parallelogram); values given in the face's local axes (ls = 2) equal the
  global values t = [g1 g2 g3] t_local (ls = 1).
"""
import os

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, f64, i64
from ssf_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
HEX = np.array([[1, 1, 1], [0, 1, 1], [0, 0, 1], [1, 0, 1],
                [1, 1, 0], [0, 1, 0], [0, 0, 0], [1, 0, 0]], float)
FACE_NODES = [(0, 3, 7, 4), (1, 0, 4, 5), (2, 1, 5, 6), (3, 2, 6, 7), (0, 1, 2, 3), (4, 7, 6, 5)]
CASES = {"elast": lambda: mesh.elastic_column(2, 2, 3, seed=21),
         "cant": lambda: mesh.cantilever(4, 2, 2)}


def _nf(X, ls, nv):
    return O.hex8_surface_load(X[:, 0], X[:, 1], X[:, 2], ls, nv)


@pytest.mark.parametrize("name", ["elast", "cant"])
def test_surface_load_bitwise_fixture(name):
    g = np.load(GOLD)
    prob = CASES[name]()
    out = O.surface_load(prob, g[f"{name}_sload_eid"], g[f"{name}_sload_ls"],
                         g[f"{name}_sload_nv"], g[f"{name}_sload_rhs0"])
    assert np.array_equal(out, g[f"{name}_sload_rhs"])


def test_surface_load_live_reference():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = O.ref()
    rng = np.random.default_rng(90)
    for _ in range(200):
        X = HEX * rng.uniform(0.3, 3.0, 3) + rng.uniform(-0.2, 0.2, HEX.shape)
        ls = rng.integers(0, 3, 6)
        nv = rng.standard_normal(72) * 10.0 ** rng.integers(-3, 6)
        b = np.zeros(24)
        R.ref_hex8_surface_load(_ptr(f64(X[:, 0])), _ptr(f64(X[:, 1])), _ptr(f64(X[:, 2])),
                                _ptr(i64(ls)), _ptr(f64(nv)), _ptr(b))
        assert np.array_equal(_nf(X, ls, nv), b)


@pytest.mark.parametrize("face", range(6))
def test_uniform_traction_on_a_box(face):
    dims = np.array([2.0, 0.5, 1.5])
    X = HEX * dims
    area = {0: dims[1] * dims[2], 1: dims[0] * dims[2], 2: dims[1] * dims[2],
            3: dims[0] * dims[2], 4: dims[0] * dims[1], 5: dims[0] * dims[1]}[face]
    t = np.array([3.0, -1.0, 7.0])
    ls = np.zeros(6, np.int64)
    ls[face] = 1
    nv = np.zeros(72)
    nv[12 * face:12 * face + 12] = np.tile(t, 4)
    nf = _nf(X, ls, nv).reshape(8, 3)
    for a in range(8):
        want = t * area / 4 if a in FACE_NODES[face] else np.zeros(3)
        assert np.abs(nf[a] - want).max() <= 1e-14 * np.abs(t).max() * area


def test_totals_and_local_axes():
    rng = np.random.default_rng(91)
    for face in range(6):
        # a parallelepiped: faces are parallelograms, 2x2 points integrate exactly
        J = np.eye(3) + rng.uniform(-0.3, 0.3, (3, 3))
        X = HEX @ J.T
        tn = rng.standard_normal((4, 3))
        ls = np.zeros(6, np.int64)
        ls[face] = 1
        nv = np.zeros(72)
        nv[12 * face:12 * face + 12] = tn.ravel()
        nf = _nf(X, ls, nv).reshape(8, 3)
        fn = FACE_NODES[face]
        P = X[list(fn)]
        # area of the parallelogram face and the mean of the bilinear field
        area = np.linalg.norm(np.cross(P[1] - P[0], P[3] - P[0]))
        assert np.abs(nf.sum(axis=0) - area * tn.mean(axis=0)).max() <= 1e-12 * area * np.abs(tn).max()
        # local axes of locglob_nodeval and the same load given in them
        gpairs = [(7, 4, 7, 3), (4, 5, 4, 0), (5, 6, 5, 1), (6, 7, 6, 2), (2, 3, 2, 1), (6, 5, 6, 7)]
        q0, q1, r0, r1 = gpairs[face]
        g1 = X[q1] - X[q0]
        g1 /= np.linalg.norm(g1)
        g2 = X[r1] - X[r0]
        g2 = g2 - (g1 @ g2) * g1
        g2 /= np.linalg.norm(g2)
        T = np.stack([g1, g2, np.cross(g1, g2)], axis=1)
        ls[face] = 2
        nv_local = np.zeros(72)
        nv_local[12 * face:12 * face + 12] = (tn @ T).ravel()  # T^T t per node
        nf2 = _nf(X, ls, nv_local).reshape(8, 3)
        assert np.abs(nf2 - nf).max() <= 1e-12 * np.abs(nf).max()


def test_surface_load_checks_catch_planted_mistakes():
    """Nodal values placed on the face in the wrong node order (a plausible slip)
    fail the live comparison and the total-force check on a distorted face."""
    X = HEX * np.array([2.0, 1.0, 1.0]) + np.array([[0.3, 0, 0]] + [[0, 0, 0]] * 7)
    tn = np.array([[1.0, 0, 0], [0, 0, 0], [0, 0, 0], [0, 0, 0]])
    ls = np.array([1, 0, 0, 0, 0, 0])
    nv = np.zeros(72)
    nv[:12] = tn.ravel()
    good = _nf(X, ls, nv)
    nv_bad = np.zeros(72)
    nv_bad[:12] = tn[[1, 0, 2, 3]].ravel()
    bad = _nf(X, ls, nv_bad)
    assert np.abs(good - bad).max() > 1e-3
