"""CPU checks that pin the oracle's node renumbering (cr_edge_adjacency,
cr_cuthill_mckee: gtopology::adjacnodes_edge GEFEL/gtopology.cpp:2632 and
cuthill_mckee_renumb :6242):

* the ordering equals the reference's own gtopology::cuthill_mckee_renumb, forward
  and reversed: the committed fixture on the four cases and, where oracle/_ref was
  built, live on further meshes, including irregular ones (elements removed, so
  degrees and ties vary);
* graph facts: the ordering is a permutation and a breadth-first order (the BFS
  level of ordering[k] never decreases), it starts at the last node of least
  degree, a node's neighbours first numbered by one parent are numbered in order
  of degree, and the reversed ordering reduces the profile of the renumbered
  matrix below that of a random numbering.
"""
import contextlib
import os

import numpy as np
import pytest

import oracle as O
from oracle import _ptr, i64
from ssf_amd import mesh

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
CASES = {"elast": lambda: mesh.elastic_column(2, 2, 3, seed=21),
         "cant": lambda: mesh.cantilever(4, 2, 2),
         "heat": lambda: mesh.heat_cube(4, seed=22),
         "quad": lambda: mesh.plate(6, 3, seed=23, nrhs=1)}


@contextlib.contextmanager
def _quiet_fd1():
    """the reference prints its graph on the C stdout: send file descriptor 1 to
    /dev/null meanwhile"""
    import sys
    sys.stdout.flush()
    saved = os.dup(1)
    dn = os.open(os.devnull, os.O_WRONLY)
    os.dup2(dn, 1)
    try:
        yield
    finally:
        import ctypes
        ctypes.CDLL(None).fflush(None)  # the C stream's buffer goes to /dev/null too
        os.dup2(saved, 1)
        os.close(dn)
        os.close(saved)


def _irregular(seed, nel=6, drop=0.25):
    """a heat cube with a random quarter of its elements removed, nodes renumbered
    compactly; the largest connected part is kept"""
    prob = mesh.heat_cube(nel, seed=seed)
    rng = np.random.default_rng(seed)
    keep = rng.uniform(size=prob.ne) > drop
    conn = prob.conn[keep]
    used = np.unique(conn)
    remap = -np.ones(prob.coords.shape[0], np.int64)
    remap[used] = np.arange(used.shape[0])
    conn = remap[conn]
    nn = used.shape[0]
    # keep the component of node 0
    ptr, adj = O.edge_adjacency(nn, conn)
    seen = np.zeros(nn, bool)
    stack = [0]
    seen[0] = True
    while stack:
        u = stack.pop()
        for v in adj[ptr[u]:ptr[u + 1]]:
            if not seen[v]:
                seen[v] = True
                stack.append(v)
    ekeep = seen[conn].all(axis=1)
    conn = conn[ekeep]
    used = np.unique(conn)
    remap = -np.ones(nn, np.int64)
    remap[used] = np.arange(used.shape[0])
    return used.shape[0], remap[conn]


@pytest.mark.parametrize("name", list(CASES))
def test_cm_fixture(name):
    g = np.load(GOLD)
    prob = CASES[name]()
    nn = prob.coords.shape[0]
    assert np.array_equal(O.cuthill_mckee(nn, prob.conn), g[f"{name}_cm_ordering"])
    assert np.array_equal(O.cuthill_mckee(nn, prob.conn, True), g[f"{name}_cmr_ordering"])


def test_cm_live_reference():
    if not O.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    R = O.ref()
    meshes = [(_irregular(s)) for s in (1, 2, 3)]
    meshes += [(p.coords.shape[0], p.conn) for p in (mesh.elastic_column(3, 4, 5, seed=4),
                                                     mesh.plate(9, 5, seed=5, nrhs=1))]
    for nn, conn in meshes:
        conn = i64(conn)
        for rev in (0, 1):
            o = np.zeros(nn, np.int64)
            with _quiet_fd1():
                R.ref_cuthill_mckee(nn, conn.shape[0], conn.shape[1], _ptr(conn), rev, _ptr(o))
            assert np.array_equal(O.cuthill_mckee(nn, conn, bool(rev)), o)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_cm_is_a_bfs_from_the_least_degree_node(seed):
    nn, conn = _irregular(seed)
    ptr, adj = O.edge_adjacency(nn, conn)
    deg = np.diff(ptr)
    o = O.cuthill_mckee(nn, conn)
    assert np.array_equal(np.sort(o), np.arange(nn))
    assert o[0] == np.flatnonzero(deg == deg.min())[-1]
    # BFS levels from the start never decrease along the ordering
    lev = -np.ones(nn, np.int64)
    lev[o[0]] = 0
    q = [o[0]]
    for u in q:
        for v in adj[ptr[u]:ptr[u + 1]]:
            if lev[v] < 0:
                lev[v] = lev[u] + 1
                q.append(v)
    assert (np.diff(lev[o]) >= 0).all()
    # children of one parent are numbered in order of degree
    num = np.empty(nn, np.int64)
    num[o] = np.arange(nn)
    parent = np.full(nn, -1)
    for k in range(nn):
        u = o[k]
        for v in adj[ptr[u]:ptr[u + 1]]:
            if parent[v] < 0 and num[v] > k:
                parent[v] = u
    for u in range(nn):
        kids = np.flatnonzero(parent == u)
        kids = kids[np.argsort(num[kids])]
        assert (np.diff(deg[kids]) >= 0).all()


def test_rcm_reduces_the_profile():
    prob = mesh.heat_cube(6, seed=7)
    nn = prob.coords.shape[0]
    ptr, adj = O.edge_adjacency(nn, prob.conn)
    rows = np.repeat(np.arange(nn), np.diff(ptr))

    def profile(order):
        num = np.empty(nn, np.int64)
        num[order] = np.arange(nn)
        i, j = num[rows], num[adj]
        lo = np.full(nn, nn)
        np.minimum.at(lo, i, np.minimum(i, j))
        return int((np.arange(nn) - np.minimum(lo, np.arange(nn))).sum())

    rng = np.random.default_rng(8)
    assert profile(O.cuthill_mckee(nn, prob.conn, True)) < 0.3 * profile(rng.permutation(nn))


def test_cm_disconnected_graph_is_reported():
    conn = np.array([[0, 1, 2, 3], [4, 5, 6, 7]])
    assert O.cuthill_mckee(8, conn) is None
