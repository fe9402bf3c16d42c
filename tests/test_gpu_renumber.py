"""GPU Cuthill-McKee renumbering (gtopology::cuthill_mckee_renumb, GEFEL/
gtopology.cpp:6242) through the C ABI: the ordering must equal the reference's
(fixture) and the oracle's exactly, forward and reversed."""
import os

import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
CASES = {"elast": lambda: mesh.elastic_column(2, 2, 3, seed=21),
         "cant": lambda: mesh.cantilever(4, 2, 2),
         "heat": lambda: mesh.heat_cube(4, seed=22),
         "quad": lambda: mesh.plate(6, 3, seed=23, nrhs=1)}


@pytest.mark.parametrize("name", list(CASES))
def test_cm_vs_reference_fixture(ctx, name):
    g = np.load(GOLD)
    prob = CASES[name]()
    nn = prob.coords.shape[0]
    assert np.array_equal(gpu.cuthill_mckee(ctx, nn, prob.conn), g[f"{name}_cm_ordering"])
    assert np.array_equal(gpu.cuthill_mckee(ctx, nn, prob.conn, reverse=True),
                          g[f"{name}_cmr_ordering"])


def _irregular(seed, nel, drop):
    prob = mesh.heat_cube(nel, seed=seed)
    rng = np.random.default_rng(seed)
    conn = prob.conn[rng.uniform(size=prob.ne) > drop]
    used = np.unique(conn)
    remap = -np.ones(prob.coords.shape[0], np.int64)
    remap[used] = np.arange(used.shape[0])
    return used.shape[0], remap[conn]


@pytest.mark.parametrize("seed", [11, 12])
def test_cm_irregular_vs_oracle(ctx, seed):
    nn, conn = _irregular(seed, 12, 0.15)
    want = O.cuthill_mckee(nn, conn)
    if want is None:  # this removal split the mesh: the device must say so
        with pytest.raises(gpu.SsfError):
            gpu.cuthill_mckee(ctx, nn, conn)
        return
    assert np.array_equal(gpu.cuthill_mckee(ctx, nn, conn), want)
    assert np.array_equal(gpu.cuthill_mckee(ctx, nn, conn, reverse=True),
                          O.cuthill_mckee(nn, conn, True))


def test_cm_larger_meshes_vs_oracle(ctx):
    for prob in (mesh.heat_cube(24, seed=13), mesh.plate(120, 40, seed=14, nrhs=1),
                 mesh.elastic_column(10, 12, 30, seed=15)):
        nn = prob.coords.shape[0]
        assert np.array_equal(gpu.cuthill_mckee(ctx, nn, prob.conn, reverse=True),
                              O.cuthill_mckee(nn, prob.conn, True)), prob.name


def test_cm_refusals(ctx):
    with pytest.raises(gpu.SsfError):  # two separate elements
        gpu.cuthill_mckee(ctx, 8, np.array([[0, 1, 2, 3], [4, 5, 6, 7]]))
    with pytest.raises(gpu.SsfError):  # node id out of range
        gpu.cuthill_mckee(ctx, 4, np.array([[0, 1, 2, 9]]))
