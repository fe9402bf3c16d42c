"""Element volume loads / heat sources on the device (ssf_elements_volume_load)
against the oracle and the reference's fixture, bit for bit, through the C ABI."""
import os

import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")


@pytest.mark.parametrize("name,make", [("elast", lambda: mesh.elastic_column(2, 2, 3, seed=21)),
                                       ("heat", lambda: mesh.heat_cube(4, seed=22))])
def test_volume_loads_match_the_reference_fixture(ctx, name, make):
    g = np.load(GOLD)
    prob = make()
    el = gpu.Elements(ctx, prob)
    rhs = el.volume_load(g[f"{name}_vload_nodval"], g[f"{name}_vload_rhs0"])
    assert np.array_equal(rhs, g[f"{name}_vload_rhs"])
    el.close()


@pytest.mark.parametrize("make", [lambda: mesh.elastic_column(16, 16, 64, seed=63),
                                  lambda: mesh.heat_cube(48, seed=64),
                                  lambda: mesh.cantilever(20, 6, 6)])
def test_volume_loads_bitwise_vs_oracle(ctx, make):
    prob = make()
    rng = np.random.default_rng(65)
    nodval = rng.standard_normal((prob.ne, prob.ndofe))
    rhs0 = rng.standard_normal(prob.n)
    el = gpu.Elements(ctx, prob)
    assert np.array_equal(el.volume_load(nodval, rhs0), O.volume_load(prob, nodval, rhs0))
    # a constant body force: the assembled loads of the free nodes add up to b V
    # minus the shares of the constrained nodes, which the oracle equally drops
    b = np.array([0.0, 0.0, -9.81e3])[: prob.ndofn] if prob.ndofn == 3 else np.array([2.0])
    nv = np.tile(b, prob.nne)[None, :].repeat(prob.ne, axis=0)
    assert np.array_equal(el.volume_load(nv), O.volume_load(prob, nv))
    el.close()


def test_volume_loads_refuse_quad4(ctx):
    prob = mesh.plate(6, 3, seed=1, nrhs=1)
    el = gpu.Elements(ctx, prob)
    with pytest.raises(gpu.SsfError):
        el.volume_load(np.zeros((prob.ne, prob.ndofe)))
    el.close()
