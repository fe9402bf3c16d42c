"""DOF numbering (gtopology::gencodnum, give_code_numbers) of the host library
through the C ABI, against the oracle's restatement (cr_codenum, cr_elem_codes,
pinned by a hand-checked example in test_oracle.py) and against the rules
themselves. Host code only: runs without a GPU."""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh


def _oracle(flags, ordering=None):
    d = np.ascontiguousarray(flags, dtype=np.int64).copy()
    o = None if ordering is None else O._ptr(np.ascontiguousarray(ordering, dtype=np.int64))
    n = O.lib().cr_codenum(d.shape[0], d.shape[1], O._ptr(d), o)
    return d, n


def _random_flags(rng, nn, ndofn):
    f = rng.choice([1, 1, 1, 0, -1, -2, 2, 3, 4, 5], size=(nn, ndofn)).astype(np.int64)
    f[f == -2] = -rng.integers(1, 50, size=int((f == -2).sum()))
    return f


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("ordered", [False, True])
def test_gencodnum_equals_the_oracle(seed, ordered):
    rng = np.random.default_rng(seed)
    nn, ndofn = int(rng.integers(1, 300)), int(rng.integers(1, 4))
    flags = _random_flags(rng, nn, ndofn)
    ordering = rng.permutation(nn) if ordered else None
    d_o, n_o = _oracle(flags, ordering)
    d_g, n_g = gpu.gencodnum(flags, ordering)
    assert n_g == n_o and np.array_equal(d_g, d_o)
    d_m, n_m = mesh.codenum(flags, ordering)
    assert n_m == n_o and np.array_equal(d_m, d_o)


def test_gencodnum_rules():
    # node order 2, 0, 1: node 2's free DOF is the first unknown; group 7 takes
    # its number at node 2 and shares it with node 1; fixed and prescribed stay
    flags = np.array([[1, 0], [7, -3], [7, 1]])
    d, n = gpu.gencodnum(flags, [2, 0, 1])
    assert n == 3
    assert d.tolist() == [[3, 0], [1, -3], [1, 2]]
    # every unknown 1..n occurs, groups once
    rng = np.random.default_rng(9)
    f = _random_flags(rng, 500, 3)
    d, n = gpu.gencodnum(f)
    pos = d[d > 0]
    assert set(pos.tolist()) == set(range(1, n + 1))
    for k in np.unique(f[f > 1]):
        assert np.unique(d[f == k]).size == 1
    assert np.array_equal(d[f <= 0], f[f <= 0])


def test_gencodnum_refuses_a_bad_ordering():
    with pytest.raises(gpu.SsfError):
        gpu.gencodnum(np.ones((4, 1)), [0, 1, 1, 3])


def test_give_code_numbers_equals_the_oracle():
    prob = mesh.cantilever(5, 3, 2)
    cn = gpu.give_code_numbers(prob.conn, prob.dof)
    co = np.zeros_like(cn)
    O.lib().cr_elem_codes(prob.ne, 8, O._ptr(O.i64(prob.conn)), 3, O._ptr(O.i64(prob.dof)),
                          O._ptr(co))
    assert np.array_equal(cn, co) and np.array_equal(cn, prob.elem_codes())
    with pytest.raises(gpu.SsfError):
        gpu.give_code_numbers(prob.conn + prob.coords.shape[0], prob.dof)
