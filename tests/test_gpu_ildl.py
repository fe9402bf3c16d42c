"""GPU incomplete LDL^T (symcomprow::incomplete_ldl, prec_ildl_scr, cg_prec with
tprec = incomdec; GEFEL/scr.cpp:958-1203) through the C ABI.

The factor and the preconditioner application run each row in the reference's
operation order, so both are compared bit for bit with the oracle and with the
reference's outputs in the fixture. The solve's dot products are summed in
another order, so x agrees to rounding and ani within one iteration."""
import os

import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "ref_vectors.npz")
CASES = {"heat": lambda: mesh.heat_cube(4, seed=22),
         "elast": lambda: mesh.elastic_column(2, 2, 3, seed=21)}


def _dev(x):
    import torch
    return torch.tensor(np.ascontiguousarray(x, dtype=np.float64), device=torch.device("cuda:0"))


def _scr(ctx, prob, device_pattern=False):
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    if device_pattern:
        m = gpu.Matrix.from_elements_device(ctx, prob.n, prob.elem_codes())
    else:
        m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    return m, adr, ci, a, rhs


def _apply(m, y):
    import torch
    yd = _dev(y)
    xd = torch.zeros_like(yd)
    torch.cuda.synchronize()
    m.ildl_apply_dev(yd.data_ptr(), xd.data_ptr())
    torch.cuda.synchronize()
    return xd.cpu().numpy()


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name", ["heat", "elast"])
def test_ildl_vs_reference_fixture(ctx, name):
    pytest.importorskip("torch")
    g = np.load(GOLD)
    prob = CASES[name]()
    m, adr, ci, a, rhs = _scr(ctx, prob)
    assert np.array_equal(m.ildl_factor(), g[f"{name}_ildl_inc"])
    assert np.array_equal(_apply(m, g[f"{name}_ildl_yv"]), g[f"{name}_ildl_pv"])
    x, info = m.solve(rhs[0], ni=2000, res=1e-24, prec=gpu.PREC_ILDL)
    assert info.status == 1
    assert _rel(x, g[f"{name}_ildl_x"]) <= 1e-10
    assert abs(info.ani - int(g[f"{name}_ildl_ani"])) <= 1
    m.close()


@pytest.mark.parametrize("make", [lambda: mesh.heat_cube(12, seed=90),
                                  lambda: mesh.elastic_column(4, 4, 8, seed=91),
                                  lambda: mesh.plate(30, 12, seed=92, nrhs=1)])
def test_ildl_factor_apply_solve_vs_oracle(ctx, make):
    pytest.importorskip("torch")
    prob = make()
    m, adr, ci, a, rhs = _scr(ctx, prob)
    inc = O.incomplete_ldl(adr, ci, a)
    assert np.array_equal(m.ildl_factor(), inc)
    y = np.random.default_rng(prob.n).standard_normal(prob.n)
    assert np.array_equal(_apply(m, y), O.prec_ildl(adr, ci, inc, y))
    # res 1e-20: the plate's entries are ~1e11, and at 1e-24 d.p falls below
    # zero = 1e-30 before the test is met (status 2 in the oracle as well)
    xo, anio, areso, sto = O.scr_cg_prec_ildl(adr, ci, a, rhs[0], 5000, 1e-20, incdec=inc)
    x, info = m.solve(rhs[0], ni=5000, res=1e-20, prec=gpu.PREC_ILDL)
    assert sto == 1 and info.status == 1
    assert _rel(x, xo) <= 1e-9
    assert abs(info.ani - anio) <= 1
    # ares = r.h / y.y at the stop, a quantity summed in another order
    assert abs(info.ares - areso) <= 1e-3 * areso
    # fewer iterations than Jacobi on the same system
    _, ij = m.solve(rhs[0], ni=5000, res=1e-20, prec=gpu.PREC_DIAG)
    assert info.ani < ij.ani
    m.close()


def test_ildl_refactors_after_a_values_change(ctx):
    pytest.importorskip("torch")
    prob = mesh.heat_cube(8, seed=93)
    m, adr, ci, a, rhs = _scr(ctx, prob)
    m.solve(rhs[0], res=1e-20, prec=gpu.PREC_ILDL)
    a2 = a * np.random.default_rng(3).uniform(0.9, 1.1, a.shape[0])
    m.set_values(a2)
    inc2 = O.incomplete_ldl(adr, ci, a2)
    assert np.array_equal(m.ildl_factor(), inc2)
    x, info = m.solve(rhs[0], ni=5000, res=1e-24, prec=gpu.PREC_ILDL)
    xo, _, _, _ = O.scr_cg_prec_ildl(adr, ci, a2, rhs[0], 5000, 1e-24, incdec=inc2)
    assert info.status == 1 and _rel(x, xo) <= 1e-10
    m.close()


def test_ildl_device_pattern_and_device_vectors(ctx):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(10, seed=94)
    m, adr, ci, a, rhs = _scr(ctx, prob, device_pattern=True)
    inc = O.incomplete_ldl(adr, ci, a)
    assert np.array_equal(m.ildl_factor(), inc)
    yd = _dev(rhs[0])
    xd = torch.full((prob.n,), 7.0, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    info = m.solve_dev(yd.data_ptr(), xd.data_ptr(), ni=5000, res=1e-24, prec=gpu.PREC_ILDL)
    xo, anio, _, _ = O.scr_cg_prec_ildl(adr, ci, a, rhs[0], 5000, 1e-24, incdec=inc)
    assert info.status == 1 and _rel(xd.cpu().numpy(), xo) <= 1e-10
    assert abs(info.ani - anio) <= 1
    # from a start vector (iv = 1): the same fixed point
    x0 = xo + 1e-3 * np.random.default_rng(5).standard_normal(prob.n)
    x1, i1 = m.solve(rhs[0], ni=5000, res=1e-24, prec=gpu.PREC_ILDL, x0=x0)
    assert i1.status == 1 and _rel(x1, xo) <= 1e-10
    m.close()


def test_ildl_conventions_and_refusals(ctx):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(5, seed=95)
    m, adr, ci, a, rhs = _scr(ctx, prob)
    x, info = m.solve(np.zeros(prob.n), prec=gpu.PREC_ILDL)
    assert info.status == 3 and info.ani == 0 and not x.any()
    _, info = m.solve(rhs[0], ni=2, res=1e-40, prec=gpu.PREC_ILDL)
    assert info.status == 0 and info.iterations == 2
    with pytest.raises(gpu.SsfError):
        m.solve(rhs[0], prec=gpu.PREC_ILDL, precision=gpu.MIXED)
    yd = _dev(rhs[0])
    with pytest.raises(gpu.SsfError):
        m.ildl_apply_dev(yd.data_ptr(), yd.data_ptr())
    adr_c, ci_c, a_c, _ = O.assemble_problem(prob, "cr")
    c = gpu.Matrix.from_csr(ctx, gpu.STORAGE_CR, adr_c, ci_c)
    c.set_values(a_c)
    with pytest.raises(gpu.SsfError):
        c.solve(rhs[0], prec=gpu.PREC_ILDL)
    with pytest.raises(gpu.SsfError):
        c.ildl_factor()
    c.close()
    m.close()
