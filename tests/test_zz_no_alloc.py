"""ssf_solve_dev_noalloc (the FP64 cg solve on the library's static scratch array,
no allocation) against ssf_solve_dev on the same matrix and right-hand side: equal
bit for bit, on two calls in a row, the second with a smaller system, for each
cg variant it serves; sizes beyond its maximum and other solvers are refused."""
import numpy as np
import pytest

from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu


def _pair(ctx, prob, storage=gpu.STORAGE_SCR, prec=gpu.PREC_DIAG, res=1e-22):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    m, rhs = gpu.build_problem_matrix(ctx, prob, storage)
    y = torch.tensor(rhs[0], dtype=torch.float64, device=dev)
    x1 = torch.zeros(prob.n, dtype=torch.float64, device=dev)
    x2 = torch.full((prob.n,), 7.0, dtype=torch.float64, device=dev)  # overwritten (iv = 0)
    torch.cuda.synchronize()
    i1 = m.solve_dev(y.data_ptr(), x1.data_ptr(), ni=20000, res=res, prec=prec)
    i2 = m.solve_dev_noalloc(y.data_ptr(), x2.data_ptr(), ni=20000, res=res, prec=prec)
    torch.cuda.synchronize()
    assert torch.equal(x1, x2), prob.name
    for f in ("status", "iterations", "ani", "ares", "rel_residual"):
        assert getattr(i1, f) == getattr(i2, f), (prob.name, f, getattr(i1, f), getattr(i2, f))
    assert i2.status in (1, 2)
    m.close()
    return i2


def test_noalloc_equals_solve_dev_on_two_calls(ctx):
    # the second call: a smaller system on the same scratch array
    _pair(ctx, mesh.heat_cube(20, seed=81))
    _pair(ctx, mesh.heat_cube(12, seed=82))


@pytest.mark.parametrize("storage,prec", [(gpu.STORAGE_SCR, gpu.PREC_NONE),
                                          (gpu.STORAGE_CR, gpu.PREC_NONE)])
def test_noalloc_other_cg_variants(ctx, storage, prec):
    _pair(ctx, mesh.cantilever(10, 3, 3), storage, prec, res=1e-12)
    _pair(ctx, mesh.plate(9, 5, seed=83, nrhs=1), storage, prec, res=1e-12)


def test_noalloc_refuses_beyond_its_maximum_and_other_solvers(ctx):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    big = mesh.heat_cube(140, seed=84)
    assert big.n > 2500000
    mb = gpu.Matrix.from_elements_device(ctx, big.n, big.elem_codes())
    tiny = torch.zeros(4, dtype=torch.float64, device=dev)  # never read: refused first
    with pytest.raises(gpu.SsfError, match="SSF_NOALLOC_MAX_N"):
        mb.solve_dev_noalloc(tiny.data_ptr(), tiny.data_ptr())
    mb.close()
    prob = mesh.heat_cube(6, seed=85)
    m, rhs = gpu.build_problem_matrix(ctx, prob)
    y = torch.tensor(rhs[0], dtype=torch.float64, device=dev)
    x = torch.zeros_like(y)
    p = gpu.SolveParams(gpu.SOLVER_CG, gpu.PREC_DIAG, 100, 1e-20, 1e-30, 0, gpu.MIXED, 0.0, 0, 0)
    info = gpu.SolveInfo()
    rc = gpu.lib().ssf_solve_dev_noalloc(m.h, gpu.ctypes.byref(p), gpu.ctypes.c_void_p(y.data_ptr()),
                                         gpu.ctypes.c_void_p(x.data_ptr()), gpu.ctypes.byref(info))
    assert rc == -5  # SSF_EUNSUP: the FP64 cg solver only
    m.close()
