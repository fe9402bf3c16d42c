"""GPU Newton-Raphson load control (gnewton_raphson, MEFEL/SRC/newtonraph.cpp:
51-268) on device vectors, through the C ABI, against the oracle's restatement.

The control decisions (accepted or rejected increments, the dlambda sequence,
solves per increment) must be the oracle's exactly; the displacements agree to
the solver's rounding."""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.tensor(np.ascontiguousarray(x, dtype=np.float64), device=torch.device("cuda:0"))


def _setup(ctx, prob):
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    K = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    K.set_values(a)
    el = gpu.Elements(ctx, prob)
    return K, el, adr, ci, a, rhs[0]


def _run(ctx, prob, K, el, fp, fc, **kw):
    import torch
    fpd, fcd = _dev(fp), _dev(fc)
    fa = torch.zeros(prob.n, dtype=torch.float64, device="cuda:0")
    ra = torch.zeros_like(fa)
    torch.cuda.synchronize()
    sres = kw.pop("sres", 1e-24)
    info, h = gpu.newton_raphson(K, el, fpd.data_ptr(), fa.data_ptr(), ra.data_ptr(),
                                 fcd.data_ptr(), res=sres, **kw)
    return info, h, fa.cpu().numpy(), ra.cpu().numpy()


def _same_decisions(h, ho):
    assert list(h["acc"]) == list(ho["acc"])
    assert list(h["nsol"]) == list(ho["nsol"])
    assert np.array_equal(h["dlam"], ho["dlam"]) and np.array_equal(h["lam"], ho["lam"])


def test_nr_increments_vs_oracle(ctx):
    pytest.importorskip("torch")
    prob = mesh.cantilever(8, 3, 3)
    K, el, adr, ci, a, fp = _setup(ctx, prob)
    fc = 0.3 * np.roll(fp, 7)
    kw = dict(incrnr=0.1, maxincrnr=0.35, check_lambdar=True, lambdar=1.3, errnr=1e-6)
    lam_o, ra_o, fa_o, ho = O.newton_raphson(prob, adr, ci, a, fp, fc=fc, **kw)
    info, h, fa, ra = _run(ctx, prob, K, el, fp, fc, **kw)
    assert info.status == 1 and info.lambda_ == lam_o
    _same_decisions(h, ho)
    assert np.abs(ra - ra_o).max() <= 1e-9 * np.abs(ra_o).max()
    assert np.abs(fa - fa_o).max() <= 1e-14 * np.abs(fa_o).max()
    el.close()
    K.close()


def test_nr_rejections_vs_oracle(ctx):
    pytest.importorskip("torch")
    prob = mesh.cantilever(4, 2, 2)
    K, el, adr, ci, a, fp = _setup(ctx, prob)
    kw = dict(incrnr=0.5, minincrnr=0.07, niilnr=2, errnr=1e-30)
    lam_o, ra_o, fa_o, ho = O.newton_raphson(prob, adr, ci, a, fp, **kw)
    info, h, fa, ra = _run(ctx, prob, K, el, fp, np.zeros(prob.n), **kw)
    assert info.status == 2 and info.lambda_ == 0.0 and info.rejected == 4
    _same_decisions(h, ho)
    assert not ra.any() and np.abs(fa).max() <= 1e-9 * np.abs(fp).max()
    el.close()
    K.close()


def test_nr_inner_iterations_vs_oracle(ctx):
    pytest.importorskip("torch")
    prob = mesh.elastic_column(3, 3, 6, seed=51)
    K, el, adr, ci, a, fp = _setup(ctx, prob)
    fp = fp if np.abs(fp).max() > 0 else np.ones(prob.n)
    kw = dict(incrnr=0.5, check_lambdar=True, lambdar=1.0, errnr=1e-9, niilnr=20)
    lam_o, ra_o, fa_o, ho = O.newton_raphson(prob, adr, ci, a, fp, sres=1e-6, **kw)
    info, h, fa, ra = _run(ctx, prob, K, el, fp, np.zeros(prob.n), sres=1e-6, **kw)
    assert info.lambda_ == lam_o == 1.0
    assert list(h["acc"]) == list(ho["acc"])
    # each correction's loose solve is rounded differently on the device, so the
    # count of corrections may differ by one
    assert np.abs(np.array(h["nsol"]) - np.array(ho["nsol"])).max() <= 1
    assert info.inner > 0
    assert np.abs(ra - ra_o).max() <= 1e-7 * np.abs(ra_o).max()
    el.close()
    K.close()


def test_nr_quad4_plate_and_refusals(ctx):
    pytest.importorskip("torch")
    prob = mesh.plate(20, 8, seed=52, nrhs=1)
    K, el, adr, ci, a, fp = _setup(ctx, prob)
    kw = dict(incrnr=0.25, check_lambdar=True, lambdar=1.0, errnr=1e-6)
    lam_o, ra_o, fa_o, ho = O.newton_raphson(prob, adr, ci, a, fp, sres=1e-20, **kw)
    info, h, fa, ra = _run(ctx, prob, K, el, fp, np.zeros(prob.n), sres=1e-20, **kw)
    assert info.lambda_ == lam_o == 1.0
    _same_decisions(h, ho)
    assert np.abs(ra - ra_o).max() <= 1e-8 * np.abs(ra_o).max()
    import torch
    z = _dev(np.zeros(prob.n))
    with pytest.raises(gpu.SsfError):  # reaction-based norms are not on this path
        gpu.newton_raphson(K, el, z.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(),
                           normt=2)
    heat = mesh.heat_cube(3, seed=1)
    Kh, elh, _, _, _, _ = _setup(ctx, heat)
    zh = _dev(np.zeros(heat.n))
    torch.cuda.synchronize()
    with pytest.raises(gpu.SsfError):  # heat element sets are not mechanical
        gpu.newton_raphson(Kh, elh, zh.data_ptr(), zh.data_ptr(), zh.data_ptr(), zh.data_ptr())
    for x in (elh, Kh, el, K):
        x.close()
