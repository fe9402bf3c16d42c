"""Batched solver (many systems x R right-hand sides in one launch sequence):
every (system, rhs) pair against the single-system solve and the oracle, and the
same result alone and among other systems."""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

RES = 1e-20


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def problems():
    return [
        mesh.plate(9, 5, seed=21, nrhs=1),
        mesh.heat_cube(6, seed=22),
        mesh.plate(20, 10, seed=23, nrhs=1),
        mesh.elastic_column(3, 4, 6, seed=24),
        mesh.plate(31, 7, seed=25, nrhs=1),
        mesh.cantilever(8, 3, 3),
    ]


def build(ctx, probs, R, seed=7):
    """Matrices assembled on the device and R right-hand sides per system: the
    problem's own load (with the Dirichlet lift) first, then seeded random loads."""
    rng = np.random.default_rng(seed)
    mats, rhs = [], []
    for p in probs:
        m, r = gpu.build_problem_matrix(ctx, p)
        y = np.zeros((R, p.n))
        y[0] = r[0]
        for k in range(1, R):
            y[k] = rng.standard_normal(p.n) * np.abs(r[0]).max()
        mats.append(m)
        rhs.append(y)
    return mats, rhs


@pytest.mark.parametrize("R", [1, 4])
def test_batch_fp64_matches_single_and_oracle(ctx, R):
    probs = problems()
    mats, rhs = build(ctx, probs, R)
    b = gpu.Batch(ctx, mats, R)
    assert b.nrows == sum(p.n for p in probs)
    xs, infos = b.solve(rhs, ni=20000, res=RES)
    for s, (p, m) in enumerate(zip(probs, mats)):
        adr, ci = m.pattern()
        a = m.get_values()
        for r in range(R):
            info = infos[s * R + r]
            x1, i1 = m.solve(rhs[s][r], ni=20000, res=RES)
            assert info.status == i1.status, (s, r, info.status, i1.status)
            assert abs(info.iterations - i1.iterations) <= max(2, i1.iterations // 50)
            assert rel(xs[s][r], x1) <= 1e-8, (s, r, rel(xs[s][r], x1))
            if r == 0:
                xo, ani_o, _, st_o = O.scr_cg_prec(adr, ci, a, rhs[s][r], 20000, RES)
                assert st_o == info.status
                assert abs(info.ani - ani_o) <= max(2, ani_o // 50)
                assert rel(xs[s][r], xo) <= 1e-8


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.MIXED])
def test_batch_alone_equals_in_mix(ctx, precision):
    probs = problems()
    R = 2
    mats, rhs = build(ctx, probs, R, seed=8)
    kw = dict(ni=20000, res=RES, precision=precision)
    xs, infos = gpu.Batch(ctx, mats, R).solve(rhs, **kw)
    for s in (0, 3, 5):
        xa, ia = gpu.Batch(ctx, [mats[s]], R).solve([rhs[s]], **kw)
        for r in range(R):
            assert np.array_equal(xa[0][r], xs[s][r]), (s, r)
            assert ia[r].iterations == infos[s * R + r].iterations
            assert ia[r].status == infos[s * R + r].status
    # another order of the same systems
    order = [5, 2, 0, 4, 1, 3]
    xo, io = gpu.Batch(ctx, [mats[k] for k in order], R).solve([rhs[k] for k in order], **kw)
    for j, s in enumerate(order):
        for r in range(R):
            assert np.array_equal(xo[j][r], xs[s][r])


@pytest.mark.parametrize("R", [1, 8])
def test_batch_mixed_precision(ctx, R):
    probs = problems()
    mats, rhs = build(ctx, probs, R, seed=9)
    b = gpu.Batch(ctx, mats, R)
    xr, _ = b.solve(rhs, ni=40000, res=1e-30)  # FP64 to its limit
    xm, im = b.solve(rhs, ni=40000, res=1e-20, precision=gpu.MIXED)
    for s in range(len(probs)):
        for r in range(R):
            info = im[s * R + r]
            assert info.status in (1, 4), (s, r, info.status)
            if info.status == 1:
                assert info.rel_residual < 1e-9
            assert rel(xm[s][r], xr[s][r]) <= 1e-8, (s, r, rel(xm[s][r], xr[s][r]))


def test_batch_zero_rhs_and_ni_exhausted(ctx):
    probs = problems()
    R = 2
    mats, rhs = build(ctx, probs, R, seed=10)
    rhs[1][1][:] = 0.0
    b = gpu.Batch(ctx, mats, R)
    ni = 37  # not a multiple of the 32-iteration chunk
    xs, infos = b.solve(rhs, ni=ni, res=RES)
    z = infos[1 * R + 1]
    assert z.status == 3 and z.iterations == 0 and not np.any(xs[1][1])
    for s, m in enumerate(mats):
        for r in range(R):
            if s == 1 and r == 1:
                continue
            info = infos[s * R + r]
            x1, i1 = m.solve(rhs[s][r], ni=ni, res=RES)
            assert info.status == i1.status
            if info.status == 0:
                assert info.iterations == ni
            # an unconverged iterate: the two FP64 recurrences differ only in the
            # summation order of their dot products. Over 37 iterations that stays
            # at 1e-15 on the plates and the heat cube, but reaches 1e-7 on the
            # elasticity systems (two CPU restatements of the recurrence with
            # different summation orders differ by as much there)
            tol = 1e-5 if probs[s].kind == mesh.KIND_HEX8_ELAST else 1e-10
            assert rel(xs[s][r], x1) <= tol
    # mixed: the zero pair stays zero
    xm, im = b.solve(rhs, ni=20000, res=RES, precision=gpu.MIXED)
    assert im[1 * R + 1].status == 3 and not np.any(xm[1][1])


def test_batch_replay_and_new_tolerance(ctx):
    probs = problems()[:4]
    mats, rhs = build(ctx, probs, 1, seed=11)
    b = gpu.Batch(ctx, mats, 1)
    x1, i1 = b.solve(rhs, ni=20000, res=1e-16)
    x2, i2 = b.solve(rhs, ni=20000, res=1e-16)
    for s in range(len(probs)):
        assert np.array_equal(x1[s][0], x2[s][0])
        assert i1[s].iterations == i2[s].iterations
    x3, i3 = b.solve(rhs, ni=20000, res=1e-24)  # the captured chunk is rebuilt
    for s in range(len(probs)):
        assert i3[s].iterations > i1[s].iterations or i3[s].status != 1
        x1s, _ = mats[s].solve(rhs[s][0], ni=20000, res=1e-24)
        assert rel(x3[s][0], x1s) <= 1e-8


def test_batch_device_vectors_and_autotune(ctx):
    torch = pytest.importorskip("torch")
    probs = problems()
    R = 4
    mats, rhs = build(ctx, probs, R, seed=12)
    b = gpu.Batch(ctx, mats, R)
    xh, ih = b.solve(rhs, ni=20000, res=RES)
    dev = torch.device("cuda:0")
    y = torch.tensor(b.pack(rhs), dtype=torch.float64, device=dev)
    x = torch.zeros_like(y)
    torch.cuda.synchronize()
    infos = b.solve_dev(y.data_ptr(), x.data_ptr(), ni=20000, res=RES)
    xd = b.unpack(x.cpu().numpy())
    for s in range(len(probs)):
        for r in range(R):
            assert np.array_equal(xd[s][r], xh[s][r])
            assert infos[s * R + r].iterations == ih[s * R + r].iterations
    fmt, lanes = b.autotune(gpu.FP64)
    assert fmt in (0, 2) and (fmt == 2 or lanes in (4, 8, 16, 32, 64))
    b.autotune(gpu.MIXED)
    xt, it = b.solve(rhs, ni=20000, res=RES)
    for s in range(len(probs)):
        for r in range(R):
            assert rel(xt[s][r], xh[s][r]) <= 1e-8
            assert it[s * R + r].status == ih[s * R + r].status


def test_batch_one_graph_launch_repeats_the_host_loop(ctx):
    """A solve after the first replays the whole solve from one graph sized by the
    last solve's chunk count; a harder right-hand side that needs more chunks
    continues from the host loop. Results equal the first solve's bit for bit."""
    probs = problems()
    mats, rhs = build(ctx, probs, 2, seed=13)
    b = gpu.Batch(ctx, mats, 2)
    # an easy right-hand side first: a multiple of x0 = A-applied smooth vector
    easy = [np.stack([m.gmxv(np.ones(p.n)), m.gmxv(np.ones(p.n))]) for m, p in zip(mats, probs)]
    x0, i0 = b.solve(easy, ni=20000, res=RES)
    for mixed in (gpu.FP64, gpu.MIXED):
        xa, ia = b.solve(rhs, ni=20000, res=RES, precision=mixed)  # more chunks than learned
        xb, ib = b.solve(rhs, ni=20000, res=RES, precision=mixed)  # one graph launch
        xc, ic = b.solve(rhs, ni=20000, res=RES, precision=mixed)
        for s in range(len(probs)):
            for r in range(2):
                assert np.array_equal(xa[s][r], xb[s][r]) and np.array_equal(xa[s][r], xc[s][r])
                q = s * 2 + r
                assert ia[q].iterations == ib[q].iterations == ic[q].iterations
                assert ia[q].status == ib[q].status == ic[q].status


def test_batch_plates_of_the_benchmark_conditioning_vs_oracle(ctx):
    """Plates like the 512-plate benchmark batch (quad4 plane stress, aspect 4,
    lognormal E with sigma_log 0.3, clamped edge; up to ~59k DOF), two load cases
    each: the FP64 batch at res 1e-20 and the mixed batch at res 1e-22 (the
    tolerances the benchmark runs) are within 1e-8 of the oracle's FP64 cg_prec
    run to 1e-30 on the oracle's own assembly of the same plate."""
    probs = [mesh.plate(340, 85, seed=41, nrhs=2), mesh.plate(160, 40, seed=42, nrhs=2),
             mesh.plate(64, 16, seed=43, nrhs=2)]
    mats, rhs = [], []
    for p in probs:
        m, r = gpu.build_problem_matrix(ctx, p)
        mats.append(m)
        rhs.append(r)
    b = gpu.Batch(ctx, mats, 2)
    x64, i64 = b.solve(rhs, ni=40000, res=1e-20)
    xmx, imx = b.solve(rhs, ni=40000, res=1e-22, precision=gpu.MIXED)
    for s, p in enumerate(probs):
        adr, ci, a, _ = O.assemble_problem(p, "scr")
        for r in range(2):
            xo, ani, ares, st = O.scr_cg_prec(adr, ci, a, rhs[s][r], 60000, 1e-30)
            assert st in (1, 2), (s, r, st)
            assert i64[s * 2 + r].status == 1 and imx[s * 2 + r].status in (1, 4)
            assert rel(x64[s][r], xo) <= 1e-8, (s, r, rel(x64[s][r], xo))
            assert rel(xmx[s][r], xo) <= 1e-8, (s, r, rel(xmx[s][r], xo))
    b.close()
    for m in mats:
        m.close()
