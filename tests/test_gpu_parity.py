"""GPU parity of the HIP path against the oracle (CPU restatement of the reference)
on the same seeded inputs, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

SMALL = {
    "hex8_elast": lambda: mesh.elastic_column(3, 4, 6, seed=11),
    "cantilever": lambda: mesh.cantilever(8, 3, 3),
    "hex8_heat": lambda: mesh.heat_cube(6, seed=12),
    "quad4": lambda: mesh.plate(9, 5, seed=13, nrhs=2),
}


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", list(SMALL))
@pytest.mark.parametrize("storage", ["scr", "cr"])
def test_pattern_and_assembly(ctx, name, storage):
    prob = SMALL[name]()
    st = gpu.STORAGE_SCR if storage == "scr" else gpu.STORAGE_CR
    m, rhs = gpu.build_problem_matrix(ctx, prob, st)
    adr, ci, a, rhs_o = O.assemble_problem(prob, storage)
    adr_g, ci_g = m.pattern()
    assert np.array_equal(adr_g, adr) and np.array_equal(ci_g, ci)
    a_g = m.get_values()
    scale = np.abs(a).max()
    assert np.abs(a_g - a).max() <= 1e-12 * scale
    assert np.abs(rhs - rhs_o).max() <= 1e-12 * max(np.abs(rhs_o).max(), 1e-300)


@pytest.mark.parametrize("name", list(SMALL))
def test_gmxv(ctx, name):
    prob = SMALL[name]()
    adr, ci, a, _ = O.assemble_problem(prob, "scr")
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    b = np.random.default_rng(1).standard_normal(prob.n)
    c = m.gmxv(b)
    ref = O.mxv_scr(adr, ci, a, b)
    assert rel(c, ref) <= 1e-14
    adr2, ci2, a2, _ = O.assemble_problem(prob, "cr")
    m2 = gpu.Matrix.from_csr(ctx, gpu.STORAGE_CR, adr2, ci2)
    m2.set_values(a2)
    assert rel(m2.gmxv(b), O.mxv_cr(adr2, ci2, a2, b)) <= 1e-14


@pytest.mark.parametrize("name", list(SMALL))
def test_pcg_fp64_parity(ctx, name):
    prob = SMALL[name]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    y = rhs[0]
    xo, ani_o, ares_o, st_o = O.scr_cg_prec(adr, ci, a, y, 20000, 1e-24)
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    x, info = m.solve(y, ni=20000, res=1e-24)
    # status 2 = |d.Ap| < zero (1e-30) before the residual test, as the reference
    assert st_o in (1, 2) and info.status == st_o
    assert rel(x, xo) <= 1e-8, rel(x, xo)
    assert abs(info.ani - ani_o) <= max(2, ani_o // 50)
    assert info.rel_residual < 1e-10


@pytest.mark.parametrize("name", ["hex8_heat", "quad4"])
def test_scr_cg_and_cr_cg(ctx, name):
    prob = SMALL[name]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    y = rhs[0]
    xo, ani_o, ares_o, st_o = O.scr_cg(adr, ci, a, y, 20000, 1e-12)
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    x, info = m.solve(y, ni=20000, res=1e-12, prec=gpu.PREC_NONE)
    assert rel(x, xo) <= 1e-8 and abs(info.ani - ani_o) <= max(2, ani_o // 50)
    adr2, ci2, a2, _ = O.assemble_problem(prob, "cr")
    xo2, ani2, ares2, st2 = O.cr_cg_prec_noprec(adr2, ci2, a2, y, 20000, 1e-12)
    m2 = gpu.Matrix.from_csr(ctx, gpu.STORAGE_CR, adr2, ci2)
    m2.set_values(a2)
    x2, info2 = m2.solve(y, ni=20000, res=1e-12, prec=gpu.PREC_NONE)
    assert rel(x2, xo2) <= 1e-8 and abs(info2.ani - ani2) <= max(2, ani2 // 50)


def test_cr_jacobi_unsupported(ctx):
    prob = SMALL["quad4"]()
    adr, ci, a, rhs = O.assemble_problem(prob, "cr")
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_CR, adr, ci)
    m.set_values(a)
    with pytest.raises(gpu.SsfError):
        m.solve(rhs[0], prec=gpu.PREC_DIAG)


@pytest.mark.parametrize("name", list(SMALL))
def test_mixed_precision(ctx, name):
    prob = SMALL[name]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    y = rhs[0]
    xo, *_ = O.scr_cg_prec(adr, ci, a, y, 50000, 1e-30)
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    x, info = m.solve(y, ni=50000, res=1e-26, precision=gpu.MIXED)
    # 1: residual test met; 4: FP64 residual floor of the system reached first
    assert info.status in (1, 4)
    assert rel(x, xo) <= 1e-8, rel(x, xo)


def test_zero_rhs_and_ni_exhausted(ctx):
    prob = SMALL["hex8_heat"]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    x, info = m.solve(np.zeros(prob.n))
    assert info.status == 3 and info.ani == 0 and not x.any()
    x, info = m.solve(rhs[0], ni=3, res=1e-30)
    assert info.status == 0 and info.ani == -1 and info.iterations == 3  # left unchanged
    xo, ani_o, ares_o, st_o = O.scr_cg_prec(adr, ci, a, rhs[0], 3, 1e-30)
    assert rel(x, xo) <= 1e-12


def test_diagonal_one_iteration(ctx):
    n = 50
    adr = np.arange(n + 1, dtype=np.int64)
    ci = np.arange(n, dtype=np.int64)
    a = np.linspace(1.0, 7.0, n)
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    y = np.random.default_rng(3).standard_normal(n)
    x, info = m.solve(y, res=1e-20)
    assert info.status == 1 and info.ani == 0
    assert rel(x, y / a) <= 1e-15


def test_deterministic(ctx):
    prob = mesh.heat_cube(10, seed=5)
    m, rhs = gpu.build_problem_matrix(ctx, prob)
    x1, i1 = m.solve(rhs[0], res=1e-24)
    x2, i2 = m.solve(rhs[0], res=1e-24)
    assert np.array_equal(x1, x2) and i1.ani == i2.ani


def test_localize_batch(ctx):
    prob = SMALL["hex8_elast"]()
    adr, ci, a, _ = O.assemble_problem(prob, "scr")
    cn = prob.elem_codes()
    ke = O.element_matrices(prob)
    m = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    m.zero()
    m.localize_batch(cn, ke.reshape(prob.ne, -1))
    assert np.abs(m.get_values() - a).max() <= 1e-12 * np.abs(a).max()


def test_cantilever_full_config(ctx):
    """Cantilever 20x20x80 hex8 (105 840 unknowns) Jacobi-PCG to res 1e-20 on both."""
    prob = mesh.cantilever()
    m, rhs = gpu.build_problem_matrix(ctx, prob)
    adr, ci = m.pattern()
    a = m.get_values()
    x, info = m.solve(rhs[0], ni=20000, res=1e-20)
    adr_o, ci_o, a_o, rhs_o = O.assemble_problem(prob, "scr")
    assert np.array_equal(adr, adr_o) and np.array_equal(ci, ci_o)
    assert np.abs(a - a_o).max() <= 1e-12 * np.abs(a_o).max()
    xo, ani_o, ares_o, st_o = O.scr_cg_prec(adr_o, ci_o, a_o, rhs_o[0], 20000, 1e-20)
    print("cantilever ani gpu", info.ani, "oracle", ani_o, "rel", rel(x, xo), flush=True)
    assert info.status == 1 and st_o == 1
    assert rel(x, xo) <= 1e-8
    assert abs(info.ani - ani_o) <= max(3, ani_o // 100)


TILINGS64 = ([(0, l, 64, 0) for l in (4, 8, 16, 32, 64)] + [(1, l, 64, 0) for l in (4, 8, 16)]
             + [(2, 0, b, i) for b in (1, 8, 64) for i in (0, 1, 2)]
             + [(2, d, b, 2) for d in (4, 8) for b in (1, 64)])
TILINGS32 = [(0, 8, 64, 0), (1, 4, 64, 0), (1, 8, 64, 1), (1, 16, 16, 1), (2, 0, 8, 0),
             (2, 0, 64, 1), (2, 0, 8, 2), (2, 0, 64, 2), (2, 4, 8, 2), (2, 8, 64, 2)]


@pytest.mark.parametrize("name", ["hex8_elast", "hex8_heat", "quad4"])
def test_spmv_tilings(ctx, name):
    """Every SpMV tiling (row-vector and stream kernels, 32/16-bit columns) gives the
    oracle's product / solution."""
    prob = SMALL[name]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    b = np.random.default_rng(2).standard_normal(prob.n)
    ref = O.mxv_scr(adr, ci, a, b)
    for t in TILINGS64:
        m.set_tune(gpu.FP64, t)
        assert rel(m.gmxv(b), ref) <= 1e-14, t
    xo, *_ = O.scr_cg_prec(adr, ci, a, rhs[0], 50000, 1e-30)
    m.set_tune(gpu.FP64, TILINGS64[-1])
    for t in TILINGS32:
        m.set_tune(gpu.MIXED, t)
        x, info = m.solve(rhs[0], ni=50000, res=1e-26, precision=gpu.MIXED)
        assert info.status in (1, 4) and rel(x, xo) <= 1e-8, (t, rel(x, xo))


def test_sell_row_patterns(ctx):
    """Row patterns (index form 2): a mesh matrix repeats few column-offset lists and
    its products equal the 32-bit form's; a pattern without repetition beyond the
    dictionary's limit keeps the column forms only."""
    prob = mesh.heat_cube(17, seed=9)  # n not a multiple of 64: rows past n in the last slice
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    b = np.random.default_rng(3).standard_normal(prob.n)
    m.set_tune(gpu.FP64, (2, 0, 64, 0))
    y0 = m.gmxv(b)
    m.set_tune(gpu.FP64, (2, 0, 64, 2))
    assert np.array_equal(m.gmxv(b), y0)  # same entries, same order
    can, act = m.spmv_bytes(gpu.FP64)
    assert act < can
    m.close()
    # every depth (value loads in flight) and block count, on heat rows (width 28)
    # and elasticity rows (width 84: several load rounds per slice, a clamped tail)
    for prob in (mesh.heat_cube(17, seed=9), mesh.elastic_column(5, 6, 7, seed=10)):
        adr, ci, a, rhs = O.assemble_problem(prob, "scr")
        m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
        m.set_values(a)
        b = np.random.default_rng(5).standard_normal(prob.n)
        m.set_tune(gpu.FP64, (2, 0, 64, 0))
        y0 = m.gmxv(b)
        for d in (2, 4, 8):
            for nb in (1, 3, 64):
                m.set_tune(gpu.FP64, (2, d, nb, 2))
                assert np.array_equal(m.gmxv(b), y0), (prob.name, d, nb)
        m.close()
    # a random pattern: every row its own offsets, more than 65535 of them
    n = 70000
    rng = np.random.default_rng(4)
    rows = [sorted(set(rng.integers(0, n, 6).tolist()) | {i}) for i in range(n)]
    adr2 = np.zeros(n + 1, np.int64)
    adr2[1:] = np.cumsum([len(r) for r in rows])
    ci2 = np.concatenate([np.array(r, np.int64) for r in rows])
    c = gpu.Matrix.from_csr(ctx, gpu.STORAGE_CR, adr2, ci2)
    c.set_values(rng.standard_normal(ci2.shape[0]))
    with pytest.raises(gpu.SsfError):
        c.set_tune(gpu.FP64, (2, 0, 64, 2))
    c.close()


def test_sell_stencil_form(ctx):
    """Stencil form (index form 3): when every row's column offsets come from one
    list of at most 32 (the heat cube's 27-point stencil) the layout gives each
    offset its own slot in every row; products and solves equal the other forms'
    (absent slots add exact zeros). A plate's offsets need a much wider layout than
    its rows: no stencil form, and asking for it is refused."""
    for prob in (mesh.heat_cube(17, seed=9), mesh.heat_cube(6, seed=10)):
        adr, ci, a, rhs = O.assemble_problem(prob, "scr")
        m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
        m.set_values(a)
        b = np.random.default_rng(6).standard_normal(prob.n)
        m.set_tune(gpu.FP64, (2, 0, 64, 0))
        y0 = m.gmxv(b)
        x0, i0 = m.solve(rhs, ni=20000, res=1e-22)
        for nb in (1, 3, 64):
            for dep in (0, 3, 4):  # one round per row, or rounds of 3 or 4 groups
                m.set_tune(gpu.FP64, (2, dep, nb, 3))
                assert np.array_equal(m.gmxv(b), y0), (prob.name, nb, dep)
        x3, i3 = m.solve(rhs, ni=20000, res=1e-22)
        assert i3.iterations == i0.iterations and np.array_equal(x3, x0)
        # the mixed solve on the stencil form
        m.set_tune(gpu.MIXED, (2, 0, 64, 3))
        xm, im = m.solve(rhs, ni=20000, res=1e-22, precision=gpu.MIXED)
        assert im.status == 1 and np.linalg.norm(xm - x0) <= 1e-8 * np.linalg.norm(x0)
        m.close()
    # a pattern whose rows share no offset list: no stencil form, and asking for
    # it is refused
    n = 3000
    rng = np.random.default_rng(12)
    low = [sorted(set(rng.integers(0, i + 1, 4).tolist()) | {i}) for i in range(n)]
    adr = np.zeros(n + 1, np.int64)
    adr[1:] = np.cumsum([len(r) for r in low])
    ci = np.concatenate([np.array(r, np.int64) for r in low])
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(rng.standard_normal(ci.shape[0]) + (ci == np.repeat(np.arange(n), np.diff(adr))) * 50)
    with pytest.raises(gpu.SsfError):
        m.set_tune(gpu.FP64, (2, 0, 64, 3))
    m.close()


def test_sell_half_stencil_form(ctx):
    """Symmetric half of the stencil form (index form 4): the lower slots of every
    row stored once, diagonal-major, and the upper entries read from the mirrored
    rows. Products equal the column forms' bit for bit at every block count, block
    size and slice map, and FP64 solves on the same block partition equal the
    stencil form's bit for bit. The mixed solve on this form runs its FP32
    recurrence in two kernels per iteration (the SpMV kernel takes the four dots of
    the step): within tolerance of the oracle in about the same iterations. The form
    is the default of a heat matrix, and refused where the storage or the stencil
    is not symmetric about the row."""
    # (cube, lower slots and diagonal): 27-point stencils; one free plane of 3 x 3 nodes
    for prob, nlow in ((mesh.heat_cube(17, seed=9), 14), (mesh.heat_cube(6, seed=10), 14),
                       (mesh.heat_cube(2, seed=3), 5)):
        adr, ci, a, rhs = O.assemble_problem(prob, "scr")
        m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
        m.set_values(a)
        n = prob.n
        # the default tiling streams the lower slots and the diagonal of every row
        can, act = m.spmv_bytes(gpu.FP64)
        assert act == nlow * 8.0 * n + 16.0 * n, (prob.name, act)
        b = np.random.default_rng(6).standard_normal(n)
        m.set_tune(gpu.FP64, (2, 0, 64, 0))
        y0 = m.gmxv(b)
        assert rel(y0, O.mxv_scr(adr, ci, a, b)) <= 1e-14
        for nb in (1, 3, 64, 512):
            for tb in (1, 2):
                for mp in (0, 1):
                    m.set_tune(gpu.FP64, (2, tb, nb, 4, mp))
                    assert np.array_equal(m.gmxv(b), y0), (prob.name, nb, tb, mp)
        for precision in (gpu.FP64, gpu.MIXED):
            m.set_tune(gpu.FP64, (2, 0, 64, 3))
            m.set_tune(gpu.MIXED, (2, 0, 64, 3))
            x3, i3 = m.solve(rhs[0], ni=20000, res=1e-22, precision=precision)
            m.set_tune(gpu.FP64, (2, 1, 64, 4))
            m.set_tune(gpu.MIXED, (2, 1, 64, 4))
            x4, i4 = m.solve(rhs[0], ni=20000, res=1e-22, precision=precision)
            if precision == gpu.FP64:
                assert i4.status == i3.status and i4.iterations == i3.iterations, prob.name
                assert np.array_equal(x4, x3), prob.name
            else:
                assert i4.status == i3.status == 1, prob.name
                assert abs(i4.iterations - i3.iterations) <= max(2, i3.iterations // 20), (
                    prob.name, i4.iterations, i3.iterations)
                assert rel(x4, x3) <= 1e-9, prob.name
        # other partitions (blocks of 512 threads, the contiguous slice map) reduce the
        # dot products in another order: solves within tolerance of the oracle
        xo, *_ = O.scr_cg_prec(adr, ci, a, rhs[0], 50000, 1e-30)
        m.set_tune(gpu.FP64, (2, 2, 8, 4, 0))
        m.set_tune(gpu.MIXED, (2, 2, 24, 4, 1))
        for precision in (gpu.FP64, gpu.MIXED):
            x4, i4 = m.solve(rhs[0], ni=20000, res=1e-22, precision=precision)
            assert i4.status in (1, 4) and rel(x4, xo) <= 1e-8, (prob.name, precision)
        # a values change rebuilds the half layout
        m.set_values(2.0 * a)
        m.set_tune(gpu.FP64, (2, 1, 64, 4))
        assert np.array_equal(m.gmxv(b), 2.0 * y0)
        m.close()
        # comprow storage (values need not be symmetric): no half form
        ac, cc, va, _ = O.assemble_problem(prob, "cr")
        c = gpu.Matrix.from_csr(ctx, gpu.STORAGE_CR, ac, cc)
        c.set_values(va)
        with pytest.raises(gpu.SsfError):
            c.set_tune(gpu.FP64, (2, 1, 64, 4))
        c.close()


@pytest.mark.parametrize("name", ["cantilever", "plate", "column"])
def test_sell_block_stencil_form(ctx, name):
    """Block stencil form: with the unknowns of a node numbered together (3 for
    hex8 elasticity, 2 for the quad4 plate), the offsets of every row counted from
    the first row of its node (27 neighbour nodes x 3 = 81, 9 x 2 = 18) form one
    list, so the layout gives each its own slot. Products equal the column forms'
    bit for bit at every block count and slice map, and the FP64 and mixed solves
    agree with the oracle. Its symmetric half (index form 4: per node the blocks
    towards the lower neighbour nodes and the diagonal block's lower triangle, the
    upper blocks read transposed from the mirrored nodes) is the default tiling
    and gives the same products and FP64 solves bit for bit."""
    # (the 40 x 40 x 4 column, z numbered fastest, spans several slices per mesh
    # plane; narrower columns have too many slices of surface nodes only for the
    # stencil form)
    prob = {"cantilever": lambda: mesh.cantilever(8, 3, 3),
            "plate": lambda: mesh.plate(20, 7, seed=11, nrhs=1),
            "column": lambda: mesh.elastic_column(40, 40, 4, seed=3)}[name]()
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
    m.set_values(a)
    B, nlow = (2, 4 * 4 + 3) if name == "plate" else (3, 13 * 9 + 6)
    can, act = m.spmv_bytes(gpu.FP64)  # the default tiling streams the half form
    assert act == nlow * 8.0 * (prob.n // B) + 16.0 * prob.n, act
    b = np.random.default_rng(7).standard_normal(prob.n)
    m.set_tune(gpu.FP64, (2, 0, 64, 0))
    y0 = m.gmxv(b)
    assert rel(y0, O.mxv_scr(adr, ci, a, b)) <= 1e-14
    x0, i0 = m.solve(rhs[0], ni=50000, res=1e-22)
    for nb in (1, 3, 64, 512):
        for mp in (0, 1):
            for dep in (0, 3, 4):  # groups of slots per round: all, 3, 4
                m.set_tune(gpu.FP64, (2, dep, nb, 3, mp))
                assert np.array_equal(m.gmxv(b), y0), (nb, mp, dep)
            for tb in (1, 2):  # the half form, blocks of 256 or 512 threads
                m.set_tune(gpu.FP64, (2, tb, nb, 4, mp))
                assert np.array_equal(m.gmxv(b), y0), (nb, mp, tb)
    # (the same 64 blocks and slice map as the column form: the same partials)
    m.set_tune(gpu.FP64, (2, 4, 64, 3, 1))
    x3, i3 = m.solve(rhs[0], ni=50000, res=1e-22)
    assert i3.iterations == i0.iterations and np.array_equal(x3, x0)
    # the half form's wavefronts take 64 nodes, not 64 rows, so its dot products
    # are reduced in another order: the solve agrees to rounding
    m.set_tune(gpu.FP64, (2, 1, 64, 4, 1))
    x4, i4 = m.solve(rhs[0], ni=50000, res=1e-22)
    assert abs(i4.iterations - i3.iterations) <= 1 and rel(x4, x3) <= 1e-12
    xo, *_ = O.scr_cg_prec(adr, ci, a, rhs[0], 50000, 1e-30)
    assert rel(x3, xo) <= 1e-8 and rel(x4, xo) <= 1e-8
    for t in ((2, 0, 64, 3, 1), (2, 1, 64, 4, 1), (2, 2, 8, 4, 0)):
        m.set_tune(gpu.MIXED, t)
        xm, im = m.solve(rhs[0], ni=50000, res=1e-22, precision=gpu.MIXED)
        assert im.status in (1, 4) and rel(xm, xo) <= 1e-8, (t, rel(xm, xo))
    # a values change rebuilds the half layout
    m.set_values(2.0 * a)
    m.set_tune(gpu.FP64, (2, 1, 64, 4))
    assert np.array_equal(m.gmxv(b), 2.0 * y0)
    m.close()


def test_autotune_keeps_parity(ctx):
    prob = mesh.heat_cube(12, seed=8)
    m, rhs = gpu.build_problem_matrix(ctx, prob)
    t64 = m.autotune(gpu.FP64, reps=2)
    t32 = m.autotune(gpu.MIXED, reps=2)
    # 16-bit columns: stream tiles only for the FP32 operator, SELL slices for both
    assert t64[0] == 2 or t64[3] == 0
    adr, ci = m.pattern()
    a = m.get_values()
    xo, *_ = O.scr_cg_prec(adr, ci, a, rhs[0], 50000, 1e-30)
    for precision in (gpu.FP64, gpu.MIXED):
        x, info = m.solve(rhs[0], ni=50000, res=1e-24, precision=precision)
        assert rel(x, xo) <= 1e-8, (precision, t64, t32)


def test_sell_slice_maps(ctx):
    """Both maps of SELL slices to wavefronts (0: a contiguous range per block;
    1: each XCD a contiguous eighth taken in turn by its wavefronts) give the same
    products in every index form, for block counts that do and do not divide
    among the 8 XCDs, and solves within tolerance of the oracle."""
    for prob in (mesh.heat_cube(17, seed=9), mesh.elastic_column(5, 6, 7, seed=10)):
        adr, ci, a, rhs = O.assemble_problem(prob, "scr")
        m = gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci)
        m.set_values(a)
        b = np.random.default_rng(7).standard_normal(prob.n)
        m.set_tune(gpu.FP64, (2, 0, 64, 0, 0))
        y0 = m.gmxv(b)
        forms = [0, 1, 2] + ([3] if prob.kind == mesh.KIND_HEX8_HEAT else [])
        for i16 in forms:
            for nb in (3, 8, 64, 512):
                for mp in (0, 1):
                    m.set_tune(gpu.FP64, (2, 4, nb, i16, mp))
                    assert np.array_equal(m.gmxv(b), y0), (prob.name, i16, nb, mp)
        xo, *_ = O.scr_cg_prec(adr, ci, a, rhs[0], 50000, 1e-30)
        for mp in (0, 1):
            t = (2, 4, 64, forms[-1], mp)
            m.set_tune(gpu.FP64, t)
            m.set_tune(gpu.MIXED, t)
            for precision in (gpu.FP64, gpu.MIXED):
                x, info = m.solve(rhs[0], ni=50000, res=1e-24, precision=precision)
                assert info.status in (1, 2, 4) and rel(x, xo) <= 1e-8, (prob.name, mp, precision)
        with pytest.raises(gpu.SsfError):
            m.set_tune(gpu.FP64, (2, 0, 64, 0, 2))
        m.close()
