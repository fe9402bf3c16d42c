"""symcomprow patterns built on the device (ssf_matrix_create_elem_device)
against the oracle's (the reference's initiate restated; integer arrays, equal
bit for bit), and the matrices on them against matrices on host patterns."""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

CASES = {
    "hex8_elast": lambda: mesh.elastic_column(5, 4, 7, seed=51),
    "cantilever": lambda: mesh.cantilever(10, 3, 4),
    "hex8_heat": lambda: mesh.heat_cube(9, seed=52),
    "quad4": lambda: mesh.plate(37, 11, seed=53, nrhs=1),
    "heat_64": lambda: mesh.heat_cube(64, seed=54),
}


def _coupled(prob):
    """two nodes share a code number (a coupled DOF), the freed number moves to
    the end so that every unknown is still referred to"""
    free = np.nonzero((prob.dof > 0).all(axis=1))[0]
    i, j = free[1], free[len(free) // 2]
    dof = prob.dof.copy()
    old = dof[j, 0]
    dof[j, 0] = dof[i, 0]
    if old != prob.n:
        dof[dof == prob.n] = old
    prob.dof = dof
    prob.n -= 1
    return prob


@pytest.mark.parametrize("name", list(CASES) + ["coupled", "shuffled"])
def test_device_pattern_equals_the_oracle(ctx, name):
    if name == "coupled":
        prob = _coupled(CASES["hex8_elast"]())
    elif name == "shuffled":  # elements in random order: the pattern does not change
        prob = CASES["hex8_heat"]()
        prob.conn = prob.conn[np.random.default_rng(55).permutation(prob.ne)]
    else:
        prob = CASES[name]()
    cn = prob.elem_codes()
    m = gpu.Matrix.from_elements_device(ctx, prob.n, cn)
    adr, ci = m.pattern()
    adr_o, ci_o = O.scr_pattern(prob.n, cn)
    assert np.array_equal(adr, adr_o) and np.array_equal(ci, ci_o)
    m.close()


@pytest.mark.parametrize("name", ["hex8_elast", "hex8_heat", "quad4"])
def test_device_pattern_matrix_assembles_and_solves_as_the_host_one(ctx, name):
    prob = CASES[name]()
    cn = prob.elem_codes()
    mh = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    md = gpu.Matrix.from_elements_device(ctx, prob.n, cn)
    lh = mh.assemble(prob)
    ld = md.assemble(prob)
    assert np.array_equal(mh.get_values(), md.get_values())
    if lh is not None:
        assert np.array_equal(lh, ld)
    rhs = prob.rhs[0] + (lh if lh is not None else 0.0)
    # the device layout takes the stencil form (no column arrays) where the host
    # one does; then both run it, else both the 16-bit column form
    tune = [2, 8, 256, 1]
    try:
        md.set_tune(gpu.FP64, tune)
    except gpu.SsfError:
        tune = [2, 0, 256, 3]
    mh.set_tune(gpu.FP64, tune)
    md.set_tune(gpu.FP64, tune)
    xh, ih = mh.solve(rhs, ni=20000, res=1e-20)
    xd, idv = md.solve(rhs, ni=20000, res=1e-20)
    assert ih.status == idv.status == 1 and ih.iterations == idv.iterations
    assert np.array_equal(xh, xd)
    # mixed precision: the SELL-only matrix takes its FP32 copy and its norm from
    # the SELL values, equal to the host matrix's from the full CSR
    mh.set_tune(gpu.MIXED, tune)
    md.set_tune(gpu.MIXED, tune)
    xh, ih = mh.solve(rhs, ni=20000, res=1e-20, precision=gpu.MIXED)
    xd, idv = md.solve(rhs, ni=20000, res=1e-20, precision=gpu.MIXED)
    assert ih.status == idv.status == 1 and ih.iterations == idv.iterations
    assert np.array_equal(xh, xd)
    # same pattern: the matrices combine (gmatrix::addgm)
    md.add(1.0, mh)
    assert np.array_equal(md.get_values(), 2.0 * mh.get_values())
    mh.close()
    md.close()


def test_device_pattern_refuses_an_unknown_no_element_refers_to(ctx):
    prob = mesh.heat_cube(4, seed=56)
    with pytest.raises(gpu.SsfError):
        gpu.Matrix.from_elements_device(ctx, prob.n + 1, prob.elem_codes())


@pytest.mark.parametrize("name", ["heat_17", "cantilever", "quad4", "heat_9"])
def test_device_layout_stencil_form_matches_the_host_layout(ctx, name):
    """The device-built layout decides on the stencil form (offsets from the row,
    or from its node's block of 2 or 3 rows) by the host layout's rule: when the
    host matrix has it, the device matrix has it with the same slots, and the
    products are equal bit for bit; else the device matrix keeps its columns."""
    prob = {"heat_17": lambda: mesh.heat_cube(17, seed=57),
            "cantilever": lambda: mesh.cantilever(8, 3, 3),  # blocks of 3, 2.0 % of padding
            "quad4": CASES["quad4"], "heat_9": CASES["hex8_heat"]}[name]()
    cn = prob.elem_codes()
    mh = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    md = gpu.Matrix.from_elements_device(ctx, prob.n, cn)
    mh.assemble(prob, lift=False)
    md.assemble(prob, lift=False)
    b = np.random.default_rng(58).standard_normal(prob.n)
    try:
        mh.set_tune(gpu.FP64, (2, 0, 64, 3))
        host_sten = True
    except gpu.SsfError:
        host_sten = False
    try:
        md.set_tune(gpu.FP64, (2, 0, 64, 3))
        dev_sten = True
    except gpu.SsfError:
        dev_sten = False
    assert host_sten == dev_sten, (name, host_sten, dev_sten)
    if name in ("heat_17", "cantilever", "quad4", "heat_9"):
        assert dev_sten, name
    if dev_sten:
        for dep in (0, 3):
            for mp in (0, 1):
                mh.set_tune(gpu.FP64, (2, dep, 64, 3, mp))
                md.set_tune(gpu.FP64, (2, dep, 64, 3, mp))
                assert np.array_equal(md.gmxv(b), mh.gmxv(b)), (name, dep, mp)
        a = mh.get_values()
        adr, ci = mh.pattern()
        ref = O.mxv_scr(adr, ci, a, b)
        assert np.linalg.norm(md.gmxv(b) - ref) <= 1e-14 * np.linalg.norm(ref)
        with pytest.raises(gpu.SsfError):  # no column arrays on the device matrix
            md.set_tune(gpu.FP64, (2, 0, 64, 0))
    mh.close()
    md.close()
