"""Coupled DOFs (gencodnum flag k > 1: several node DOFs share one code number)
through the device assembly, against the oracle's localize / assemble_init
(GEFEL/scr.cpp:422-458, TRFEL/SRC/globmatt.cpp:1635), which adds every
contribution of every element, duplicates included, in element order.

Three couplings are covered for each element kind:
  * periodic across a thin mesh (one element across the tied axis): two nodes of
    one element share code numbers, so its entries meet in one entry;
  * periodic across a thicker mesh: elements that share no node share unknowns;
  * one unknown tied over a whole face: hundreds of elements add into the same
    entries.
"""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu


def _tie(flags, a_nodes, b_nodes, next_group):
    """give node a_nodes[k] and b_nodes[k] one coupled group per free DOF"""
    for a, b in zip(a_nodes, b_nodes):
        for j in range(flags.shape[1]):
            if flags[a, j] == 1 and flags[b, j] == 1:
                flags[a, j] = flags[b, j] = next_group
                next_group += 1
    return next_group


def _pairs(coords, axis):
    """nodes on the lower and upper face of `axis`, matched by the other coordinates"""
    lo = np.nonzero(np.isclose(coords[:, axis], coords[:, axis].min()))[0]
    hi = np.nonzero(np.isclose(coords[:, axis], coords[:, axis].max()))[0]
    other = [k for k in range(3) if k != axis]
    ko = np.lexsort(tuple(coords[lo][:, k] for k in other))
    kh = np.lexsort(tuple(coords[hi][:, k] for k in other))
    return lo[ko], hi[kh]


def elast_problem(nex, ney, nez, periodic_y=True, tie_face=False, seed=71):
    coords, conn = mesh.hex_grid(nex, ney, nez, 2.0, 1.0, 1.0)
    nn, ne = coords.shape[0], conn.shape[0]
    rng = np.random.default_rng(seed)
    flags = np.ones((nn, 3), dtype=np.int64)
    x0 = np.isclose(coords[:, 0], 0.0)
    flags[x0] = 0
    flags[x0, 0] = -1  # prescribed u_x = pv[0] on x = 0
    g = 2
    if periodic_y:
        a, b = _pairs(coords, 1)
        g = _tie(flags, a, b, g)
    if tie_face:  # u_z of the whole x = L face is one unknown
        face = np.nonzero(np.isclose(coords[:, 0], coords[:, 0].max()) & (flags[:, 2] == 1))[0]
        flags[face, 2] = g
    dof, n = mesh.codenum(flags)
    e = mesh.lognormal(rng, ne, 30e9, 0.3)
    nu = rng.uniform(0.2, 0.3, ne)
    f = rng.standard_normal((1, n)) * 1e3
    return mesh.Problem(mesh.KIND_HEX8_ELAST, coords, conn, dof, n, e, nu, np.ones(ne),
                        np.array([1e-3]), f, name=f"coupled_elast_{nex}x{ney}x{nez}")


def heat_problem(nel, nex=None, seed=72):
    nex = nel if nex is None else nex
    coords, conn = mesh.hex_grid(nex, nel, nel, 1.0, 1.0, 1.0)
    nn, ne = coords.shape[0], conn.shape[0]
    rng = np.random.default_rng(seed)
    flags = np.ones((nn, 1), dtype=np.int64)
    kz = np.rint(coords[:, 2] * nel).astype(int)
    flags[kz == 0, 0] = -1
    flags[kz == nel, 0] = -2
    a, b = _pairs(coords, 0)  # periodic in x
    _tie(flags, a, b, 2)
    dof, n = mesh.codenum(flags)
    k = mesh.lognormal(rng, ne, 1.0, 0.5)
    return mesh.Problem(mesh.KIND_HEX8_HEAT, coords, conn, dof, n, k, np.zeros(ne), np.ones(ne),
                        np.array([0.0, 100.0]), np.zeros((1, n)), name=f"coupled_heat_{nex}x{nel}")


def plate_problem(nex, ney, seed=73):
    coords, conn = mesh.quad_grid(nex, ney, float(nex) / ney, 1.0)
    nn, ne = coords.shape[0], conn.shape[0]
    rng = np.random.default_rng(seed)
    flags = np.ones((nn, 2), dtype=np.int64)
    flags[np.isclose(coords[:, 0], 0.0)] = 0
    a, b = _pairs(coords, 1)  # periodic in y
    _tie(flags, a, b, 2)
    dof, n = mesh.codenum(flags)
    e = mesh.lognormal(rng, ne, 30e9, 0.3)
    f = rng.standard_normal((1, n))
    return mesh.Problem(mesh.KIND_QUAD4_PSTRESS, coords, conn, dof, n, e, np.full(ne, 0.25),
                        np.ones(ne), np.zeros(0), f, name=f"coupled_plate_{nex}x{ney}")


CASES = {
    "elast_in_element": lambda: elast_problem(5, 1, 3),
    "elast_across": lambda: elast_problem(4, 3, 3),
    "elast_face_tie": lambda: elast_problem(3, 12, 12, periodic_y=False, tie_face=True),
    "heat_in_element": lambda: heat_problem(4, nex=1),
    "heat_across": lambda: heat_problem(5),
    "plate_in_element": lambda: plate_problem(7, 1),
    "plate_across": lambda: plate_problem(9, 6),
}


def _has_repeats(cn):
    for row in cn:
        p = row[row > 0]
        if np.unique(p).size < p.size:
            return True
    return False


@pytest.mark.parametrize("name", list(CASES))
def test_coupled_assembly_and_solve_match_the_oracle(ctx, name):
    prob = CASES[name]()
    cn = prob.elem_codes()
    if "in_element" in name:
        assert _has_repeats(cn)
    m, rhs = gpu.build_problem_matrix(ctx, prob)
    adr, ci, a, rhs_o = O.assemble_problem(prob, "scr")
    adr_g, ci_g = m.pattern()
    assert np.array_equal(adr_g, adr) and np.array_equal(ci_g, ci)
    a_g = m.get_values()
    assert np.abs(a_g - a).max() <= 1e-12 * np.abs(a).max()
    scale = max(np.abs(rhs_o).max(), 1e-300)
    assert np.abs(rhs - rhs_o).max() <= 1e-12 * scale
    xo, ani_o, _, st_o = O.scr_cg_prec(adr, ci, a, rhs_o[0], 40000, 1e-24)
    x, info = m.solve(rhs[0], ni=40000, res=1e-24)
    assert st_o in (1, 2) and info.status == st_o
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    m.close()


@pytest.mark.parametrize("name", ["elast_in_element", "heat_across"])
def test_coupled_device_pattern_assembles_as_the_host_one(ctx, name):
    prob = CASES[name]()
    cn = prob.elem_codes()
    mh = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    md = gpu.Matrix.from_elements_device(ctx, prob.n, cn)
    lh = mh.assemble(prob)
    ld = md.assemble(prob)
    assert np.array_equal(mh.get_values(), md.get_values())
    if lh is not None:
        assert np.array_equal(lh, ld)
    mh.close()
    md.close()


@pytest.mark.parametrize("name", ["cantilever_20x10x10", "elast_in_element", "heat_across"])
def test_localize_batch_matches_the_oracle(ctx, name):
    """gmatrix::localized of the oracle's element matrices, on meshes with many
    boundary elements (codes <= 0 are skipped, not shared)"""
    prob = mesh.cantilever(20, 10, 10) if name.startswith("cantilever") else CASES[name]()
    cn = prob.elem_codes()
    ke = O.element_matrices(prob)
    adr, ci = O.scr_pattern(prob.n, cn)
    a = O.scr_assemble(prob.n, adr, ci, cn, ke)
    m = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    m.localize_batch(cn, ke.reshape(prob.ne, -1))
    assert np.abs(m.get_values() - a).max() <= 1e-14 * np.abs(a).max()
    m.close()
