"""A compiled C++ caller of the boundary: tests/cxx/gmatrix_driver drives the
gmatrix_gpu mirror (ssf_amd/host) as SIFEL's element loop drives gmatrix
(initiate from code numbers, localized per element, solve_system, gmxv). Its
assembled values, solution and product against the oracle's."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from ssf_amd import build, mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["heat", "elast", "plate"])
def test_cxx_gmatrix_caller_matches_the_oracle(tmp_path, name):
    prob = {"heat": lambda: mesh.heat_cube(7, seed=91),
            "elast": lambda: mesh.cantilever(8, 3, 3),
            "plate": lambda: mesh.plate(11, 6, seed=92, nrhs=1)}[name]()
    cn = prob.elem_codes()
    ke = O.element_matrices(prob)
    adr, ci, a, rhs = O.assemble_problem(prob, "scr")
    ne, nd = cn.shape
    res, ni = 1e-24, 20000
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        np.array([11, prob.n, ne, nd, ni], np.int64).tofile(f)
        np.array([res], np.float64).tofile(f)
        cn.astype(np.int64).tofile(f)
        ke.astype(np.float64).tofile(f)
        rhs[0].astype(np.float64).tofile(f)
    drv = build.DRIVER if os.path.exists(build.DRIVER) else build.build_driver()
    r = subprocess.run([drv, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw = open(outp, "rb").read()
    status, ani, negm = np.frombuffer(raw[:24], np.int64)
    ares = np.frombuffer(raw[24:32], np.float64)[0]
    v = np.frombuffer(raw[32:], np.float64)
    n = prob.n
    x, ax, ag = v[:n], v[n:2 * n], v[2 * n:]
    assert negm == a.shape[0]
    assert np.abs(ag - a).max() <= 1e-14 * np.abs(a).max()
    xo, ani_o, ares_o, st_o = O.scr_cg_prec(adr, ci, a, rhs[0], ni, res)
    assert status == st_o and abs(ani - ani_o) <= max(2, ani_o // 50)
    assert np.linalg.norm(x - xo) <= 1e-8 * np.linalg.norm(xo)
    # A x at the solution cancels down to rounding level: its error is bounded by
    # the rounding of the sums, i.e. relative to |A| |x|
    axo = O.mxv_scr(adr, ci, a, x)
    scale = O.mxv_scr(adr, ci, np.abs(a), np.abs(x))
    assert np.linalg.norm(ax - axo) <= 1e-13 * np.linalg.norm(scale)
