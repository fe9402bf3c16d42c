"""Served solves: requests admitted under an HBM budget and solved with
continuous batching; each result against the single-system solve and the oracle,
and the same alone and in the mix."""
import time

import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu

RES = 1e-20


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def requests():
    return [
        mesh.heat_cube(6, seed=31),
        mesh.elastic_column(3, 4, 6, seed=32),
        mesh.plate(9, 5, seed=33, nrhs=1),
        mesh.cantilever(8, 3, 3),
        mesh.heat_cube(9, seed=34),
        mesh.plate(30, 12, seed=35, nrhs=1),
    ]


def test_server_matches_single_and_oracle(ctx):
    probs = requests()
    srv = gpu.Server(budget_bytes=1 << 30, res=RES, chunk=16)
    try:
        tickets = [srv.submit(p) for p in probs]
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    # the heat cubes (27-point stencils) and the cantilever (27 nodes of 3
    # unknowns) solve on the half stencil form; the 3 x 4 x 6 column (too much
    # surface for the stencil form) and the plates keep their slices
    assert st["half_form_requests"] == 3, st
    for p, (x, info, times) in zip(probs, out):
        m, rhs = gpu.build_problem_matrix(ctx, p)
        x1, i1 = m.solve(rhs[0], ni=100000, res=RES)
        assert info.status == i1.status == 1
        assert abs(info.iterations - i1.iterations) <= max(2, i1.iterations // 50)
        assert rel(x, x1) <= 1e-8, (p.name, rel(x, x1))
        assert times[0] <= times[1] <= times[2] <= times[3]
    # against the oracle on the first request
    adr, ci, a, rhs_o = O.assemble_problem(probs[0], "scr")
    xo, ani_o, _, st_o = O.scr_cg_prec(adr, ci, a, rhs_o[0], 100000, RES)
    assert st_o == out[0][1].status and abs(out[0][1].ani - ani_o) <= 2
    assert rel(out[0][0], xo) <= 1e-8


def test_server_alone_equals_in_mix():
    probs = requests()
    target = probs[3]
    srv = gpu.Server(budget_bytes=1 << 30, res=RES, chunk=8)
    try:
        x_alone, i_alone, _ = srv.wait(srv.submit(target))
    finally:
        srv.close()
    srv = gpu.Server(budget_bytes=1 << 30, res=RES, chunk=8, prep_threads=3)
    try:
        tickets = []
        for k, p in enumerate(probs):
            tickets.append(srv.submit(p))
            time.sleep(0.01 * k)  # arrivals while others are iterating
        tickets.append(srv.submit(target))
        for p in probs[:3]:
            tickets.append(srv.submit(p))
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    # the target appears twice in the mix (as request 3 and as request 6)
    for j in (3, 6):
        x, info, _ = out[j]
        assert np.array_equal(x, x_alone)
        assert info.iterations == i_alone.iterations and info.status == i_alone.status
    # the repeated requests give the same results as their first runs
    for j in range(3):
        assert np.array_equal(out[7 + j][0], out[j][0])
    assert st["finished"] == len(tickets)
    assert st["reserved_bytes"] == 0


def test_server_budget_serialises_and_rejects():
    probs = requests()
    big = mesh.heat_cube(9, seed=34)
    # about one request's bytes: requests wait for the previous ones to retire
    srv = gpu.Server(budget_bytes=8 << 20, res=RES, chunk=16)
    try:
        tickets = [srv.submit(p) for p in probs[:3]]
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
        assert st["peak_bytes"] <= 8 << 20
        for x, info, _ in out:
            assert info.status == 1
        huge = mesh.heat_cube(40, seed=1)
        t = srv.submit(huge)
        with pytest.raises(gpu.SsfError):
            srv.wait(t)
        # the server keeps serving after a rejected request
        x, info, _ = srv.wait(srv.submit(big))
        assert info.status == 1
    finally:
        srv.close()


def test_server_ni_and_zero_rhs():
    p = mesh.cantilever(8, 3, 3)
    srv = gpu.Server(budget_bytes=1 << 30, res=RES, ni=20, chunk=16)
    try:
        x, info, _ = srv.wait(srv.submit(p))
        assert info.status == 0 and info.iterations == 20
        z = mesh.plate(9, 5, seed=33, nrhs=1)
        x0, i0, _ = srv.wait(srv.submit(z, rhs=np.zeros(z.n)))
        assert i0.status == 3 and i0.iterations == 0 and not np.any(x0)
    finally:
        srv.close()


def test_server_mixed_matches_oracle_and_alone_equals_in_mix():
    """The mixed-precision served form: every result within 1e-8 of the oracle's
    FP64 cg_prec solution, and bitwise the same alone and among other requests."""
    probs = requests()
    target = probs[1]
    kw = dict(budget_bytes=1 << 30, res=1e-22, chunk=16, precision=gpu.MIXED)
    srv = gpu.Server(**kw)
    try:
        x_alone, i_alone, _ = srv.wait(srv.submit(target))
    finally:
        srv.close()
    srv = gpu.Server(prep_threads=3, **kw)
    try:
        tickets = []
        for k, p in enumerate(probs):
            tickets.append(srv.submit(p))
            time.sleep(0.01 * k)
        tickets.append(srv.submit(target))
        out = [srv.wait(t) for t in tickets]
        st = srv.stats()
    finally:
        srv.close()
    for p, (x, info, _) in zip(probs, out):
        adr, ci, a, rhs_o = O.assemble_problem(p, "scr")
        xo, _, _, st_o = O.scr_cg_prec(adr, ci, a, rhs_o[0], 100000, 1e-30)
        assert info.status in (1, 4) and info.outer >= 1, (p.name, info.status)
        assert rel(x, xo) <= 1e-8, (p.name, rel(x, xo))
    for j in (1, len(probs)):
        x, info, _ = out[j]
        assert np.array_equal(x, x_alone)
        assert info.iterations == i_alone.iterations and info.outer == i_alone.outer
    assert st["reserved_bytes"] == 0
    assert st["half_form_requests"] == 3, st  # the target (3 x 4 x 6 column) keeps its slices


def test_server_mixed_ni_and_zero_rhs():
    p = mesh.cantilever(8, 3, 3)
    srv = gpu.Server(budget_bytes=1 << 30, res=1e-22, ni=20, chunk=16, precision=gpu.MIXED)
    try:
        x, info, _ = srv.wait(srv.submit(p))
        assert info.status == 0 and info.iterations >= 20
        z = mesh.plate(9, 5, seed=33, nrhs=1)
        x0, i0, _ = srv.wait(srv.submit(z, rhs=np.zeros(z.n)))
        assert i0.status == 3 and i0.iterations == 0 and not np.any(x0)
    finally:
        srv.close()
