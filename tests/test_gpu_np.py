"""GPU nonstationary heat (capacity assembly, gmatrix::scalgm / addgm on the
device values, one_step_linear and the d-form step of linear_nonstat_solv_dform
on device vectors) against the oracle, through the C ABI."""
import numpy as np
import pytest

import oracle as O
from ssf_amd import gpu, mesh

pytestmark = pytest.mark.gpu


def _dev(x):
    import torch
    return torch.tensor(np.ascontiguousarray(x, dtype=np.float64), device=torch.device("cuda:0"))


def _mats(ctx, prob, c, rho, from_oracle):
    """K (with its lift in rhs), C and a third matrix A of the same pattern."""
    cn = prob.elem_codes()
    K = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    C = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    A = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, cn)
    adr, ci, aK, rhs = O.assemble_problem(prob, "scr")
    aC = O.scr_assemble(prob.n, adr, ci, cn, O.capacity_matrices(prob, c, rho))
    if from_oracle:
        K.set_values(aK)
        C.set_values(aC)
    else:
        lift = K.assemble(prob)
        rhs = (prob.rhs[0] + lift)[None, :] if lift is not None else prob.rhs
        C.assemble_capacity(prob, c, rho)
    return K, C, A, adr, ci, aK, aC, rhs


def test_capacity_assembly_vs_oracle(ctx):
    prob = mesh.heat_cube(7, seed=40)
    rng = np.random.default_rng(40)
    c, rho = rng.uniform(0.5, 2.0, prob.ne), rng.uniform(0.5, 2.0, prob.ne)
    K, C, A, adr, ci, aK, aC, _ = _mats(ctx, prob, c, rho, from_oracle=False)
    a = C.get_values()
    assert np.abs(a - aC).max() <= 1e-12 * np.abs(aC).max()
    # without densities: rho = 1
    C2 = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, prob.n, prob.elem_codes())
    C2.assemble_capacity(prob, c)
    aC1 = O.scr_assemble(prob.n, adr, ci, prob.elem_codes(), O.capacity_matrices(prob, c))
    assert np.abs(C2.get_values() - aC1).max() <= 1e-12 * np.abs(aC1).max()
    for m in (K, C, A, C2):
        m.close()


def test_scale_and_add_are_the_reference_loops(ctx):
    prob = mesh.heat_cube(6, seed=41)
    c = np.full(prob.ne, 1.5)
    K, C, A, adr, ci, aK, aC, _ = _mats(ctx, prob, c, None, from_oracle=True)
    K.scale(0.37)
    K.add(1.0, C)
    ref = aK * 0.37
    ref += 1.0 * aC
    assert np.array_equal(K.get_values(), ref)
    # a matrix of another pattern is refused
    other = gpu.Matrix.from_elements(ctx, gpu.STORAGE_SCR, *(lambda p: (p.n, p.elem_codes()))(
        mesh.heat_cube(5, seed=1)))
    with pytest.raises(gpu.SsfError):
        K.add(1.0, other)
    for m in (K, C, A, other):
        m.close()


@pytest.mark.parametrize("alpha,dt", [(0.5, 0.01), (1.0, 0.05)])
def test_np_steps_vs_oracle(ctx, alpha, dt):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(9, seed=42)
    rng = np.random.default_rng(42)
    c, rho = rng.uniform(0.5, 2.0, prob.ne), rng.uniform(0.5, 2.0, prob.ne)
    K, C, A, adr, ci, aK, aC, rhs = _mats(ctx, prob, c, rho, from_oracle=True)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    Td, vd, fd = _dev(T), _dev(v), _dev(rhs[0])
    torch.cuda.synchronize()
    for step in range(4):
        st, ani, _, aA = O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], T, v, 20000, 1e-22)
        info = gpu.np_step_linear(K, C, A, alpha, dt, fd.data_ptr(), Td.data_ptr(), vd.data_ptr(),
                                  ni=20000, res=1e-22)
        assert info.status == st == 1
        assert abs(info.ani - ani) <= 2
        Tg, vg = Td.cpu().numpy(), vd.cpu().numpy()
        assert np.linalg.norm(Tg - T) <= 1e-10 * np.linalg.norm(T), step
        assert np.linalg.norm(vg - v) <= 1e-9 * np.linalg.norm(v), step
        if step == 0:
            assert np.array_equal(A.get_values(), aA)  # scalmat + addmat, bit for bit
    for m in (K, C, A):
        m.close()


@pytest.mark.parametrize("alpha,dt", [(0.5, 0.01), (1.0, 0.05), (0.75, 0.02)])
def test_np_dform_steps_vs_oracle(ctx, alpha, dt):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(9, seed=46)
    rng = np.random.default_rng(46)
    c, rho = rng.uniform(0.5, 2.0, prob.ne), rng.uniform(0.5, 2.0, prob.ne)
    K, C, A, adr, ci, aK, aC, rhs = _mats(ctx, prob, c, rho, from_oracle=True)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    Td, vd, fd = _dev(T), _dev(v), _dev(rhs[0])
    torch.cuda.synchronize()
    for step in range(4):
        st, ani, _, aA = O.np_step_linear(adr, ci, aK, aC, alpha, dt, rhs[0], T, v, 20000, 1e-22,
                                          form="d")
        info = gpu.np_step_linear(K, C, A, alpha, dt, fd.data_ptr(), Td.data_ptr(), vd.data_ptr(),
                                  ni=20000, res=1e-22, form="d")
        assert info.status == st == 1
        assert abs(info.ani - ani) <= 2
        Tg, vg = Td.cpu().numpy(), vd.cpu().numpy()
        assert np.linalg.norm(Tg - T) <= 1e-10 * np.linalg.norm(T), step
        assert np.linalg.norm(vg - v) <= 1e-8 * np.linalg.norm(v), step
        if step == 0:
            assert np.array_equal(A.get_values(), aA)
    for m in (K, C, A):
        m.close()


def test_np_dform_one_unknown_and_errors(ctx):
    torch = pytest.importorskip("torch")
    adr, ci = np.array([0, 1]), np.array([0])
    k, m, alpha, dt, T0, N = 3.0, 2.0, 0.75, 0.4, 5.0, 12
    mats = [gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci) for _ in range(3)]
    mats[0].set_values(np.array([k]))
    mats[1].set_values(np.array([m]))
    Td, vd, fd = _dev([T0]), _dev([-k / m * T0]), _dev([0.0])
    torch.cuda.synchronize()
    for _ in range(N):
        gpu.np_step_linear(mats[0], mats[1], mats[2], alpha, dt, fd.data_ptr(), Td.data_ptr(),
                           vd.data_ptr(), ni=10, res=1e-30, form="d")
    lam = k / m
    r = (1 - (1 - alpha) * dt * lam) / (1 + alpha * dt * lam)
    assert abs(Td.cpu().numpy()[0] - T0 * r ** N) <= 1e-13 * T0
    assert abs(vd.cpu().numpy()[0] + lam * Td.cpu().numpy()[0]) <= 1e-12 * T0
    with pytest.raises(gpu.SsfError):
        gpu.np_step_linear(mats[0], mats[1], mats[2], 0.0, dt, fd.data_ptr(), Td.data_ptr(),
                           vd.data_ptr(), form="d")
    with pytest.raises(gpu.SsfError):  # A must be a third matrix
        gpu.np_step_linear(mats[0], mats[1], mats[0], alpha, dt, fd.data_ptr(), Td.data_ptr(),
                           vd.data_ptr(), form="d")
    for mm in mats:
        mm.close()


def test_np_forms_agree_mixed_precision(ctx):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(12, seed=47)
    c = np.full(prob.ne, 1.0)
    K, C, A, adr, ci, aK, aC, rhs = _mats(ctx, prob, c, None, from_oracle=False)
    Tv, vv, fd = _dev(np.zeros(prob.n)), _dev(np.zeros(prob.n)), _dev(rhs[0])
    Td, vd = _dev(np.zeros(prob.n)), _dev(np.zeros(prob.n))
    torch.cuda.synchronize()
    for _ in range(3):
        gpu.np_step_linear(K, C, A, 0.5, 0.02, fd.data_ptr(), Tv.data_ptr(), vv.data_ptr(),
                           res=1e-24)
        info = gpu.np_step_linear(K, C, A, 0.5, 0.02, fd.data_ptr(), Td.data_ptr(), vd.data_ptr(),
                                  res=1e-22, precision=gpu.MIXED, form="d")
        assert info.status == 1
    a, b = Tv.cpu().numpy(), Td.cpu().numpy()
    assert np.linalg.norm(a - b) <= 1e-8 * np.linalg.norm(a)
    for m in (K, C, A):
        m.close()


def test_np_device_assembled_and_mixed_precision(ctx):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(12, seed=43)
    c = np.full(prob.ne, 1.0)
    K, C, A, adr, ci, aK, aC, rhs = _mats(ctx, prob, c, None, from_oracle=False)
    T, v = np.zeros(prob.n), np.zeros(prob.n)
    Td, vd, fd = _dev(T), _dev(v), _dev(rhs[0])
    Tm, vm = _dev(T), _dev(v)
    torch.cuda.synchronize()
    for _ in range(3):
        O.np_step_linear(adr, ci, aK, aC, 0.5, 0.02, rhs[0], T, v, 20000, 1e-22)
        gpu.np_step_linear(K, C, A, 0.5, 0.02, fd.data_ptr(), Td.data_ptr(), vd.data_ptr(),
                           res=1e-22)
        info = gpu.np_step_linear(K, C, A, 0.5, 0.02, fd.data_ptr(), Tm.data_ptr(), vm.data_ptr(),
                                  res=1e-20, precision=gpu.MIXED)
        assert info.status == 1
    for x in (Td, Tm):
        assert np.linalg.norm(x.cpu().numpy() - T) <= 1e-8 * np.linalg.norm(T)
    for m in (K, C, A):
        m.close()


def test_np_one_unknown_amplification(ctx):
    torch = pytest.importorskip("torch")
    adr, ci = np.array([0, 1]), np.array([0])
    k, m, alpha, dt, T0, N = 3.0, 2.0, 0.75, 0.4, 5.0, 12
    mats = [gpu.Matrix.from_csr(ctx, gpu.STORAGE_SCR, adr, ci) for _ in range(3)]
    mats[0].set_values(np.array([k]))
    mats[1].set_values(np.array([m]))
    Td, vd, fd = _dev([T0]), _dev([-k / m * T0]), _dev([0.0])
    torch.cuda.synchronize()
    for _ in range(N):
        gpu.np_step_linear(mats[0], mats[1], mats[2], alpha, dt, fd.data_ptr(), Td.data_ptr(),
                           vd.data_ptr(), ni=10, res=1e-30)
    lam = k / m
    r = (1 - (1 - alpha) * dt * lam) / (1 + alpha * dt * lam)
    assert abs(Td.cpu().numpy()[0] - T0 * r ** N) <= 1e-13 * T0
    for mm in mats:
        mm.close()


def test_np_reforms_a_after_values_change(ctx):
    """A is formed once per (K, C, dt alpha); scaling K forms it again."""
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(6, seed=44)
    c = np.full(prob.ne, 2.0)
    K, C, A, adr, ci, aK, aC, rhs = _mats(ctx, prob, c, None, from_oracle=True)
    Td, vd, fd = _dev(np.zeros(prob.n)), _dev(np.zeros(prob.n)), _dev(rhs[0])
    torch.cuda.synchronize()
    gpu.np_step_linear(K, C, A, 1.0, 0.1, fd.data_ptr(), Td.data_ptr(), vd.data_ptr())
    a1 = A.get_values()
    K.scale(2.0)
    gpu.np_step_linear(K, C, A, 1.0, 0.1, fd.data_ptr(), Td.data_ptr(), vd.data_ptr())
    ref = aK * 2.0
    ref *= 0.1
    ref += 1.0 * aC
    assert not np.array_equal(a1, A.get_values()) and np.array_equal(A.get_values(), ref)
    for m in (K, C, A):
        m.close()


def test_np_backward_euler_reaches_the_stationary_solution(ctx):
    torch = pytest.importorskip("torch")
    prob = mesh.heat_cube(16, seed=45)
    c = np.full(prob.ne, 1.0)
    K, C, A, adr, ci, aK, aC, rhs = _mats(ctx, prob, c, None, from_oracle=False)
    xs, _ = K.solve(rhs[0], res=1e-26)
    Td, vd, fd = _dev(np.zeros(prob.n)), _dev(np.zeros(prob.n)), _dev(rhs[0])
    torch.cuda.synchronize()
    for _ in range(3):
        gpu.np_step_linear(K, C, A, 1.0, 1e9, fd.data_ptr(), Td.data_ptr(), vd.data_ptr(),
                           res=1e-26)
    assert np.abs(Td.cpu().numpy() - xs).max() <= 1e-7 * np.abs(xs).max()
    for m in (K, C, A):
        m.close()
