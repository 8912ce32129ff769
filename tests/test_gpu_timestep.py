"""The callers of the operator apply on the device: Kutta RK3 (eager and
HIP-graph replay), backward Euler by FP64-refined mixed-precision GMRES, and
the apply_A / time_advance / asgard mirror -- checked against the reference's
gold trajectories and the CPU oracle."""
import numpy as np
import pytest

import oracle as O
from kmx import coeffs as K
from kmx import grid as G
from test_coeffs_gold import central_flux

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def setup(pde, deg, grid_type="SG", max_lev=8):
    lev_vec = [d.lev for d in pde.dims]
    tab = G.hash_table_sparse_nd(lev_vec, max_lev, grid_type)
    mats = K.restrict(K.get_coeff_mats(pde, deg, 0.0, max_lev), pde, deg)
    return tab, mats


@pytest.mark.parametrize("name,pfn,lev,deg", [("continuity2_sg_l2_d2", K.continuity2, 2, 2),
                                              ("continuity2_sg_l4_d3", K.continuity2, 4, 3),
                                              ("continuity3_sg_l2_d2", K.continuity3, 2, 2)])
def test_rk3_on_gpu_matches_gold(kmx_gpu, gold, name, pfn, lev, deg):
    import torch

    from kmx import timestep as TS

    pde = central_flux(pfn(lev))
    tab, mats = setup(pde, deg)
    plan = kmx_gpu.Plan(mats, tab.idx1d, deg)
    src = TS.SeparableSource(pde, deg, tab.idx1d)
    f = torch.from_numpy(K.md_eval_function(pde, deg, tab.idx1d, pde.initial_conditions, 0.0)).cuda()
    dt = pde.set_dt(pde, 0.01)
    for i in range(5):
        f = TS.rk3_step(lambda v: plan.apply(v), f, i * dt, dt, src)
        g = gold[f"time_advance/{name}_t{i}"].ravel()
        assert rel(f.cpu().numpy(), g) <= 1e-12, i


def test_rk3_graph_replays_match_eager(kmx_gpu):
    """The captured step, replayed with eager device work in between, gives the
    eager result on every replay (no memset nodes, state updated in place)."""
    import torch

    from kmx import timestep as TS

    pde = K.continuity2(3)
    tab, mats = setup(pde, 2)
    plan = kmx_gpu.Plan(mats, tab.idx1d, 2)
    src = TS.SeparableSource(pde, 2, tab.idx1d)
    rng = np.random.default_rng(0)
    f0 = torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
    dt = pde.set_dt(pde, 0.01)
    g = TS.RK3Graph(plan, f0, dt, src)
    fe = f0.clone()
    for i in range(6):
        fg = g.step(i * dt)
        junk = torch.randn(4096, device="cuda") * 2  # eager work between replays
        torch.cuda.synchronize()
        fe = TS.rk3_step(lambda v: plan.apply(v), fe, i * dt, dt, src)
        assert rel(fg.cpu().numpy(), fe.cpu().numpy()) <= 1e-14, i
        del junk


def test_backward_euler_gmres_matches_gold(kmx_gpu, gold):
    """time_advance.m:267-349 semantics; FP32 inner GMRES refined to 1e-12 in FP64."""
    import torch

    from kmx import timestep as TS

    pde = central_flux(K.continuity2(2))
    tab, mats = setup(pde, 2)
    plan = kmx_gpu.Plan(mats, tab.idx1d, 2)
    src = TS.SeparableSource(pde, 2, tab.idx1d)
    f = torch.from_numpy(K.md_eval_function(pde, 2, tab.idx1d, pde.initial_conditions, 0.0)).cuda()
    dt = pde.set_dt(pde, 0.01)
    for i in range(5):
        f, hist, inner = TS.backward_euler(plan, f, dt, src((i + 1) * dt), tol=1e-12)
        assert hist[-1] <= 1e-12
        g = gold[f"time_advance/continuity2_implicit_l2_d2_t{i}"].ravel()
        assert rel(f.cpu().numpy(), g) <= 1e-11, (i, hist)


def test_backward_euler_refinement_reaches_fp64_residual(kmx_gpu):
    """A stiff diffusion step (dt >> explicit limit): FP32 inner applies alone
    stall near 1e-7; refinement with FP64 residuals reaches 1e-10."""
    import torch

    from kmx import timestep as TS

    pde = K.diffusion_nd(2, 4)
    tab, mats = setup(pde, 3)
    plan = kmx_gpu.Plan(mats, tab.idx1d, 3)
    rng = np.random.default_rng(1)
    f0 = torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
    dt = 100 * pde.set_dt(pde, 0.01)
    f1, hist, inner = TS.backward_euler(plan, f0, dt, tol=1e-10, inner_tol=1e-5)
    assert hist[-1] <= 1e-10, hist
    assert len(hist) >= 2  # at least one refinement sweep was needed
    # against a dense FP64 solve of the same system (oracle-assembled A)
    A = O.build_A(O.ApplyInputs(mats, tab.idx1d, 3))
    ref = np.linalg.solve(np.eye(A.shape[0]) - dt * A, f0.cpu().numpy())
    assert rel(f1.cpu().numpy(), ref) <= 1e-8


def test_2d_continuity2_rk3(kmx_gpu):
    """2D continuity (pdes/continuity2.m, LF = -1), SG lev 3, deg 2, explicit RK3,
    5 steps, CFL 0.01: device run == oracle run; L2 error vs the exact solution."""
    import torch

    from kmx import asgard

    err, fval = asgard.asgard(K.continuity2, lev=3, deg=2, num_steps=5, CFL=0.01)
    pde = K.continuity2(3)
    tab, mats = setup(pde, 2)
    inp = O.ApplyInputs(mats, tab.idx1d, 2)
    f = K.md_eval_function(pde, 2, tab.idx1d, pde.initial_conditions, 0.0)
    dt = pde.set_dt(pde, 0.01)
    src = lambda t: K.md_eval_function(pde, 2, tab.idx1d, pde.sources, t)  # noqa: E731
    for i in range(5):
        t = i * dt
        k1 = O.apply_A(inp, f) + src(t)
        y1 = f + dt * 0.5 * k1
        k2 = O.apply_A(inp, y1) + src(t + dt / 2)
        y2 = f + dt * (-1.0 * k1 + 2.0 * k2)
        k3 = O.apply_A(inp, y2) + src(t + dt)
        f = f + dt * (k1 / 6 + 2 * k2 / 3 + k3 / 6)
    assert rel(fval, f) <= 1e-13
    assert err is not None and err < 1e-3  # asgard_test.m-style accuracy for this tiny grid
    # the mirror's apply_A on host vectors equals the device apply
    pde2, table, A_data, f0, _ = asgard.build(K.continuity2, asgard.OPTS(lev=3, deg=2))
    x = np.random.default_rng(3).standard_normal(f0.size)
    y = asgard.apply_A(pde2, asgard.OPTS(lev=3, deg=2), A_data, x, 2)
    assert rel(y, O.apply_A(O.ApplyInputs(pde2.mats, table.idx1d, 2), x)) <= 1e-13
    del torch


@pytest.mark.parametrize("name,pfn,lev,deg", [("diffusion1_sg_l2_d2", K.diffusion1_gold, 2, 2),
                                              ("diffusion1_sg_l3_d3", K.diffusion1_gold, 3, 3),
                                              ("diffusion2_sg_l2_d2", K.diffusion2, 2, 2),
                                              ("diffusion2_sg_l3_d3", K.diffusion2, 3, 3)])
def test_rk3_dirichlet_on_gpu_matches_gold(kmx_gpu, gold, name, pfn, lev, deg):
    """Non-periodic RK3 on the device: the inhomogeneous Dirichlet terms
    (boundary_condition_vector.m) enter every stage through the separable source
    (eager steps and graph replays), the states after 1..5 steps match the gold
    trajectory and the oracle's RK3 within 1e-12."""
    import torch

    from kmx import asgard as AS
    from kmx import timestep as TS

    opts = AS.OPTS(lev=lev, deg=deg)
    pde, table, A_data, f0, dt = AS.build(pfn, opts)
    src = TS.SeparableSource(pde, deg, table.idx1d)
    assert src.num_boundary == 2 * pde.ndim
    plan = kmx_gpu.Plan(pde.mats, table.idx1d, deg)
    inp = O.ApplyInputs(pde.mats, table.idx1d, deg)
    f = torch.from_numpy(f0).cuda()
    g3 = TS.RK3Graph(plan, f, dt, source=src)
    fo = f0.copy()
    for i in range(5):
        t = i * dt
        f = AS.time_advance(pde, opts, A_data, f, t, dt, deg, table)  # the reference-named entry
        fg = g3.step(t)
        rhs = lambda tt: src.host(tt)  # noqa: E731
        k1 = O.apply_A(inp, fo) + rhs(t)
        y1 = fo + dt * 0.5 * k1
        k2 = O.apply_A(inp, y1) + rhs(t + dt / 2)
        y2 = fo + dt * (-k1 + 2 * k2)
        k3 = O.apply_A(inp, y2) + rhs(t + dt)
        fo = fo + dt * (k1 / 6 + 2 * k2 / 3 + k3 / 6)
        g = gold[f"time_advance/{name}_t{i}"].ravel()
        for got in (f.cpu().numpy(), fg.cpu().numpy()):
            assert rel(got, g) <= 1e-12, (i, rel(got, g))
            assert rel(got, fo) <= 1e-12, i


@pytest.mark.parametrize("n,k", [(7, 0), (300, 3), (2000, 12), (32064, 17), (32768, 40), (32769, 5), (40000, 70),
                                 (100003, 39), (5000, 126), (5000, 127)])
def test_arnoldi_step_matches_two_pass_gram_schmidt(kmx_gpu, n, k):
    """kmx_arnoldi_step against the same step in numpy: w = s0 v_k + s1 u, two
    classical Gram-Schmidt passes against v_0..v_k, h = both passes' coefficients,
    v_{k+1} = w / |w|; h within 1e-13 of |w|, v_{k+1} within 1e-13, and v_{k+1}
    orthogonal to v_0..v_k to 1e-14.  Sizes on both sides of n = 32768 (chunks of
    at most 256 values and k <= 126: rows held in registers; otherwise rows
    streamed) and every row bucket of the register kernels (k + 2 <= 8, 16, 32,
    64, 128).  A u in the span of v_0..v_k (s1 = 1, s0 = 0)
    breaks down: |w| ~ 0."""
    import torch

    from kmx import _lib
    from kmx.plan import arnoldi_step

    rng = np.random.default_rng(n + k)
    q, _ = np.linalg.qr(rng.standard_normal((n, k + 1)))
    V = np.zeros((k + 2, n))
    V[: k + 1] = q.T
    u = rng.standard_normal(n)
    s0, s1 = 1.0, -0.37
    w = s0 * V[k] + s1 * u
    h1 = V[: k + 1] @ w
    w = w - V[: k + 1].T @ h1
    h2 = V[: k + 1] @ w
    w = w - V[: k + 1].T @ h2
    nrm = np.linalg.norm(w)
    Vd = torch.from_numpy(V).cuda()
    hd = torch.zeros(k + 2, dtype=torch.float64, device="cuda")
    work = torch.empty(_lib.arnoldi_work(n), dtype=torch.float64, device="cuda")
    arnoldi_step(n, k, Vd, torch.from_numpy(u).cuda(), s0, s1, hd, work)
    h = hd.cpu().numpy()
    Vn = Vd.cpu().numpy()
    assert np.max(np.abs(h[: k + 1] - (h1 + h2))) <= 1e-13 * max(1.0, np.abs(h1).max())
    assert abs(h[k + 1] - nrm) <= 1e-13 * nrm
    assert np.max(np.abs(Vn[k + 1] - w / nrm)) <= 1e-13
    assert np.max(np.abs(Vn[: k + 1] @ Vn[k + 1])) <= 1e-14 * np.sqrt(n)
    assert np.array_equal(Vn[: k + 1], V[: k + 1])  # the basis rows are left alone
    # breakdown: u in the span of the basis
    u2 = V[: k + 1].T @ rng.standard_normal(k + 1)
    arnoldi_step(n, k, Vd, torch.from_numpy(u2).cuda(), 0.0, 1.0, hd, work)
    h = hd.cpu().numpy()
    assert h[k + 1] <= 1e-13 * np.linalg.norm(u2)


def test_arnoldi_step_rejects_bad_arguments(kmx_gpu):
    import torch

    from kmx import _lib

    V = torch.zeros((3, 10), dtype=torch.float64, device="cuda")
    u = torch.zeros(10, dtype=torch.float64, device="cuda")
    h = torch.zeros(3, dtype=torch.float64, device="cuda")
    work = torch.empty(_lib.arnoldi_work(10), dtype=torch.float64, device="cuda")
    from kmx.plan import arnoldi_step

    with pytest.raises(ValueError):
        arnoldi_step(10, 2, V, u, 1.0, 1.0, h, work)  # V has k+1 rows, not k+2
    with pytest.raises(kmx_gpu.KmxError):
        arnoldi_step(10, _lib.ARNOLDI_MAX_K + 1, torch.zeros((_lib.ARNOLDI_MAX_K + 3, 10), dtype=torch.float64,
                                                             device="cuda"), u,
                     1.0, 1.0, torch.zeros(_lib.ARNOLDI_MAX_K + 3, dtype=torch.float64, device="cuda"), work)


def test_backward_euler_second_call_reuses_the_solver(kmx_gpu):
    """The GMRES solver is kept per plan and step size: a second BE step with the
    same dt reuses its graphs and gives the same FP64-refined result."""
    import torch

    from kmx import timestep as TS

    pde = K.continuity2(4)
    mats = K.get_coeff_mats(pde, 3)
    tab = G.hash_table_sparse_nd([4, 4], 4, "SG")
    plan = kmx_gpu.Plan(mats, tab.idx1d, 3, None, None, index_base=1, fp32=True)
    f0 = torch.from_numpy(np.random.default_rng(1).standard_normal(plan.dof)).cuda()
    dt = 50 * pde.set_dt(pde, 0.01)
    f1, hist1, it1 = TS.backward_euler(plan, f0, dt, tol=1e-10)
    assert len(TS._SOLVERS[plan]) == 1
    f2, hist2, it2 = TS.backward_euler(plan, f0, dt, tol=1e-10)
    assert len(TS._SOLVERS[plan]) == 1
    assert hist1[-1] <= 1e-10 and hist2[-1] <= 1e-10
    assert torch.equal(f1, f2) and it1 == it2
