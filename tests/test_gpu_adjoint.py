"""Differentiation through the device apply (3D variable-coefficient advection-diffusion, lev 5 deg 3): the adjoint kernel
behind a torch.autograd.Function, RK3 with checkpointed recomputation,
gradients of the final-time L2 misfit w.r.t. the initial coefficients and the
per-dimension scalings, checked against central finite differences in FP64."""
import numpy as np
import pytest

from kmx import coeffs as K
from kmx import grid as G

pytestmark = pytest.mark.gpu


def build(lev, deg):
    pde = K.advection_diffusion_varcoeff(3, lev)
    table = G.hash_table_nd([lev] * 3, "SG")
    rp, cols = G.connect_nd(table, lev, lev)
    mats = K.get_coeff_mats(pde, deg, 0.0, None)
    return pde, table, rp, cols, mats


@pytest.mark.parametrize("lev,deg,nsteps,every,sweep", [(3, 2, 6, 2, False), (5, 3, 20, 5, False), (3, 2, 6, 2, True)])
def test_gradient_vs_central_fd(kmx_gpu, lev, deg, nsteps, every, sweep):
    """sweep: forward and adjoint applies on the separable dimension sweep (the
    terms of each dimension folded into one matrix, refolded at every scale change)."""
    import torch

    from kmx import autograd as AG

    pde, table, rp, cols, mats = build(lev, deg)
    plan = kmx_gpu.Plan(mats, table.idx1d, deg, rp, cols, adjoint=True, level_blocks=sweep)
    if sweep:
        plan.set_sweep("fp64", True)
    groups = [[d, 3 + d] for d in range(3)]  # advection + diffusion terms of dim d
    op = AG.ScaledOperator(plan, groups)
    dt = pde.set_dt(pde, 0.01)
    f0 = torch.from_numpy(K.md_eval_function(pde, deg, table.idx1d, pde.initial_conditions, 0.0)).cuda()
    rng = np.random.default_rng(7)
    target = f0 + 0.1 * torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
    alpha0 = torch.tensor([1.0, 0.8, 1.2], dtype=torch.float64, device="cuda")

    def loss(f, a):
        return AG.misfit(AG.rk3_trajectory(op, f, a, dt, nsteps, every), target)

    f = f0.clone().requires_grad_(True)
    a = alpha0.clone().requires_grad_(True)
    L = loss(f, a)
    L.backward()
    gf, ga = f.grad.detach(), a.grad.detach()
    h = 1e-6
    # central differences carry O(h^2) truncation and eps*|L|/h cancellation error
    slack = 50 * np.finfo(float).eps * abs(float(L.detach())) / h
    with torch.no_grad():
        v = torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
        v /= torch.linalg.vector_norm(v)
        fd = (loss(f0 + h * v, alpha0) - loss(f0 - h * v, alpha0)) / (2 * h)
        ad = torch.dot(gf, v)
        assert abs(float(fd - ad)) <= 1e-6 * abs(float(ad)) + slack, (float(fd), float(ad), slack)
        for g in range(3):
            e = torch.zeros(3, dtype=torch.float64, device="cuda")
            e[g] = h
            fdg = (loss(f0, alpha0 + e) - loss(f0, alpha0 - e)) / (2 * h)
            assert abs(float(fdg - ga[g])) <= 1e-6 * abs(float(ga[g])) + slack, (g, float(fdg), float(ga[g]), slack)


def test_adjoint_function_matches_transpose(kmx_gpu):
    """Backward of one apply: dL/dx = A^T gy and dL/dalpha_g = gy . A_g x."""
    import torch

    from kmx import autograd as AG

    pde, table, rp, cols, mats = build(3, 2)
    plan = kmx_gpu.Plan(mats, table.idx1d, 2, rp, cols, adjoint=True)
    op = AG.ScaledOperator(plan, [[0, 3], [1, 4], [2, 5]])
    rng = np.random.default_rng(1)
    x = torch.from_numpy(rng.standard_normal(plan.dof)).cuda().requires_grad_(True)
    alpha = torch.tensor([0.5, 1.5, 2.0], dtype=torch.float64, device="cuda", requires_grad=True)
    gy = torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
    y = AG.apply(op, x, alpha)
    (y * gy).sum().backward()
    # reference: dense A_g from unit vectors is too big; use the bilinear identity
    ref_ga = []
    for g in range(3):
        ref_ga.append(float(torch.dot(gy, op.group_apply(x.detach(), g))))
    assert np.allclose(alpha.grad.cpu().numpy(), ref_ga, rtol=1e-13)
    w = torch.from_numpy(rng.standard_normal(plan.dof)).cuda()
    op.set_scales([0.5, 1.5, 2.0])
    Aw = plan.apply(w)
    assert abs(float(torch.dot(x.grad, w) - torch.dot(gy, Aw))) <= 1e-12 * float(
        torch.linalg.vector_norm(gy) * torch.linalg.vector_norm(Aw) + 1)
