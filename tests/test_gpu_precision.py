"""Precision study on the 6D diffusion operator: pdes/diffusion2.m's LDG
chain generalised to 6 dims (6 terms, one per dim), sparse grid lev 4
(hash_table_nD + connect_nD: N = 501, P = 247,785), deg 2 (64-value blocks).

* FP64 apply vs the oracle: 1e-13 relative (2-norm).
* FP32 apply (FP32 operands and accumulation, FP64 x / y) vs the oracle:
  the error is recorded; bar 5e-6 (FP32 unit roundoff 6e-8 times the
  accumulation depth).
* Backward Euler (time_advance.m:328-349's system) with FP64 GMRES residuals
  around FP32 inner applies, refined to an FP64 relative residual of 1e-10,
  against the same system solved with FP64 applies only (to 1e-13).
"""
import numpy as np
import pytest

import oracle as O
from kmx import coeffs as K
from kmx import grid as G

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def diffusion6(kmx_gpu):
    pde = K.diffusion_nd(6, 4)
    mats = K.get_coeff_mats(pde, 2)
    tab = G.hash_table_nd([4] * 6, "SG")
    rp, cols = G.connect_nd(tab, 4, 4, "SG")
    plan = kmx_gpu.Plan(mats, tab.idx1d, 2, rp, cols, index_base=1, fp32=True)
    assert (plan.N, plan.P, plan.B, plan.T) == (501, 247785, 64, 6)
    return pde, mats, tab, rp, cols, plan


def test_6d_apply_fp64_and_fp32_vs_oracle(diffusion6):
    import torch

    pde, mats, tab, rp, cols, plan = diffusion6
    x = np.random.default_rng(4).standard_normal(plan.dof)
    yo = O.apply_A(O.ApplyInputs(mats, tab.idx1d, 2, rp, cols), x, nthreads=8)
    xd = torch.from_numpy(x).cuda()
    e64 = rel(plan.apply(xd, prec="fp64").cpu().numpy(), yo)
    e32 = rel(plan.apply(xd, prec="fp32").cpu().numpy(), yo)
    print(f"6D lev4 deg2 diffusion: FP64 rel err {e64:.2e}, FP32 rel err {e32:.2e}")
    assert e64 <= 1e-13
    assert e32 <= 5e-6


def test_6d_backward_euler_fp32_inner_refined_to_fp64(diffusion6):
    import torch

    from kmx import timestep as TS

    pde, mats, tab, rp, cols, plan = diffusion6
    f0 = torch.from_numpy(np.random.default_rng(5).standard_normal(plan.dof)).cuda()
    dt = 100 * pde.set_dt(pde, 0.01)  # far beyond the explicit limit
    f1, hist, inner = TS.backward_euler(plan, f0, dt, tol=1e-10, inner_prec="fp32", inner_tol=1e-5)
    print(f"BE 6D: FP64 residual history {hist}, {inner} FP32 inner iterations")
    assert hist[-1] <= 1e-10, hist
    ref, hist64, _ = TS.backward_euler(plan, f0, dt, tol=1e-13, inner_prec="fp64", inner_tol=1e-13)
    assert hist64[-1] <= 1e-13, hist64
    assert rel(f1.cpu().numpy(), ref.cpu().numpy()) <= 1e-8


def test_6d_apply_bf16_mfma_vs_oracle(kmx_gpu, diffusion6):
    """BF16 operands on the matrix cores (v_mfma_f32_16x16x16_bf16) with FP32
    accumulation, on the separable dimension sweep: the error against the FP64
    oracle is recorded; bar 3e-2 (BF16 unit roundoff 3.9e-3 on both operands,
    FP32 sums), and above 1e-5 (the operands really are rounded to BF16)."""
    import torch

    pde, mats, tab, rp, cols, _ = diffusion6
    plan = kmx_gpu.Plan(mats, tab.idx1d, 2, rp, cols, index_base=1, fp32=True, adjoint=True, level_blocks=True)
    assert plan.sweep_info()["separable"] and plan.sweep("bf16")["available"]
    x = np.random.default_rng(4).standard_normal(plan.dof)
    yo = O.apply_A(O.ApplyInputs(mats, tab.idx1d, 2, rp, cols), x, nthreads=8)
    xd = torch.from_numpy(x).cuda()
    e16 = rel(plan.apply(xd, prec="bf16").cpu().numpy(), yo)
    e32 = rel(plan.apply(xd, prec="fp32").cpu().numpy(), yo)
    print(f"6D lev4 deg2 diffusion: BF16-MFMA rel err {e16:.2e} (FP32 {e32:.2e})")
    assert 1e-5 < e16 <= 3e-2
    # the adjoint apply runs the transposed BF16 values: against the FP64 pair-path adjoint
    wd = torch.from_numpy(np.random.default_rng(6).standard_normal(plan.dof)).cuda()
    t16 = plan.apply_transpose(wd, prec="bf16").cpu().numpy()
    t64 = plan.apply_transpose(wd, prec="fp64").cpu().numpy()
    print(f"  adjoint: BF16 rel err {rel(t16, t64):.2e}")
    assert 1e-5 < rel(t16, t64) <= 3e-2


def _separable_case(nd, lev, deg, active, conn, seed):
    """Random 1D operator in each term's active dimension, the identity elsewhere."""
    rng = np.random.default_rng(seed)
    if conn:
        h = O.hash_table_nd(nd, lev, "SG")
        rp, cols = O.connect_nd(nd, lev, "SG", lev, lev)
    else:
        h = O.hash_table_sparse_nd([lev] * nd, lev, "SG")
        rp = cols = None
    n = deg * 2 ** lev
    mats = [[rng.standard_normal((n, n)) if k == a else np.eye(n) for k in range(nd)] for a in active]
    x = rng.standard_normal(h["idx1d"].shape[0] * deg ** nd)
    return mats, h["idx1d"], rp, cols, x


@pytest.mark.parametrize("nd,lev,deg,conn,active", [
    (2, 4, 2, False, (0, 1)), (3, 4, 3, True, (0, 1, 2)), (3, 3, 4, False, (2, 2, 0)), (2, 6, 3, False, (1, 0, 0)),
])
def test_bf16_mfma_shapes_vs_oracle(kmx_gpu, nd, lev, deg, conn, active):
    """Separable shapes of several degrees (steps of four positions crossing level
    runs, packed column tiles and split tiles): BF16 within 3e-2 of the oracle."""
    import torch

    mats, idx, rp, cols, x = _separable_case(nd, lev, deg, active, conn, seed=20 + nd + lev)
    yo = O.apply_A(O.ApplyInputs(mats, idx, deg, rp, cols), x, nthreads=8)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
    assert plan.sweep("bf16")["available"]
    xd = torch.from_numpy(x).cuda()
    e16 = rel(plan.apply(xd, prec="bf16").cpu().numpy(), yo)
    e32 = rel(plan.apply(xd, prec="fp32").cpu().numpy(), yo)
    print(f"{nd}D lev {lev} deg {deg}: BF16 {e16:.2e}, FP32 {e32:.2e}")
    assert 1e-5 < e16 <= 3e-2, e16
    assert e32 <= 5e-6


def test_bf16_needs_the_separable_sweep(kmx_gpu, diffusion6):
    import torch

    plan = diffusion6[-1]  # built without KMX_PLAN_LEVEL_BLOCKS: no sweep
    assert not plan.sweep("bf16")["available"]
    with pytest.raises(kmx_gpu.KmxError):
        plan.apply(torch.zeros(plan.dof, dtype=torch.float64, device="cuda"), prec="bf16")
    rng = np.random.default_rng(2)
    mats, idx, rp, cols, _ = _separable_case(3, 3, 2, (0, 1), True, seed=3)
    mats[0][1] = rng.standard_normal(mats[0][1].shape)  # term 0 now acts in two dimensions
    general = kmx_gpu.Plan(mats, idx, 2, rp, cols, index_base=1, fp32=True, level_blocks=True)
    assert general.sweep_info()["available"] and not general.sweep_info()["separable"]
    with pytest.raises(kmx_gpu.KmxError):
        general.apply(torch.zeros(general.dof, dtype=torch.float64, device="cuda"), prec="bf16")
