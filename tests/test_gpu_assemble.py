"""GPU parity of the assembled operator (kmx_plan_assemble: apply_A.m's build_A
branch, apply_A.m:150-155, 170-180, 246-248; krond.m:7-17) against the oracle's
dense build (orc_build_A), through the C ABI.

Bars: bit-exact against the oracle for unit term scales (both multiply each
entry's factors dimension by dimension and add the terms in order, without fused
multiply-adds); 1e-15 relative per entry with term scales (the scale multiplies
the product instead of the first factor); and the assembled matrix times x equals
the matrix-free FP64 apply within 1e-13 relative (2-norm).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def make_case(nd, lev, deg, T, conn, seed):
    rng = np.random.default_rng(seed)
    if conn:
        h = O.hash_table_nd(nd, lev, "SG")
        rp, cols = O.connect_nd(nd, lev, "SG", lev, lev)
    else:
        h = O.hash_table_sparse_nd([lev] * nd, max(lev, 1), "SG")
        rp = cols = None
    n = deg * 2 ** lev
    mats = [[rng.standard_normal((n, n)) for _ in range(nd)] for _ in range(T)]
    return mats, h["idx1d"], rp, cols


def dense_from_blocks(blocks, N, B, rp, cols):
    """Dense A from the per-pair blocks ([q, c, r] = entry (r, c) of pair q)."""
    if rp is None:
        e = np.repeat(np.arange(N), N)
        j = np.tile(np.arange(N), N)
    else:
        e = np.repeat(np.arange(N), np.diff(np.asarray(rp)))
        j = np.asarray(cols) - 1
    A = np.zeros((N * B, N * B))
    for q in range(len(e)):
        A[e[q] * B:(e[q] + 1) * B, j[q] * B:(j[q] + 1) * B] = blocks[q].T
    return A


@pytest.mark.parametrize("nd,lev,deg,T,conn", [(2, 3, 2, 2, True), (3, 2, 3, 3, True), (2, 3, 3, 1, False),
                                               (1, 4, 4, 2, False), (4, 2, 2, 2, True)])
def test_assembled_blocks_equal_oracle(kmx_gpu, nd, lev, deg, T, conn):
    import torch

    mats, idx, rp, cols = make_case(nd, lev, deg, T, conn, seed=31 + nd + lev)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols)
    blocks = plan.assemble().cpu().numpy()
    assert blocks.shape == (plan.P, plan.B, plan.B)
    A = dense_from_blocks(blocks, plan.N, plan.B, rp, cols)
    A_ref = O.build_A(O.ApplyInputs(mats, idx, deg, rp, cols))
    assert np.array_equal(A, A_ref)
    x = np.random.default_rng(5).standard_normal(plan.dof)
    y = plan.apply(torch.from_numpy(x).cuda(), prec="fp64").cpu().numpy()
    assert np.linalg.norm(A @ x - y) <= 1e-13 * np.linalg.norm(y)


def test_assembled_with_term_scales(kmx_gpu):
    mats, idx, rp, cols = make_case(2, 3, 3, 2, True, seed=3)
    plan = kmx_gpu.Plan(mats, idx, 3, rp, cols)
    plan.set_term_scale(0, 0.75)
    plan.set_term_scale(1, -2.5)
    A = dense_from_blocks(plan.assemble().cpu().numpy(), plan.N, plan.B, rp, cols)
    scaled = [[mats[0][0] * 0.75, mats[0][1]], [mats[1][0] * -2.5, mats[1][1]]]
    A_ref = O.build_A(O.ApplyInputs(scaled, idx, 3, rp, cols))
    assert np.max(np.abs(A - A_ref)) <= 1e-15 * np.max(np.abs(A_ref))


def test_asgard_build_A_sparse_and_dense(kmx_gpu):
    """kmx.asgard.build_A (the reference-named mirror) on continuity2's own
    operators: sparse (use_sparse_A) and dense agree, equal the oracle's build,
    and A f equals apply_A."""
    from kmx import asgard as AS
    from kmx import coeffs as K

    opts = AS.OPTS(lev=3, deg=2)
    pde, table, A_data, f0, _ = AS.build(K.continuity2, opts)
    As = AS.build_A(pde, opts, A_data, 2, sparse=True)
    Ad = AS.build_A(pde, opts, A_data, 2, sparse=False)
    assert np.array_equal(As.toarray(), Ad)
    assert np.array_equal(Ad, O.build_A(O.ApplyInputs(pde.mats, table.idx1d, 2)))
    f = np.random.default_rng(2).standard_normal(f0.size)
    y = AS.apply_A(pde, opts, A_data, f, 2)
    assert np.linalg.norm(Ad @ f - y) <= 1e-13 * np.linalg.norm(y)
