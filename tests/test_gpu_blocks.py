"""GPU parity of the dimension-sweep applies built on the level-block structure
(KMX_PLAN_LEVEL_BLOCKS) against the oracle.

Elements sharing a level multi-index form dense cell tensors; when every
connected pair of such blocks is the tensor product of per-dimension 1D
connections (connect_nD.m:63-106 builds them that way, an all-pairs list
trivially), the plan records the structure and builds a dimension sweep that
sums exactly the terms apply_A.m:100-237 sums, in another order.  FP64 bar:
1e-13 relative (2-norm); FP32: 5e-6.  Pair lists without that structure keep
the pair-by-pair kernels.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 5e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def case(nd, lev, deg, T, conn, seed):
    rng = np.random.default_rng(seed)
    if conn:
        h = O.hash_table_nd(nd, lev, "SG")
        rp, cols = O.connect_nd(nd, lev, "SG", lev, lev)
    else:
        h = O.hash_table_sparse_nd([lev] * nd, max(lev, 1), "SG")
        rp = cols = None
    n = deg * 2 ** lev
    mats = [[rng.standard_normal((n, n)) for _ in range(nd)] for _ in range(T)]
    x = rng.standard_normal(h["idx1d"].shape[0] * deg ** nd)
    return mats, h["idx1d"], rp, cols, x


def oracle_y(mats, idx, deg, rp, cols, x):
    return O.apply_A(O.ApplyInputs(mats, idx, deg, rp, cols), x, nthreads=8)


def test_unstructured_pair_list_keeps_the_pair_kernels(kmx_gpu):
    """A pair list that is not a per-block tensor product (random subset of
    connect_nD): no sweep, and the pair-by-pair apply still matches."""
    import torch

    mats, idx, rp, cols, x = case(3, 4, 2, 2, True, seed=5)
    rng = np.random.default_rng(3)
    keep = rng.random(len(cols)) < 0.7
    nrp = np.zeros_like(rp)
    nrp[1:] = np.cumsum([keep[rp[i]:rp[i + 1]].sum() for i in range(len(rp) - 1)])
    ncols = cols[keep]
    plan = kmx_gpu.Plan(mats, idx, 2, nrp, ncols, index_base=1, fp32=True, level_blocks=True)
    assert not plan.sweep_info()["available"]
    with pytest.raises(kmx_gpu.KmxError):
        plan.set_sweep("fp32", True)
    yo = oracle_y(mats, idx, 2, nrp, ncols, x)
    xd = torch.from_numpy(x).cuda()
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        assert rel(plan.apply(xd, prec=prec).cpu().numpy(), yo) <= tol, prec
    plan.destroy()


def test_plans_are_pair_by_pair_by_default(kmx_gpu):
    import torch

    mats, idx, rp, cols, x = case(3, 4, 2, 2, True, seed=9)
    plan = kmx_gpu.Plan(mats, idx, 2, rp, cols, index_base=1, fp32=True)
    assert not plan.sweep_info()["available"]
    assert not plan.sweep("fp64")["active"]
    y = plan.apply(torch.from_numpy(x).cuda(), prec="fp64").cpu().numpy()
    assert rel(y, oracle_y(mats, idx, 2, rp, cols, x)) <= TOL64


def test_update_term_and_scale_reach_the_general_sweep(kmx_gpu):
    import torch

    mats, idx, rp, cols, x = case(3, 5, 2, 3, True, seed=13)
    plan = kmx_gpu.Plan(mats, idx, 2, rp, cols, index_base=1, fp32=True, level_blocks=True)
    assert plan.sweep_info()["available"] and not plan.sweep_info()["separable"]
    xd = torch.from_numpy(x).cuda()
    rng = np.random.default_rng(1)
    new = rng.standard_normal(mats[1][2].shape)
    plan.update_term(1, 2, new)
    plan.set_term_scale(0, -0.5)
    mats2 = [list(m) for m in mats]
    mats2[1][2] = new
    mats2[0] = [mats[0][0] * -0.5, mats[0][1], mats[0][2]]
    yo = oracle_y(mats2, idx, 2, rp, cols, x)
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        plan.set_sweep(prec, True)
        assert rel(plan.apply(xd, prec=prec).cpu().numpy(), yo) <= tol, prec
        plan.set_sweep(prec, False)
        assert rel(plan.apply(xd, prec=prec).cpu().numpy(), yo) <= tol, prec
    plan.destroy()


SWEEP_CASES = [
    (2, 3, 2, 2, False), (2, 3, 2, 2, True), (2, 5, 3, 3, True), (2, 4, 4, 2, False), (2, 5, 1, 2, True),
    (2, 4, 3, 3, True), (3, 4, 2, 3, True), (3, 3, 3, 3, False), (3, 5, 3, 3, True), (3, 4, 4, 2, True),
    (4, 3, 2, 4, True), (5, 2, 2, 2, True), (6, 3, 2, 6, True),
]


@pytest.mark.parametrize("nd,lev,deg,T,conn", SWEEP_CASES)
def test_dimension_sweep_matches_oracle(kmx_gpu, nd, lev, deg, T, conn):
    """The dimension-sweep path (each 1D operator split into its down and up
    parts by level, intermediates on the grid's own level tuples, matrix cores)
    against the oracle: FP64 1e-13, FP32 5e-6."""
    import torch

    mats, idx, rp, cols, x = case(nd, lev, deg, T, conn, seed=31 + nd * 10 + lev)
    yo = oracle_y(mats, idx, deg, rp, cols, x)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
    info = plan.sweep("fp64")
    assert info["available"], plan.sweep_info()
    xd = torch.from_numpy(x).cuda()
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        plan.set_sweep(prec, True)
        assert plan.sweep(prec)["active"]
        y = plan.apply(xd, prec=prec).cpu().numpy()
        assert rel(y, yo) <= tol, (prec, rel(y, yo))
        y2 = plan.apply(xd, prec=prec).cpu().numpy()  # a second apply reuses the boxes
        assert np.array_equal(y, y2)
        plan.set_sweep(prec, False)
    plan.destroy()


def test_dimension_sweep_3d_lev7_and_scales(kmx_gpu):
    import torch

    mats, idx, rp, cols, x = case(3, 7, 3, 3, True, seed=2)
    yo = oracle_y(mats, idx, 3, rp, cols, x)
    plan = kmx_gpu.Plan(mats, idx, 3, rp, cols, index_base=1, fp32=True, level_blocks=True)
    assert plan.sweep("fp64")["available"]
    xd = torch.from_numpy(x).cuda()
    plan.set_sweep("fp64", True)
    assert rel(plan.apply(xd, prec="fp64").cpu().numpy(), yo) <= TOL64
    plan.set_term_scale(1, 0.25)
    mats2 = [list(m) for m in mats]
    mats2[1] = [mats[1][0] * 0.25, mats[1][1], mats[1][2]]
    assert rel(plan.apply(xd, prec="fp64").cpu().numpy(), oracle_y(mats2, idx, 3, rp, cols, x)) <= TOL64


def test_dimension_sweep_adapted_grid(kmx_gpu):
    import torch
    from kmx import grid as G

    rng = np.random.default_rng(12)
    base = G.hash_table_sparse_nd([4, 4, 4], 6, "SG")
    table = G.refine_random(base, 6, 0.6, rng)
    lev = int(table.lev.max())
    n = 2 * 2 ** lev
    mats = [[rng.standard_normal((n, n)) for _ in range(3)] for _ in range(2)]
    x = rng.standard_normal(table.num_elem * 8)
    yo = oracle_y(mats, table.idx1d, 2, None, None, x)
    plan = kmx_gpu.Plan(mats, table.idx1d, 2, None, None, index_base=1, fp32=True, level_blocks=True)
    if not plan.sweep("fp64")["available"]:
        pytest.skip("refined level set is not downward closed")
    plan.set_sweep("fp64", True)
    assert rel(plan.apply(torch.from_numpy(x).cuda(), prec="fp64").cpu().numpy(), yo) <= TOL64


ADAPTED_SWEEP_CASES = [  # (dims, base level, refined max level, degree, terms, seed)
    (2, 3, 4, 2, 2, 1), (2, 3, 5, 3, 1, 2), (3, 3, 4, 2, 1, 3), (2, 4, 5, 2, 1, 4),
]


@pytest.mark.parametrize("nd,base_lev,ref_lev,deg,T,seed", ADAPTED_SWEEP_CASES)
def test_dimension_sweep_adapted_shapes(kmx_gpu, nd, base_lev, ref_lev, deg, T, seed):
    """Partially filled level blocks (absent cells read as zero through the box
    map) on the dimension-sweep path, FP64 and FP32, against the oracle."""
    import torch
    from kmx import grid as G

    rng = np.random.default_rng(seed)
    table = G.refine_random(G.hash_table_sparse_nd([base_lev] * nd, base_lev + 2, "SG"), ref_lev, 0.6, rng)
    n = deg * 2 ** int(table.lev.max())
    mats = [[rng.standard_normal((n, n)) for _ in range(nd)] for _ in range(T)]
    x = rng.standard_normal(table.num_elem * deg ** nd)
    yo = oracle_y(mats, table.idx1d, deg, None, None, x)
    plan = kmx_gpu.Plan(mats, table.idx1d, deg, None, None, index_base=1, fp32=True, level_blocks=True)
    assert plan.sweep("fp64")["available"]
    xd = torch.from_numpy(x).cuda()
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        plan.set_sweep(prec, True)
        assert rel(plan.apply(xd, prec=prec).cpu().numpy(), yo) <= tol, prec
    plan.destroy()


SEPARABLE_CASES = [  # (dims, level, degree, active dim of each term (-1: the identity term), connect_nD)
    (2, 4, 3, (0, 1, 1), True), (3, 4, 2, (0, 1, 2, 2, -1), False), (3, 5, 3, (2, 0, 1), True),
    (6, 3, 2, (0, 1, 2, 3, 4, 5), True),
]


def separable_case(nd, lev, deg, active, conn, seed):
    """Random 1D operator in each term's active dimension, the identity in the others."""
    mats, idx, rp, cols, x = case(nd, lev, deg, len(active), conn, seed)
    n = deg * 2 ** lev
    for t, a in enumerate(active):
        for k in range(nd):
            if k != a:
                mats[t][k] = np.eye(n)
    return mats, idx, rp, cols, x


@pytest.mark.parametrize("nd,lev,deg,active,conn", SEPARABLE_CASES)
def test_separable_sweep_matches_oracle(kmx_gpu, nd, lev, deg, active, conn):
    """Terms acting in one dimension each (identity factors elsewhere; one term
    that is the identity throughout): the sweep runs one whole-operator op per
    dimension, x to y; FP64 1e-13 and FP32 5e-6 against the oracle, fewer flops
    than the general sweep of the same grid, and term scales applied."""
    import torch

    mats, idx, rp, cols, x = separable_case(nd, lev, deg, active, conn, seed=71 + nd * 10 + lev)
    yo = oracle_y(mats, idx, deg, rp, cols, x)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
    info = plan.sweep("fp64")
    assert info["available"]
    gen_mats = case(nd, lev, deg, len(active), conn, seed=1)[0]
    general = kmx_gpu.Plan(gen_mats, idx, deg, rp, cols, index_base=1, fp32=False, level_blocks=True)
    assert info["flops"] < general.sweep("fp64")["flops"]
    general.destroy()
    xd = torch.from_numpy(x).cuda()
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        plan.set_sweep(prec, True)
        y = plan.apply(xd, prec=prec).cpu().numpy()
        assert rel(y, yo) <= tol, (prec, rel(y, yo))
    plan.set_term_scale(0, 0.5)
    mats2 = [list(m) for m in mats]
    mats2[0] = [0.5 * mats[0][0]] + list(mats[0][1:])
    assert rel(plan.apply(xd, prec="fp64").cpu().numpy(), oracle_y(mats2, idx, deg, rp, cols, x)) <= TOL64
    plan.destroy()


def test_separable_sweep_update_term(kmx_gpu):
    """kmx_plan_update_term on a separable sweep: a new matrix in a term's active
    dimension is used at once; replacing a factor the sweep takes as the identity
    makes the sweep refuse to apply (the pair path still applies the new operator)."""
    import torch

    nd, lev, deg = 3, 4, 2
    mats, idx, rp, cols, x = separable_case(nd, lev, deg, (0, 1, 2), True, seed=5)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
    xd = torch.from_numpy(x).cuda()
    plan.set_sweep("fp64", True)
    rng = np.random.default_rng(6)
    n = deg * 2 ** lev
    new = rng.standard_normal((n, n))
    plan.update_term(1, 1, new)
    mats[1][1] = new
    assert rel(plan.apply(xd, prec="fp64").cpu().numpy(), oracle_y(mats, idx, deg, rp, cols, x)) <= TOL64
    other = rng.standard_normal((n, n))
    plan.update_term(2, 0, other)  # term 2 acted in dimension 2 only
    mats[2][0] = other
    with pytest.raises(RuntimeError):
        plan.apply(xd, prec="fp64")
    plan.set_sweep("fp64", False)
    assert rel(plan.apply(xd, prec="fp64").cpu().numpy(), oracle_y(mats, idx, deg, rp, cols, x)) <= TOL64
    plan.destroy()


def test_separable_terms_on_the_general_sweep(kmx_gpu, monkeypatch):
    """KMX_SWEEP_SEPARABLE=0 at plan creation: separable terms run the general
    down/up sweep (identity factors as ordinary matrices); same oracle bar."""
    import torch

    mats, idx, rp, cols, x = separable_case(3, 4, 2, (0, 1, 2, -1), True, seed=9)
    yo = oracle_y(mats, idx, 2, rp, cols, x)
    monkeypatch.setenv("KMX_SWEEP_SEPARABLE", "0")
    plan = kmx_gpu.Plan(mats, idx, 2, rp, cols, index_base=1, fp32=True, level_blocks=True)
    monkeypatch.delenv("KMX_SWEEP_SEPARABLE")
    sep = kmx_gpu.Plan(mats, idx, 2, rp, cols, index_base=1, fp32=True, level_blocks=True)
    assert plan.sweep("fp64")["flops"] > sep.sweep("fp64")["flops"]  # the general form does more work
    xd = torch.from_numpy(x).cuda()
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        plan.set_sweep(prec, True)
        assert rel(plan.apply(xd, prec=prec).cpu().numpy(), yo) <= tol, prec
    plan.destroy()
    sep.destroy()


def test_separable_batch_refuses_stale_member(kmx_gpu):
    """A batch on merged sweeps refuses to apply once a member's identity factor
    was replaced, and applies again on the pair path after the sweep is off."""
    import torch

    mats, idx, rp, cols, x = separable_case(2, 4, 3, (0, 1), True, seed=10)
    plan = kmx_gpu.Plan(mats, idx, 3, rp, cols, index_base=1, fp32=True, level_blocks=True)
    batch = kmx_gpu.Batch([plan])
    batch.set_sweep("fp64", True)
    xd = torch.from_numpy(x).cuda()
    y = torch.empty_like(xd)
    batch.apply([xd], [y], prec="fp64")
    assert rel(y.cpu().numpy(), oracle_y(mats, idx, 3, rp, cols, x)) <= TOL64
    n = mats[0][1].shape[0]
    other = np.random.default_rng(11).standard_normal((n, n))
    plan.update_term(0, 1, other)  # term 0 acted in dimension 0 only
    mats[0][1] = other
    with pytest.raises(RuntimeError):
        batch.apply([xd], [y], prec="fp64")
    batch.set_sweep("fp64", False)
    batch.apply([xd], [y], prec="fp64")
    assert rel(y.cpu().numpy(), oracle_y(mats, idx, 3, rp, cols, x)) <= TOL64
    batch.destroy()
    plan.destroy()


@pytest.mark.parametrize("nd,lev,deg,active,conn", SEPARABLE_CASES)
def test_separable_sweep_transpose(kmx_gpu, nd, lev, deg, active, conn):
    """The adjoint apply on a separable sweep (transposed combined matrices): the
    pair lists are symmetric, so A^T w is the forward sum with every 1D matrix
    transposed; FP64 1e-13 / FP32 5e-6 against the oracle of that sum, and the
    pair-by-pair adjoint kernel agrees.  Without KMX_PLAN_ADJOINT the transpose
    runs only while the sweep is on."""
    import torch

    mats, idx, rp, cols, x = separable_case(nd, lev, deg, active, conn, seed=91 + nd * 10 + lev)
    mats_t = [[np.ascontiguousarray(m.T) for m in tm] for tm in mats]
    yo = oracle_y(mats_t, idx, deg, rp, cols, x)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, adjoint=True, level_blocks=True)
    xd = torch.from_numpy(x).cuda()
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        plan.set_sweep(prec, True)
        assert rel(plan.apply_transpose(xd, prec=prec).cpu().numpy(), yo) <= tol, prec
        plan.set_sweep(prec, False)
        assert rel(plan.apply_transpose(xd, prec=prec).cpu().numpy(), yo) <= tol, (prec, "pairs")
    plan.destroy()
    bare = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
    with pytest.raises(RuntimeError):
        bare.apply_transpose(xd, prec="fp64")
    bare.set_sweep("fp64", True)
    assert rel(bare.apply_transpose(xd, prec="fp64").cpu().numpy(), yo) <= TOL64
    bare.destroy()
