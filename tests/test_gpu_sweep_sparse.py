"""GPU parity of the block-sparse separable dimension sweep.

The separable sweep sums, per dimension, the terms acting there (their other
factors the identity) as one combined operator, keeping only the deg x deg
blocks above 1e-15 of each term's largest entry.  Checked here against the
oracle (the reference's pair-by-pair apply_A.m:100-237 restated in C++):

* the 64-instance adapted advection-diffusion batch (levels 4-8 in 2D and 3D,
  deg 2-4, all element pairs), every instance of one merged-sweep batched apply:
  FP64 <= 1e-13, FP32 <= 5e-6, BF16 operands with FP32 accumulation <= 3e-2
  (2-norm relative);
* the reference's own gold coefficient matrices, whose MASS factors are the
  identity only up to 2.7e-14 .. 2.3e-13 (M\\M, pterm_coeff_matrix.m:20): the
  FP32 sweep takes them as the identity and stays within 5e-6 of the oracle fed
  the gold matrices; FP64 keeps the pair path (<= 1e-13) unless the factors are
  declared the identity, and then matches the oracle fed exact identities;
* term scales changed after a batch's sweep is switched on reach every member
  (separable, general sweep, pair path) at the next batched apply;
* a pair list whose identity dimensions lack self-connections is not taken as
  separable (the general sweep runs and matches the oracle).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 5e-6
TOL16 = 3e-2  # BF16 operands (unit roundoff 3.9e-3), FP32 accumulation
CHUNKS = 8


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def scaled_mats(inst):
    """The instance's terms with its scale folded into the first factor (what the batch applies)."""
    return [[inst["scales"][t] * inst["mats"][t][0]] + list(inst["mats"][t][1:]) for t in range(len(inst["mats"]))]


@pytest.fixture(scope="module")
def batch64(kmx_gpu):
    """One merged-sweep batched apply, FP64 and FP32, of the 64 adapted instances on x ~ N(0,1)."""
    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=64, seed=0)
    sysm = BT.BatchedSystem(inst, sweep=("fp64", "fp32"))
    for prec in ("fp64", "fp32"):
        sw = sysm.sweep[prec]
        print(f"batch64: {prec} merged sweep members {sw['members']} of {len(inst)}", flush=True)
        assert sw["active"] and sw["members"] == len(inst), sw
    assert sysm.sweep["bf16"]["active"] and sysm.sweep["bf16"]["members"] == len(inst)
    x = torch.from_numpy(np.random.default_rng(11).standard_normal(sysm.dof)).cuda()
    ys = {}
    for prec in ("fp64", "fp32", "bf16"):
        y = torch.empty_like(x)
        sysm.apply(x, y, prec)
        torch.cuda.synchronize()
        ys[prec] = y.cpu().numpy()
    info = [p.sweep_info() for p in sysm.plans]
    return inst, sysm.offsets, x.cpu().numpy(), ys, info


@pytest.mark.parametrize("chunk", range(CHUNKS))
def test_batch64_each_instance_matches_oracle(batch64, chunk):
    inst, off, x, ys, info = batch64
    per = (len(inst) + CHUNKS - 1) // CHUNKS
    for i in range(chunk * per, min(len(inst), (chunk + 1) * per)):
        it = inst[i]
        assert info[i]["separable"] and info[i]["fp64"], info[i]
        xi = x[off[i]:off[i + 1]]
        yo = O.apply_A(O.ApplyInputs(scaled_mats(it), it["table"].idx1d, it["deg"]), xi, nthreads=O.num_threads())
        e64 = rel(ys["fp64"][off[i]:off[i + 1]], yo)
        e32 = rel(ys["fp32"][off[i]:off[i + 1]], yo)
        e16 = rel(ys["bf16"][off[i]:off[i + 1]], yo)
        print(f"instance {i}: {it['nd']}D lev {it['lev']} deg {it['deg']} N {it['table'].num_elem} "
              f"kept blocks {info[i]['kept_blocks']}/{info[i]['total_blocks']}: fp64 {e64:.2e} fp32 {e32:.2e} "
              f"bf16 {e16:.2e}", flush=True)
        assert e64 <= TOL64, (i, e64)
        assert e32 <= TOL32, (i, e32)
        assert 1e-5 < e16 <= TOL16, (i, e16)


GOLD_SETS = [  # (name, levels, deg, terms): the gold coefficient files of the reference's PDEs
    ("continuity2_coefficients_l4_4_d3", [4, 4], 3, 2),
    ("continuity2_coefficients_l4_5_d3", [4, 5], 3, 2),
    ("diffusion2_coefficients_l3_3_d5", [3, 3], 5, 2),
    ("diffusion2_coefficients_l2_3_d5", [2, 3], 5, 2),
    ("continuity3_coefficients_l2_3_2_d4", [2, 3, 2], 4, 3),
]


@pytest.mark.parametrize("name,levs,deg,T", GOLD_SETS)
def test_gold_near_identity_factors(kmx_gpu, gold, name, levs, deg, T):
    import torch

    nd = len(levs)
    mats = [[gold[f"coefficients/{name}_{t + 1}_{d + 1}"] for d in range(nd)] for t in range(T)]
    for d in range(nd):
        assert mats[0][d].shape == (deg * 2 ** levs[d],) * 2
    dev = [[float(np.abs(m - np.eye(m.shape[0])).max()) for m in row] for row in mats]
    near = [[v < 1e-11 for v in row] for row in dev]
    assert all(sum(not v for v in row) == 1 for row in near), dev  # one active factor per term
    h = O.hash_table_sparse_nd(levs, max(levs), "SG")
    idx = h["idx1d"]
    x = np.random.default_rng(5).standard_normal(idx.shape[0] * deg ** nd)
    xd = torch.from_numpy(x).cuda()
    yo = O.apply_A(O.ApplyInputs(mats, idx, deg), x, nthreads=O.num_threads())

    plan = kmx_gpu.Plan(mats, idx, deg, None, None, index_base=1, fp32=True, level_blocks=True)
    info = plan.sweep_info()
    worst = max(v for row, nr in zip(dev, near) for v, n in zip(row, nr) if n)
    print(f"{name}: identity factors deviate by up to {worst:.2e}; sweep {info}", flush=True)
    assert info["separable"] and not info["fp64"]
    assert info["identity_dev"] == pytest.approx(worst, rel=1e-12)
    assert plan.sweep("fp32")["available"] and not plan.sweep("fp64")["available"]
    with pytest.raises(kmx_gpu.KmxError):
        plan.set_sweep("fp64", True)
    e64 = rel(plan.apply(xd, prec="fp64").cpu().numpy(), yo)  # the pair path: the matrices as given
    plan.set_sweep("fp32", True)
    e32 = rel(plan.apply(xd, prec="fp32").cpu().numpy(), yo)
    print(f"{name}: fp64 pair path {e64:.2e}, fp32 separable sweep {e32:.2e} vs the oracle", flush=True)
    assert e64 <= TOL64 and e32 <= TOL32
    plan.destroy()

    # declared identities (MASS pterms with g = 1): the FP64 sweep takes them as exactly I
    declared = kmx_gpu.Plan(mats, idx, deg, None, None, index_base=1, fp32=True, level_blocks=True, identity=near)
    info = declared.sweep_info()
    assert info["separable"] and info["fp64"] and info["identity_dev"] == 0.0
    declared.set_sweep("fp64", True)
    y = declared.apply(xd, prec="fp64").cpu().numpy()
    exact = [[np.eye(m.shape[0]) if n else m for m, n in zip(row, nr)] for row, nr in zip(mats, near)]
    ye = O.apply_A(O.ApplyInputs(exact, idx, deg), x, nthreads=O.num_threads())
    print(f"{name}: declared-identity fp64 sweep {rel(y, ye):.2e} vs the oracle with exact identities, "
          f"{rel(y, yo):.2e} vs the oracle with the gold factors", flush=True)
    assert rel(y, ye) <= TOL64
    assert rel(y, yo) <= 1e-12
    declared.destroy()
    with pytest.raises(kmx_gpu.KmxError):  # a declared factor must be the identity
        bad = [[1.0 - 0.5 * n for n in row] for row in near]
        kmx_gpu.Plan([[m * b if n else m for m, b, n in zip(row, br, nr)] for row, br, nr in zip(mats, bad, near)],
                     idx, deg, None, None, index_base=1, level_blocks=True, identity=near)


def test_scale_change_reaches_every_batch_member(kmx_gpu):
    import torch

    rng = np.random.default_rng(21)
    plans, inputs = [], []
    # a separable member (one active factor per term), a general-sweep member (every
    # factor random) and a pair-path member (no level blocks)
    for kind, nd, lev, deg in (("sep", 2, 5, 3), ("general", 2, 4, 2), ("pairs", 3, 3, 2)):
        h = O.hash_table_sparse_nd([lev] * nd, lev, "SG")
        n = deg * 2 ** lev
        if kind == "sep":
            mats = [[rng.standard_normal((n, n)) if d == t else np.eye(n) for d in range(nd)] for t in range(nd)]
        else:
            mats = [[rng.standard_normal((n, n)) for _ in range(nd)] for _ in range(2)]
        p = kmx_gpu.Plan(mats, h["idx1d"], deg, None, None, index_base=1, fp32=True, level_blocks=kind != "pairs")
        if kind != "pairs":
            assert p.sweep_info()["separable"] == (kind == "sep")
        plans.append(p)
        inputs.append((mats, h["idx1d"], deg))
    b = kmx_gpu.Batch(plans)
    b.set_sweep("fp64", True)
    assert b.sweep("fp64")["members"] == 2
    xs = [torch.from_numpy(rng.standard_normal(p.dof)).cuda() for p in plans]
    ys = [torch.empty_like(x) for x in xs]
    for step in range(2):
        scales = [rng.uniform(0.5, 2.0, size=len(mats)) for mats, _, _ in inputs]
        for p, sc in zip(plans, scales):
            for t, s in enumerate(sc):
                p.set_term_scale(t, float(s))
        b.apply(xs, ys, prec="fp64")
        torch.cuda.synchronize()
        for (mats, idx, deg), sc, x, y in zip(inputs, scales, xs, ys):
            sm = [[sc[t] * mats[t][0]] + list(mats[t][1:]) for t in range(len(mats))]
            yo = O.apply_A(O.ApplyInputs(sm, idx, deg), x.cpu().numpy(), nthreads=O.num_threads())
            assert rel(y.cpu().numpy(), yo) <= TOL64, step
    b.destroy()
    for p in plans:
        p.destroy()


def test_identity_dimension_without_self_connections_is_not_separable(kmx_gpu):
    import torch

    nd, lev, deg = 2, 3, 2
    h = O.hash_table_sparse_nd([lev] * nd, lev, "SG")
    idx = h["idx1d"]
    N = idx.shape[0]
    # pairs that differ in the dimension-1 index: a tensor product of the 1D masks
    # (dim 1: all but the diagonal, dim 2: all), without self-connections in dim 1
    rows = [[j for j in range(N) if idx[i, 0] != idx[j, 0]] for i in range(N)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    cols = np.asarray([j + 1 for r in rows for j in r], dtype=np.int32)
    n = deg * 2 ** lev
    rng = np.random.default_rng(3)
    mats = [[rng.standard_normal((n, n)), np.eye(n)], [np.eye(n), rng.standard_normal((n, n))]]
    x = rng.standard_normal(N * deg ** nd)
    yo = O.apply_A(O.ApplyInputs(mats, idx, deg, rp, cols), x, nthreads=O.num_threads())
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, level_blocks=True)
    info = plan.sweep_info()
    print(f"no self-connections in dim 1: sweep {info}", flush=True)
    assert not info["separable"]
    if info["available"]:
        plan.set_sweep("fp64", True)
    y = plan.apply(torch.from_numpy(x).cuda(), prec="fp64").cpu().numpy()
    assert rel(y, yo) <= TOL64
    plan.destroy()


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16"])
def test_fused_rk3_matches_applies(kmx_gpu, prec):
    """kmx_batch_rk3 (stage arithmetic fused into the merged sweeps, graph-replayed)
    against the same RK3 steps built from batched applies and FP64 vector updates
    (time_advance.m:151-204): FP64 within 1e-13 of the FP64 pair path, FP32
    within 1e-6 of it, BF16 operands within 1e-2 (reported; BF16 rounds each
    apply's operands to 3.9e-3); every replay checked against the eager steps."""
    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=16, seed=8, lev2d=(4, 7), lev3d=(4, 6))
    ref = BT.BatchedSystem(inst)
    fused = BT.BatchedSystem(inst, sweep=(prec,))
    assert fused.sweep[prec]["members"] == len(inst)
    g = BT.BatchRK3Fused(fused, prec=prec, steps_per_replay=2)
    gref = BT.BatchRK3Graph(ref, prec="fp64")
    eager = fused.f.clone()
    tol = {"fp64": 1e-13, "fp32": 1e-6, "bf16": 1e-2}[prec]
    for r in range(4):
        g.step()
        fused.rk3(2, prec, f=eager)  # the same fused steps, eager
        for _ in range(2):
            gref.step()
        torch.cuda.synchronize()
        assert torch.equal(fused.f, eager), r
        e = rel(fused.f.cpu().numpy(), ref.f.cpu().numpy())
        print(f"fused rk3 {prec} after {2 * (r + 1)} steps: {e:.2e} vs the fp64 pair path", flush=True)
        assert e <= tol, (r, e)
    # the instances as two batches on streams of their own: the same states
    g2 = BT.BatchRK3Fused(fused, prec=prec, steps_per_replay=2, split=2)
    assert len(g2.batches) == 2 and sorted(i for grp in g2.groups for i in grp) == list(range(len(inst)))
    f_before = fused.f.clone()
    g.step()
    one = fused.f.clone()
    fused.f.copy_(f_before)
    g2.step()
    torch.cuda.synchronize()
    assert torch.equal(fused.f, one)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_launch_layout_switches_keep_results(kmx_gpu, prec):
    """The merged sweep's work list dealt to the XCDs in list order
    (KMX_SEP_XCD_BLOCK=1) or in runs (the default), and the fused RK3 stage
    launches one box value per thread (KMX_RK3_VEC=0) or four (the default), give
    bit-identical applies and RK3 states: they change where and how values move,
    not what is summed or in which order."""
    import os

    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=8, seed=21, lev2d=(4, 6), lev3d=(4, 5))

    def run():
        sysm = BT.BatchedSystem(inst, sweep=(prec,))
        x = torch.from_numpy(np.random.default_rng(4).standard_normal(sysm.dof)).cuda()
        y = torch.empty_like(x)
        sysm.apply(x, y, prec)
        f = sysm.f.clone()
        sysm.rk3(3, prec, f=f)
        torch.cuda.synchronize()
        return y, f

    saved = {k: os.environ.get(k) for k in ("KMX_SEP_XCD_BLOCK", "KMX_RK3_VEC")}
    try:
        for k in saved:
            os.environ.pop(k, None)
        y0, f0 = run()
        os.environ["KMX_SEP_XCD_BLOCK"] = "1"
        y1, f1 = run()
        os.environ.pop("KMX_SEP_XCD_BLOCK")
        os.environ["KMX_RK3_VEC"] = "0"
        y2, f2 = run()
    finally:
        for k, v in saved.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    assert torch.equal(y0, y1) and torch.equal(f0, f1)
    assert torch.equal(y0, y2) and torch.equal(f0, f2)
    assert float(torch.linalg.vector_norm(f0)) > 0.0
