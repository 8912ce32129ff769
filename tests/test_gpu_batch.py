"""Many instances in one persistent launch (the 64-instance adapted-batch shape, reduced size):
batched apply == per-plan applies == oracle, graph-replayed RK3 == eager RK3,
and the FP32 batched trajectory stays within 1e-6 of the FP64 one."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def system(kmx_gpu):
    from kmx import batch as BT

    inst = BT.make_instances(count=12, seed=3, lev2d=(4, 6), lev3d=(3, 4))
    return BT.BatchedSystem(inst)


def test_batched_apply_matches_plans_and_oracle(system):
    import torch

    x = torch.randn(system.dof, dtype=torch.float64, device="cuda")
    for prec, tol in (("fp64", 1e-13), ("fp32", 5e-6)):
        y = torch.empty_like(x)
        system.apply(x, y, prec)
        torch.cuda.synchronize()
        for i, (p, inst) in enumerate(zip(system.plans, system.instances)):
            xi = x[system.offsets[i]:system.offsets[i + 1]].contiguous()
            yi = y[system.offsets[i]:system.offsets[i + 1]].cpu().numpy()
            assert rel(yi, p.apply(xi, prec=prec).cpu().numpy()) <= tol
            if i < 3 and prec == "fp64":
                mats = [[inst["scales"][t] * inst["mats"][t][0]] + list(inst["mats"][t][1:])
                        for t in range(len(inst["mats"]))]
                yo = O.apply_A(O.ApplyInputs(mats, inst["table"].idx1d, inst["deg"]), xi.cpu().numpy(), nthreads=8)
                assert rel(yi, yo) <= tol


def test_graph_rk3_matches_eager_and_fp32_error(kmx_gpu):
    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=12, seed=4, lev2d=(4, 6), lev3d=(3, 4))
    s64 = BT.BatchedSystem(inst)
    s32 = BT.BatchedSystem(inst)
    g64 = BT.BatchRK3Graph(s64, prec="fp64")
    g32 = BT.BatchRK3Graph(s32, prec="fp32")
    fe = s64.f.clone()
    for _ in range(10):
        g64.step()
        g32.step()
        fe = g64.eager_step(fe)
        torch.cuda.synchronize()
        assert rel(s64.f.cpu().numpy(), fe.cpu().numpy()) <= 1e-14
    err = rel(s32.f.cpu().numpy(), s64.f.cpu().numpy())
    print(f"fp32 batched RK3 vs fp64 after 10 steps: {err:.3e}")
    assert err <= 1e-6


def _lb_case(nd, lev, deg, T, conn, seed, scale=None):
    """Random 1D operators on a sparse grid (connect_nD pairs or all pairs)."""
    rng = np.random.default_rng(seed)
    if conn:
        h = O.hash_table_nd(nd, lev, "SG")
        rp, cols = O.connect_nd(nd, lev, "SG", lev, lev)
        idx = h["idx1d"]
    else:
        idx = O.hash_table_sparse_nd([lev] * nd, max(lev, 1), "SG")["idx1d"]
        rp = cols = None
    n = deg * 2 ** lev
    mats = [[rng.standard_normal((n, n)) for _ in range(nd)] for _ in range(T)]
    x = rng.standard_normal(idx.shape[0] * deg ** nd)
    return dict(mats=mats, idx=idx, deg=deg, rp=rp, cols=cols, x=x)


def _oracle(c, scales=None):
    mats = c["mats"]
    if scales is not None:
        mats = [[scales[t] * m[0]] + list(m[1:]) for t, m in enumerate(mats)]
    return O.apply_A(O.ApplyInputs(mats, c["idx"], c["deg"], c["rp"], c["cols"]), c["x"], nthreads=8)


def test_batched_sweep_mixed_dims_matches_oracle(kmx_gpu):
    """Merged dimension sweeps (one launch per phase for every plan that has a
    sweep path) next to plans on the pair path: 2D to 4D (and a 1D plan, which
    has no sweep), connect_nD and all pairs, an adapted grid with partially filled level blocks, a plan listed
    twice, one without level blocks and one whose terms each act in one
    dimension (the separable form: one op per dimension); FP64 1e-13 and FP32 5e-6 against the
    oracle, bit-identical on a second apply, and the pair path back after
    switching the sweep off."""
    import torch
    from kmx import grid as G

    kmx = kmx_gpu
    cases = [_lb_case(1, 4, 3, 2, False, 51), _lb_case(2, 4, 2, 2, True, 52), _lb_case(3, 3, 3, 3, True, 53),
             _lb_case(4, 3, 2, 4, True, 54), _lb_case(2, 3, 4, 2, False, 55)]
    rng = np.random.default_rng(3)  # a refinement whose level set stays downward closed
    table = G.refine_random(G.hash_table_sparse_nd([3] * 3, 5, "SG"), 4, 0.6, rng)
    n = 2 * 2 ** int(table.lev.max())
    cases.append(dict(mats=[[rng.standard_normal((n, n)) for _ in range(3)] for _ in range(2)], idx=table.idx1d,
                      deg=2, rp=None, cols=None, x=rng.standard_normal(table.num_elem * 8)))
    sep = _lb_case(3, 4, 3, 3, True, 57)  # separable: term t acts in dimension t only
    for t in range(3):
        for k in range(3):
            if k != t:
                sep["mats"][t][k] = np.eye(sep["mats"][t][k].shape[0])
    cases.append(sep)
    plans = [kmx.Plan(c["mats"], c["idx"], c["deg"], c["rp"], c["cols"], index_base=1, fp32=True,
                      level_blocks=True) for c in cases]
    # level blocks need d >= 2: the 1D plan stays on the pair path
    assert [p.sweep("fp64")["available"] for p in plans] == [False] + [True] * 6
    scales = [1.0, 0.5, 1.0]
    for t, s in enumerate(scales):
        plans[2].set_term_scale(t, s)
    plain = kmx.Plan(cases[1]["mats"], cases[1]["idx"], 2, cases[1]["rp"], cases[1]["cols"], index_base=1, fp32=True)
    members = plans + [plain, plans[2]]  # the repeated plan runs its sweep for its first slot only
    cases_all = cases + [cases[1], cases[2]]
    want = [_oracle(c, scales if c is cases[2] else None) for c in cases_all]
    batch = kmx.Batch(members)
    xs = [torch.from_numpy(c["x"]).cuda() for c in cases_all]
    for prec, tol in (("fp64", 1e-13), ("fp32", 5e-6)):
        batch.set_sweep(prec, True)
        info = batch.sweep(prec)
        assert info == {"members": 6, "active": True, "phases": 7}, info
        ys = [torch.full_like(x, np.nan) for x in xs]
        batch.apply(xs, ys, prec=prec)
        got = [y.cpu().numpy() for y in ys]
        for i, (g, w) in enumerate(zip(got, want)):
            assert rel(g, w) <= tol, (prec, i, rel(g, w))
        batch.apply(xs, ys, prec=prec)
        for g, y in zip(got, ys):
            assert np.array_equal(g, y.cpu().numpy())
        batch.set_sweep(prec, False)
        assert not batch.sweep(prec)["active"]
        batch.apply(xs, ys, prec=prec)
        for i, (y, w) in enumerate(zip(ys, want)):
            assert rel(y.cpu().numpy(), w) <= tol, (prec, "pairs", i)
    batch.destroy()
    with pytest.raises(RuntimeError):
        kmx.Batch([plain]).set_sweep("fp32", True)  # no plan with a sweep path


def test_graph_rk3_batched_sweep(kmx_gpu):
    """The RK3 step of adapted instances on merged sweeps, replayed from a HIP
    graph: every replay equals the eager step, the FP64 sweep trajectory tracks
    the FP64 pair trajectory, and the FP32 sweep stays within 1e-6 of it."""
    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=12, seed=5, lev2d=(4, 6), lev3d=(3, 4))
    ref = BT.BatchedSystem(inst)
    s64 = BT.BatchedSystem(inst, sweep=("fp64",))
    s32 = BT.BatchedSystem(inst, sweep=("fp32",))
    assert s64.sweep["fp64"]["members"] == 12 and s32.sweep["fp32"]["members"] == 12
    gref = BT.BatchRK3Graph(ref, prec="fp64")
    g64 = BT.BatchRK3Graph(s64, prec="fp64")
    g32 = BT.BatchRK3Graph(s32, prec="fp32")
    fe = s64.f.clone()
    for _ in range(8):
        gref.step()
        g64.step()
        g32.step()
        torch.cuda.synchronize()
        fe = g64.eager_step(fe)
        torch.cuda.synchronize()
        assert rel(s64.f.cpu().numpy(), fe.cpu().numpy()) <= 1e-14
    assert rel(s64.f.cpu().numpy(), ref.f.cpu().numpy()) <= 1e-12
    err = rel(s32.f.cpu().numpy(), ref.f.cpu().numpy())
    print(f"fp32 merged-sweep RK3 vs fp64 pairs after 8 steps: {err:.3e}")
    assert err <= 1e-6


def test_batch_autotune_keeps_results(kmx_gpu):
    """kmx_batch_autotune picks one of the timed load-batch counts for the merged
    sweep and the batched apply still matches the oracle per instance (FP32 bar)."""
    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=8, seed=4, lev2d=(4, 6), lev3d=(3, 4))
    sysm = BT.BatchedSystem(inst, sweep=("fp32",))
    x = torch.from_numpy(np.random.default_rng(6).standard_normal(sysm.dof)).cuda()
    y0 = sysm.apply(x, torch.empty_like(x), "fp32").cpu().numpy()
    wb = sysm.batch.autotune("fp32")
    assert wb in (4, 8, 16)
    y1 = sysm.apply(x, torch.empty_like(x), "fp32").cpu().numpy()
    assert rel(y1, y0) <= 1e-6
    xs = x.cpu().numpy()
    for i, (p, it) in enumerate(zip(sysm.plans, inst)):
        lo, hi = sysm.offsets[i], sysm.offsets[i + 1]
        mats = [[m * s if d == 0 else m for d, m in enumerate(ms)] for ms, s in zip(it["mats"], it["scales"])]
        yo = O.apply_A(O.ApplyInputs(mats, it["table"].idx1d, it["deg"]), xs[lo:hi], nthreads=8)
        assert rel(y1[lo:hi], yo) <= 5e-6, i
