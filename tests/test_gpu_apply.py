"""GPU parity of the HIP operator apply against the CPU oracle (apply_A.m's
explicit branch), through the C ABI.

Bars: plain FP64 within 1e-13 relative (2-norm) of the oracle; FP32 variants
within 5e-6 relative (FP32 unit roundoff 6e-8 times the accumulation depth).
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL64 = 1e-13
TOL32 = 5e-6


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make_case(nd, lev, deg, T, conn, seed, scale_mass=False):
    """Random FP64 operator inputs in the reference's layout (1-based indices)."""
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


def run(kmx, mats, idx, deg, rp, cols, x, prec, adjoint=False):
    import torch

    plan = kmx.Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, adjoint=adjoint)
    xd = torch.from_numpy(x).cuda()
    y = plan.apply(xd, prec=prec)
    torch.cuda.synchronize()
    return plan, y.cpu().numpy()


def oracle_y(mats, idx, deg, rp, cols, x, threads=8):
    inp = O.ApplyInputs(mats, idx, deg, rp, cols)
    return O.apply_A(inp, x, nthreads=threads)


# 2D continuity-shaped operator, SG lev 3, deg 2, two terms; all pairs
# (new hash, the reference default) and connect_nD (old hash)
@pytest.mark.parametrize("conn", [False, True])
def test_2d_lev3_shape_fp64(kmx_gpu, conn):
    mats, idx, rp, cols, x = make_case(2, 3, 2, 2, conn, seed=1)
    _, y = run(kmx_gpu, mats, idx, 2, rp, cols, x, "fp64")
    yo = oracle_y(mats, idx, 2, rp, cols, x)
    assert rel(y, yo) <= TOL64


@pytest.mark.parametrize(
    "nd,lev,deg,T,conn",
    [
        (1, 4, 2, 2, False), (1, 3, 5, 1, False), (1, 2, 8, 2, False),
        (2, 3, 1, 2, True), (2, 3, 3, 2, True), (2, 3, 4, 3, False), (2, 2, 6, 1, False),
        (3, 3, 2, 3, True), (3, 3, 3, 3, True), (3, 2, 4, 2, False),
        (4, 2, 2, 4, True), (4, 2, 3, 2, True),  # deg 3 in 4D: generic kernel
        (5, 2, 2, 2, True), (6, 2, 2, 6, True), (6, 1, 3, 1, False),
    ],
)
def test_shapes_fp64_fp32(kmx_gpu, nd, lev, deg, T, conn):
    mats, idx, rp, cols, x = make_case(nd, lev, deg, T, conn, seed=nd * 100 + lev * 10 + deg)
    yo = oracle_y(mats, idx, deg, rp, cols, x)
    plan, y64 = run(kmx_gpu, mats, idx, deg, rp, cols, x, "fp64")
    assert rel(y64, yo) <= TOL64, (nd, lev, deg)
    import torch

    xd = torch.from_numpy(x).cuda()
    y32 = plan.apply(xd, prec="fp32").cpu().numpy()
    assert rel(y32, yo) <= TOL32


def test_3d_lev7_full_size_fp64(kmx_gpu):
    """3D, lev 7, deg 3, 3 terms, connect_nD (P = 1,903,000): FP64 within 1e-13."""
    import torch

    mats, idx, rp, cols, x = make_case(3, 7, 3, 3, True, seed=2)
    assert len(cols) == 1903000
    plan = kmx_gpu.Plan(mats, idx, 3, rp, cols, fp32=True, adjoint=True)
    xd = torch.from_numpy(x).cuda()
    y64 = plan.apply(xd, prec="fp64").cpu().numpy()
    yo = oracle_y(mats, idx, 3, rp, cols, x, threads=16)
    r64 = rel(y64, yo)
    print(f"3D lev 7 fp64 rel err {r64:.3e}")
    assert r64 <= TOL64
    y32 = plan.apply(xd, prec="fp32").cpu().numpy()
    print(f"3D lev 7 fp32 rel err {rel(y32, yo):.3e}")
    assert rel(y32, yo) <= TOL32
    # adjoint identity <A x, w> = <x, A^T w> at full size
    w = np.random.default_rng(5).standard_normal(x.size)
    wd = torch.from_numpy(w).cuda()
    aTw = plan.apply_transpose(wd, prec="fp64").cpu().numpy()
    lhs, rhs = float(np.dot(y64, w)), float(np.dot(x, aTw))
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(y64) * np.linalg.norm(w)
    # linearity A(2x - 3w) = 2Ax - 3Aw
    z = plan.apply(2 * xd - 3 * wd, prec="fp64").cpu().numpy()
    yw = plan.apply(wd, prec="fp64").cpu().numpy()
    assert rel(z, 2 * y64 - 3 * yw) <= 1e-13


def test_3d_lev7_continuity3_operators(kmx_gpu):
    """The same size with continuity3's own 1D operators (upwind DIV, MASS;
    old hash + connect_nD, deg 3): the FP64 drop-in within 1e-13 of the oracle,
    FP32 within 5e-6, and the separable sweep (FP32 and FP64) likewise."""
    import torch
    from kmx import coeffs, grid

    nd, lev, deg = 3, 7, 3
    table = grid.hash_table_nd([lev] * nd, "SG")
    rp, cols = grid.connect_nd(table, lev, lev, "SG")
    mats, _ = coeffs.operator_mats("continuity3", nd, lev, deg)
    x = np.random.default_rng(11).standard_normal(table.num_elem * deg ** nd)
    plan = kmx_gpu.Plan(mats, table.idx1d, deg, rp, cols, fp32=True, level_blocks=True)
    xd = torch.from_numpy(x).cuda()
    yo = oracle_y(mats, table.idx1d, deg, rp, cols, x, threads=16)
    r64 = rel(plan.apply(xd, prec="fp64").cpu().numpy(), yo)
    r32 = rel(plan.apply(xd, prec="fp32").cpu().numpy(), yo)
    print(f"continuity3 lev 7: fp64 {r64:.2e}, fp32 {r32:.2e}")
    assert r64 <= TOL64 and r32 <= TOL32
    assert plan.sweep_info()["separable"]
    for prec, tol in (("fp32", TOL32), ("fp64", TOL64)):
        assert plan.sweep(prec)["available"], prec
        plan.set_sweep(prec, True)
        rs = rel(plan.apply(xd, prec=prec).cpu().numpy(), yo)
        print(f"  separable sweep {prec}: {rs:.2e}")
        assert rs <= tol
        plan.set_sweep(prec, False)


def test_autotune_keeps_results(kmx_gpu):
    """Every kernel variant the autotuner may pick computes the same sum."""
    import torch

    mats, idx, rp, cols, x = make_case(3, 4, 3, 3, True, seed=8)
    plan = kmx_gpu.Plan(mats, idx, 3, rp, cols)
    xd = torch.from_numpy(x).cuda()
    before = {p: plan.apply(xd, prec=p).cpu().numpy() for p in ("fp64", "fp32")}
    chosen = plan.autotune()
    assert chosen["fp64"]["threads"] in (128, 256, 512)
    yo = oracle_y(mats, idx, 3, rp, cols, x)
    for p, tol in (("fp64", TOL64), ("fp32", TOL32)):
        after = plan.apply(xd, prec=p).cpu().numpy()
        assert rel(after, yo) <= tol
        assert rel(after, before[p]) <= tol
    # a new plan of the same shape starts from the tuned choice
    plan2 = kmx_gpu.Plan(mats, idx, 3, rp, cols)
    assert plan2.variant("fp64") == {k: chosen["fp64"][k] for k in ("threads", "lds_staged", "lds_mode")}


@pytest.mark.parametrize("nd,lev,deg,conn", [(6, 2, 2, True), (5, 2, 2, True), (3, 2, 4, False)])
def test_lane_split_variants(kmx_gpu, nd, lev, deg, conn):
    """Blocks split over deg lanes (DPP exchange): every variant the autotuner
    can pick, both precisions, forward and transpose, against the oracle."""
    import torch

    mats, idx, rp, cols, x = make_case(nd, lev, deg, 3, conn, seed=60 + nd)
    yo = oracle_y(mats, idx, deg, rp, cols, x)
    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, adjoint=True)
    chosen = plan.autotune()
    xd = torch.from_numpy(x).cuda()
    assert rel(plan.apply(xd, prec="fp64").cpu().numpy(), yo) <= TOL64, chosen
    assert rel(plan.apply(xd, prec="fp32").cpu().numpy(), yo) <= TOL32, chosen
    A = O.build_A(O.ApplyInputs(mats, idx, deg, rp, cols)) if plan.dof <= 6000 else None
    if A is not None:
        assert rel(plan.apply_transpose(xd, prec="fp64").cpu().numpy(), A.T @ x) <= TOL64


@pytest.mark.parametrize("nd,lev,deg,conn", [(2, 3, 2, True), (3, 3, 3, True), (2, 3, 3, False), (4, 2, 3, True)])
def test_transpose_matches_assembled(kmx_gpu, nd, lev, deg, conn):
    mats, idx, rp, cols, x = make_case(nd, lev, deg, 2, conn, seed=7 + nd)
    A = O.build_A(O.ApplyInputs(mats, idx, deg, rp, cols))
    import torch

    plan = kmx_gpu.Plan(mats, idx, deg, rp, cols, adjoint=True)
    yT = plan.apply_transpose(torch.from_numpy(x).cuda(), prec="fp64").cpu().numpy()
    assert rel(yT, A.T @ x) <= TOL64
    yT32 = plan.apply_transpose(torch.from_numpy(x).cuda(), prec="fp32").cpu().numpy()
    assert rel(yT32, A.T @ x) <= TOL32


def test_edge_cases(kmx_gpu):
    import torch

    # a single element (lev 0): one 1x1-block pair per term
    rng = np.random.default_rng(0)
    mats = [[rng.standard_normal((3, 3)) for _ in range(2)]]
    idx = np.array([[1, 1]])
    x = rng.standard_normal(9)
    _, y = run(kmx_gpu, mats, idx, 3, None, None, x, "fp64")
    assert rel(y, oracle_y(mats, idx, 3, None, None, x)) <= TOL64
    # ragged pair list with empty rows: those rows must come out exactly 0
    mats, idx, _, _, x = make_case(2, 3, 2, 2, False, seed=3)
    N = idx.shape[0]
    rows = [sorted(rng.choice(N, size=int(rng.integers(0, 6)), replace=False) + 1) if i % 3 else [] for i in range(N)]
    rp = np.concatenate([[0], np.cumsum([len(r) for r in rows])]).astype(np.int64)
    cols = np.concatenate([np.asarray(r, dtype=np.int32) for r in rows if len(r)] or [np.zeros(0, np.int32)])
    _, y = run(kmx_gpu, mats, idx, 2, rp, cols, x, "fp64")
    yo = oracle_y(mats, idx, 2, rp, cols, x)
    assert rel(y, yo) <= TOL64
    for i in range(0, N, 3):
        assert np.all(y[i * 4:(i + 1) * 4] == 0.0)
    # 0-based indices give the same answer
    plan0 = kmx_gpu.Plan(mats, idx - 1, 2, rp, cols - 1, index_base=0)
    y0 = plan0.apply(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(y0, y)


def test_host_entry_and_update_term(kmx_gpu):
    mats, idx, rp, cols, x = make_case(3, 3, 3, 3, True, seed=11)
    plan = kmx_gpu.Plan(mats, idx, 3, rp, cols, adjoint=True)
    y = plan.apply_host(x)
    assert rel(y, oracle_y(mats, idx, 3, rp, cols, x)) <= TOL64
    # time-dependent coefficient refresh (get_coeff_mats TD=1): new matrix in term 1, dim 2
    new = np.random.default_rng(12).standard_normal(mats[1][2].shape)
    plan.update_term(1, 2, new)
    mats[1][2] = new
    assert rel(plan.apply_host(x), oracle_y(mats, idx, 3, rp, cols, x)) <= TOL64
    # per-term scale
    plan.set_term_scale(0, 0.5)
    ref = oracle_y([[0.5 * mats[0][0], mats[0][1], mats[0][2]], mats[1], mats[2]], idx, 3, rp, cols, x)
    assert rel(plan.apply_host(x), ref) <= TOL64


def test_errors(kmx_gpu):
    import torch

    from kmx._lib import KMX_E_ARG, KMX_E_SHAPE, KmxError

    mats, idx, rp, cols, x = make_case(2, 2, 2, 1, False, seed=4)
    plan = kmx_gpu.Plan(mats, idx, 2, fp32=False)
    xd = torch.from_numpy(x).cuda()
    with pytest.raises(KmxError) as e:
        plan.apply(xd, prec="fp32")  # no FP32 blocks
    assert e.value.code == KMX_E_ARG
    with pytest.raises(KmxError) as e:
        plan.apply_transpose(xd)  # no adjoint data
    assert e.value.code == KMX_E_ARG
    with pytest.raises(KmxError) as e:
        plan.apply(xd, y=xd)  # aliasing
    assert e.value.code == KMX_E_ARG
    bad = idx.copy()
    bad[0, 0] = 99
    with pytest.raises(KmxError) as e:
        kmx_gpu.Plan(mats, bad, 2)
    assert e.value.code == KMX_E_ARG
    with pytest.raises(KmxError) as e:
        kmx_gpu.Plan([[np.zeros((5, 5)), np.zeros((5, 5))]], idx, 2)  # 5 is not a multiple of deg
    assert e.value.code == KMX_E_SHAPE


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7, 8])
def test_kronmult_gpu(kmx_gpu, k):
    """kron_multd.m / kronmult{k}.m on the device, rectangular factors (Kronmult/test1.m);
    k = 7, 8 against the oracle's restatement of kron_multd.m's recursion."""
    import torch

    rng = np.random.default_rng(20 + k)
    shapes = [(2, 3), (3, 4), (4, 5), (5, 6), (6, 7), (3, 2), (2, 2), (3, 2)][:k]
    A = [rng.random(s) for s in shapes]
    X = rng.random((int(np.prod([s[1] for s in shapes])), 2))
    Y = kmx_gpu.kronmult([torch.from_numpy(a).cuda() for a in A], torch.from_numpy(X).cuda()).cpu().numpy()
    assert rel(Y, O.kron_multd(A, X)) <= 1e-14


def test_gemm_vbatched_and_batch_lists(kmx_gpu):
    import torch

    rng = np.random.default_rng(31)
    ms, ns, ks = [3, 5, 1, 4], [2, 4, 3, 4], [4, 2, 5, 3]
    As = [np.asfortranarray(rng.random((m, k))) for m, k in zip(ms, ks)]
    Bs = [np.asfortranarray(rng.random((n, k))) for n, k in zip(ns, ks)]  # transB = 'T'
    Cs = [np.asfortranarray(rng.random((m, n))) for m, n in zip(ms, ns)]
    dA = [torch.from_numpy(a.ravel(order="F")).cuda() for a in As]
    dB = [torch.from_numpy(b.ravel(order="F")).cuda() for b in Bs]
    dC = [torch.from_numpy(c.ravel(order="F").copy()).cuda() for c in Cs]
    kmx_gpu.gemm_vbatched("N", "T", ms, ns, ks, 1.5, dA, ms, dB, ns, -0.5, dC, ms)
    for a, b, c, d, m, n in zip(As, Bs, Cs, dC, ms, ns):
        got = d.cpu().numpy().reshape((m, n), order="F")
        assert np.allclose(got, 1.5 * a @ b.T - 0.5 * c, rtol=1e-14, atol=1e-14)
    # replay kronmult3_batch's stage lists stage by stage through gemm_vbatched
    shapes = [(2, 3), (3, 4), (4, 5)]
    mats = [rng.random(s) for s in shapes]
    nvec = 2
    X = rng.random((60, nvec))
    descs, work = kmx_gpu.kronmult_batch_plan([s[0] for s in shapes], [s[1] for s in shapes], nvec)
    bufs = {0: torch.from_numpy(X.ravel(order="F")).cuda(), 1: torch.zeros(24 * nvec, dtype=torch.float64).cuda(),
            2: torch.zeros(max(1, work), dtype=torch.float64).cuda(), 3: torch.zeros(1, dtype=torch.float64).cuda()}
    for i, m in enumerate(mats):
        bufs[4 + i] = torch.from_numpy(m.ravel(order="F")).cuda()
    for s in sorted({d["stage"] for d in descs}):
        st = [d for d in descs if d["stage"] == s]
        for ta in {d["transA"] for d in st}:
            for tb in {d["transB"] for d in st}:
                g = [d for d in st if d["transA"] == ta and d["transB"] == tb]
                ptr = lambda b, o: bufs[b].data_ptr() + 8 * o  # noqa: E731
                kmx_gpu.gemm_vbatched(ta, tb, [d["m"] for d in g], [d["n"] for d in g], [d["k"] for d in g], 1.0,
                                      [ptr(d["bufA"], d["offA"]) for d in g], [d["lda"] for d in g],
                                      [ptr(d["bufB"], d["offB"]) for d in g], [d["ldb"] for d in g], 0.0,
                                      [ptr(d["bufC"], d["offC"]) for d in g], [d["ldc"] for d in g])
    Y = bufs[1].cpu().numpy().reshape((24, nvec), order="F")
    assert rel(Y, np.kron(np.kron(mats[0], mats[1]), mats[2]) @ X) <= 1e-14


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
def test_gemm_vbatched_tiled_shapes(kmx_gpu, ta, tb):
    """gemm_vbatched.m's C = alpha op(A) op(B) + beta C on GEMMs that span several
    64 x 64 tiles of C and 16-wide chunks of k, sizes multiples of neither, leading
    dimensions above the minimum, k = 0 and beta = 0 (C not read: NaN there stays
    out), every transpose combination: within 1e-13 of numpy."""
    import torch

    rng = np.random.default_rng(77)
    shapes = [(1, 1, 1), (3, 5, 7), (64, 64, 16), (65, 130, 17), (200, 33, 150), (16, 80, 0), (129, 2, 40)]
    alpha, beta = 1.25, -0.75
    A, B, C, dA, dB, dC, lda, ldb, ldc, want = [], [], [], [], [], [], [], [], [], []
    for q, (m, n, k) in enumerate(shapes):
        ra, ca = (k, m) if ta == "T" else (m, k)
        rb, cb = (n, k) if tb == "T" else (k, n)
        la, lb, lc = max(1, ra) + 3, max(1, rb) + 1, m + 2
        a = rng.standard_normal((la, max(1, ca)))
        b = rng.standard_normal((lb, max(1, cb)))
        c = rng.standard_normal((lc, n))
        opa = (a[:ra, :ca].T if ta == "T" else a[:ra, :ca])
        opb = (b[:rb, :cb].T if tb == "T" else b[:rb, :cb])
        bq = 0.0 if q == 2 else beta
        w = c.copy()
        w[:m, :] = alpha * (opa @ opb) + (bq * c[:m, :] if bq else 0.0)
        if q == 2:
            c[:m, :] = np.nan  # beta = 0: C must not be read
        want.append((w, bq))
        for lst, arr in ((dA, a), (dB, b), (dC, c)):
            lst.append(torch.from_numpy(arr.ravel(order="F").copy()).cuda())
        lda.append(la)
        ldb.append(lb)
        ldc.append(lc)
    ms, ns, ks = [s_[0] for s_ in shapes], [s_[1] for s_ in shapes], [s_[2] for s_ in shapes]
    # beta differs for GEMM 2: run it on its own, the rest as one batch
    rest = [q for q in range(len(shapes)) if q != 2]
    for idx, bq in ((rest, beta), ([2], 0.0)):
        kmx_gpu.gemm_vbatched(ta, tb, [ms[q] for q in idx], [ns[q] for q in idx], [ks[q] for q in idx], alpha,
                              [dA[q] for q in idx], [lda[q] for q in idx], [dB[q] for q in idx],
                              [ldb[q] for q in idx], bq, [dC[q] for q in idx], [ldc[q] for q in idx])
    for q, (m, n, k) in enumerate(shapes):
        got = dC[q].cpu().numpy().reshape((ldc[q], n), order="F")
        w = want[q][0]
        assert np.all(np.isfinite(got[:m, :])), q
        assert np.max(np.abs(got[:m, :] - w[:m, :])) <= 1e-13 * max(1.0, np.max(np.abs(w[:m, :]))), q
        if q != 2:
            assert np.array_equal(got[m:, :], w[m:, :]), q  # rows past m untouched


def test_batch_matches_single_plans(kmx_gpu):
    """One persistent launch over instances of mixed (d, lev, deg) == their own applies."""
    import torch

    cases = [(2, 3, 2, 2, False), (3, 2, 3, 3, True), (2, 4, 4, 2, False), (3, 3, 2, 2, True), (2, 2, 3, 1, False),
             (4, 2, 2, 2, True)]  # the 4D one is served outside the persistent kernel
    plans, xs, ys, refs = [], [], [], []
    for c, (nd, lev, deg, T, conn) in enumerate(cases):
        mats, idx, rp, cols, x = make_case(nd, lev, deg, T, conn, seed=40 + c)
        p = kmx_gpu.Plan(mats, idx, deg, rp, cols)
        plans.append(p)
        xs.append(torch.from_numpy(x).cuda())
        ys.append(torch.empty_like(xs[-1]))
        refs.append(oracle_y(mats, idx, deg, rp, cols, x))
    b = kmx_gpu.Batch(plans)
    for prec, tol in (("fp64", TOL64), ("fp32", TOL32)):
        b.apply(xs, ys, prec=prec)
        torch.cuda.synchronize()
        for y, r in zip(ys, refs):
            assert rel(y.cpu().numpy(), r) <= tol, prec


def test_plain_c_caller(kmx_gpu):
    """dg-sparsegrid_amd/examples/kmx_demo.c: a C program (the shape of a MEX
    gateway) drives the ABI with host arrays and checks against its own
    direct sum."""
    import os
    import subprocess

    from conftest import PKG

    exe = os.path.join(PKG, "lib", "kmx_demo")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "kmx_demo ok" in r.stdout
