"""Pin the CPU oracle against the reference's own gold fixtures and against
closed forms (no GPU).  Fixture keys follow gold/generated-inputs/<dir>/<file>."""
import numpy as np
import pytest

import oracle as O
from kmx import coeffs as K


# ---------------------------------------------------------------- PERMUTE/
# gold/gold_permute.m:1-90 -- dims/ns/ords and the per-level cases it wrote
@pytest.mark.parametrize("dims,n,o", [(1, 1, 0), (2, 4, 1), (4, 6, 0), (6, 8, 1)])
def test_perm_leq_eq_max(gold, dims, n, o):
    for name, fn in (("perm_leq", O.perm_leq), ("perm_eq", O.perm_eq), ("perm_max", O.perm_max)):
        key = f"permutations/{name}_{dims}_{n}_{o}"
        r = fn(dims, n, o)
        if key in gold.files:
            assert np.array_equal(gold[key].astype(int), r), key
        ckey = key + "_count"
        if ckey in gold.files:
            assert int(gold[ckey][0, 0]) == r.shape[0], ckey


@pytest.mark.parametrize("i,levels", list(enumerate([[3, 3], [1, 4], [1, 5, 8], [10, 6, 9, 10], [2, 10, 1, 5, 4, 7]], 1)))
def test_perm_d(gold, i, levels):
    s, d = i % 2, len(levels)
    cases = [
        (f"permutations/perm_leq_d_{d}_{s}", O.perm_leq_d(d, levels, max(levels), s)),
        (f"permutations/perm_eq_d_{d}_{s}", O.perm_eq_d(d, levels, max(levels), s)),
        (f"permutations/perm_max_d_{d}_{s}", O.perm_max(d, levels, s)),
    ]
    seen = 0
    for key, r in cases:
        if key in gold.files:
            seen += 1
            assert np.array_equal(gold[key].astype(int), r), key
    assert seen >= 2


def test_index_leq_max(gold):
    lists = [np.arange(2, 4), np.arange(0, 5), np.arange(0, 4), np.arange(1, 6)]
    r = O.index_leq_max(4, lists, 10, 4)
    assert np.array_equal(gold["permutations/index_leq_max_4d_10s_4m"].astype(int), r)
    assert int(gold["permutations/index_leq_max_4d_10s_4m_count"][0, 0]) == r.shape[0]


# ------------------------------------------------------------ element table
@pytest.mark.parametrize("levels", [[7], [5, 2], [3, 2, 3]])
@pytest.mark.parametrize("grid", ["SG", "FG"])
def test_element_table(gold, levels, grid):
    # gold/gold_hash_table.m:12-64 (max_adapt_lev = 7)
    d = len(levels)
    h = O.hash_table_sparse_nd(levels, 7, grid)
    assert np.array_equal(gold[f"element_table/table_{d}d_{grid}"].astype(int), np.hstack([h["lev"], h["cell"]]))
    assert np.array_equal(gold[f"element_table/ids_{d}d_{grid}"].ravel().astype(np.uint64), h["ids"] - 1)


def test_lev_cell_to_1d(gold):
    keys = [k for k in gold.files if k.startswith("element_table/1d_index_")]
    assert len(keys) == 8
    for k in keys:
        lev, cell = map(int, k.split("_")[-2:])
        assert int(gold[k][0, 0]) == O.lev_cell_to_1d(lev, cell)


# ------------------------------------------------------------- connectivity
@pytest.mark.parametrize("lev", [2, 3, 8])
def test_connect_1d(gold, lev):
    assert np.array_equal(gold[f"connectivity/connect_1_{lev}"].astype(np.uint8), O.connect_1d(lev))


@pytest.mark.parametrize("nd,lev,grid,levsum,levmax,n", [(2, 3, "FG", 6, 3, 64), (3, 4, "SG", 4, 4, 104)])
def test_connect_nd(gold, nd, lev, grid, levsum, levmax, n):
    # gold/gold_connect.m:31-56 (connect_nD called with 5 args -> sparse-grid filter)
    rp, cols = O.connect_nd(nd, lev, grid, levsum, levmax)
    assert len(rp) - 1 == n
    for i in range(n):
        g = gold[f"connectivity/connect_n_{nd}_{lev}_{grid}_{i + 1}"].ravel().astype(int)
        assert np.array_equal(g, cols[rp[i]: rp[i + 1]]), i


def test_config_pair_counts():
    # the pair-list sizes the configurations are quoted on
    for nd, lev, P in [(2, 3, 392), (3, 5, 63664)]:
        rp, cols = O.connect_nd(nd, lev, "SG", lev, lev)
        assert len(cols) == P


# ---------------------------------------------------------------- kron
def _kron_all(A):
    K = A[0]
    for a in A[1:]:
        K = np.kron(K, a)
    return K


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5, 6, 7])
def test_kron_multd_closed_form(k):
    # Kronmult/test1.m:29-97 pattern (rectangular factors, nvec = 2); k = 7 uses
    # kron_multd's transpose recursion (kron_multd.m:43-86)
    rng = np.random.default_rng(k)
    shapes = [(2, 3), (3, 4), (4, 5), (5, 6), (6, 7), (3, 2), (2, 2)][:k]
    if k == 7:
        shapes = [(2, 3), (3, 2), (2, 2), (2, 3), (3, 2), (2, 2), (2, 3)]
    A = [rng.random(s) for s in shapes]
    X = rng.random((int(np.prod([s[1] for s in shapes])), 2))
    Y = O.kron_multd(A, X)
    Yr = _kron_all(A) @ X
    assert np.max(np.abs(Y - Yr)) <= 1e-13 * np.max(np.abs(Yr))


def test_kron_multd_method2_branch():
    # kronmult2 picks method 2 when it is cheaper (kronmult2.m:33-39)
    rng = np.random.default_rng(7)
    A = [rng.random((5, 2)), rng.random((1, 6))]
    X = rng.random((12, 3))
    assert np.allclose(O.kron_multd(A, X), np.kron(A[0], A[1]) @ X, rtol=1e-14, atol=1e-14)


def test_krond():
    rng = np.random.default_rng(3)
    A = [rng.random((2, 3)), rng.random((3, 2)), rng.random((2, 2))]
    assert np.array_equal(O.krond(A), _kron_all(A))


# ---------------------------------------------------------------- apply_A
def _rand_inputs(nd, lev, deg, T, conn, seed):
    rng = np.random.default_rng(seed)
    if conn:
        h = O.hash_table_nd(nd, lev, "SG")
        rp, cols = O.connect_nd(nd, lev, "SG", lev, lev)
    else:
        h = O.hash_table_sparse_nd([lev] * nd, 8, "SG")
        rp = cols = None
    n = deg * 2 ** lev
    mats = [[rng.standard_normal((n, n)) for _ in range(nd)] for _ in range(T)]
    inp = O.ApplyInputs(mats, h["idx1d"], deg, rp, cols)
    x = rng.standard_normal(inp.N * inp.B)
    return inp, x, mats, h


@pytest.mark.parametrize("nd,lev,deg,T,conn", [(2, 3, 2, 2, False), (2, 3, 2, 2, True), (3, 3, 3, 3, True),
                                               (1, 4, 4, 2, False), (4, 2, 2, 2, True)])
def test_apply_matrix_free_equals_assembled(nd, lev, deg, T, conn):
    """apply_A.m:185-201 (kron_multd) vs apply_A.m:170-180 (krond build_A)."""
    inp, x, _, _ = _rand_inputs(nd, lev, deg, T, conn, seed=nd * 10 + lev)
    y0 = O.apply_A(inp, x, mode=0)
    y1 = O.apply_A(inp, x, mode=1)
    assert np.array_equal(y0, y1)  # in-place chain == kron_multd recursion, bit for bit
    A = O.build_A(inp)
    yA = A @ x
    assert np.linalg.norm(y0 - yA) <= 1e-14 * np.linalg.norm(yA)
    y8 = O.apply_A(inp, x, mode=0, nthreads=4)
    assert np.array_equal(y0, y8)


@pytest.mark.parametrize("nd,lev,deg", [(2, 3, 2), (2, 2, 3), (3, 2, 2)])
def test_apply_full_grid_closed_form(nd, lev, deg):
    """All pairs: y = S (sum_t kron_d A_{t,d}) S^T x with S the element gather
    (the full-grid form of apply_A.m:74-98 / fast_2d_matrix_apply.m)."""
    inp, x, mats, h = _rand_inputs(nd, lev, deg, 2, False, seed=5 + nd)
    n = deg * 2 ** lev
    K = sum(_kron_all(tm) for tm in mats)
    B = deg ** nd
    rows = np.zeros(inp.N * B, dtype=np.int64)
    for e in range(inp.N):
        for a in range(B):
            digits = np.unravel_index(a, (deg,) * nd)
            r = 0
            for d in range(nd):
                r = r * n + (h["idx1d"][e, d] - 1) * deg + digits[d]
            rows[e * B + a] = r
    y_ref = K[np.ix_(rows, rows)] @ x
    y = O.apply_A(inp, x)
    assert np.linalg.norm(y - y_ref) <= 1e-13 * np.linalg.norm(y_ref)


def test_apply_separable_identity_case():
    """All mass factors = I: term t reduces to a 1D operator along dim t."""
    nd, lev, deg = 2, 3, 2
    h = O.hash_table_sparse_nd([lev] * nd, 8, "SG")
    n = deg * 2 ** lev
    rng = np.random.default_rng(11)
    D1 = rng.standard_normal((n, n))
    I = np.eye(n)
    inp = O.ApplyInputs([[D1, I]], h["idx1d"], deg)
    x = rng.standard_normal(inp.N * inp.B)
    y = O.apply_A(inp, x)
    # textbook: y_e[a1,a2] = sum_{j: I2(j)==I2(e)} sum_b1 D1[I1(e)deg+a1, J1(j)deg+b1] x_j[b1,a2]
    idx = h["idx1d"] - 1
    yr = np.zeros_like(y)
    for e in range(inp.N):
        for j in range(inp.N):
            if idx[j, 1] != idx[e, 1]:
                continue
            blk = D1[idx[e, 0] * deg:(idx[e, 0] + 1) * deg, idx[j, 0] * deg:(idx[j, 0] + 1) * deg]
            yr[e * 4:(e + 1) * 4] += (blk @ x[j * 4:(j + 1) * 4].reshape(deg, deg)).ravel()
    assert np.linalg.norm(y - yr) <= 1e-14 * np.linalg.norm(yr)


def test_planted_error_is_caught():
    """A wrong block index (I <-> J swapped in one dim) must fail the assembled check."""
    inp, x, mats, h = _rand_inputs(2, 3, 2, 1, False, seed=99)
    y = O.apply_A(inp, x)
    bad = O.ApplyInputs([[mats[0][0].T.copy(), mats[0][1]]], h["idx1d"], 2)
    yb = O.apply_A(bad, x)
    assert np.linalg.norm(y - yb) > 1e-3 * np.linalg.norm(y)


def test_gemm_vbatched_closed_form():
    rng = np.random.default_rng(2)
    ms, ns, ks = [3, 2, 4], [2, 5, 1], [4, 3, 2]
    As = [rng.random((k, m)) for m, k in zip(ms, ks)]  # transA = 'T'
    Bs = [rng.random((k, n)) for n, k in zip(ns, ks)]
    Cs = [np.asfortranarray(rng.random((m, n))) for m, n in zip(ms, ns)]
    C0 = [c.copy() for c in Cs]
    O.gemm_vbatched("T", "N", ms, ns, ks, 2.0, As, Bs, 0.5, Cs)
    for a, b, c, c0 in zip(As, Bs, Cs, C0):
        assert np.allclose(c, 2.0 * a.T @ b + 0.5 * c0, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("nd,lev,deg", [(1, 4, 3), (2, 3, 2), (3, 2, 2)])
def test_wavelet_to_realspace_closed_form(nd, lev, deg):
    """wavelet_to_realspace.m:44-79 against numpy: sum over elements of
    kron(column blocks of Meval_d) times the element's coefficients."""
    rng = np.random.default_rng(nd * 10 + lev)
    h = O.hash_table_sparse_nd([lev] * nd, lev, "SG")
    idx = h["idx1d"]
    n = deg * 2 ** lev
    npts = [5 + 2 * d for d in range(nd)]
    meval = [rng.standard_normal((npts[d], n)) for d in range(nd)]
    B = deg ** nd
    fw = rng.standard_normal(idx.shape[0] * B)
    fr = O.wavelet_to_realspace(meval, idx, deg, fw)
    ref = np.zeros(int(np.prod(npts)))
    for e in range(idx.shape[0]):
        K = np.ones((1, 1))
        for d in range(nd):
            I = idx[e, d] - 1
            K = np.kron(K, meval[d][:, I * deg:(I + 1) * deg])
        ref += K @ fw[e * B:(e + 1) * B]
    assert np.linalg.norm(fr - ref) <= 1e-13 * np.linalg.norm(ref)
    # a planted mistake (one element's block from the neighbouring 1D index) is caught
    bad = idx.copy()
    bad[-1, 0] = bad[-1, 0] % (2 ** lev) + 1
    assert np.linalg.norm(O.wavelet_to_realspace(meval, bad, deg, fw) - ref) > 1e-6 * np.linalg.norm(ref)


# gold_wavelet_to_realspace.m:1-3 (lev, deg), max_lev 8 (OPTS.m:24), sparse grid;
# domains pdes/continuity1.m:16, continuity2.m:27-29, continuity3.m:24-34
REALSPACE_GOLD = [("continuity_1", 7, 8, [(-1, 1)]),
                  ("continuity_2", 5, 4, [(-1, 1), (-2, 2)]),
                  ("continuity_3", 4, 3, [(-1, 1), (-2, 2), (-3, 3)])]


@pytest.mark.parametrize("name,deg,lev,doms", REALSPACE_GOLD)
def test_wavelet_to_realspace_gold(gold, name, deg, lev, doms):
    """wavelet_to_realspace.m end to end against the reference's gold vectors: the
    evaluation matrices of matrix_plot_D.m (restated in kmx.coeffs), the new-hash
    element table and f = 2 * (0, 1, ..., N deg^d - 1) (gold_wavelet_to_realspace.m:
    36-42).  The 3D vector is stored as every 8th value."""
    key = f"transformations/wavelet_to_realspace_{name}"
    stride = int(gold[key + "/stride"][0])
    ref = gold[key].ravel()
    idx = O.hash_table_sparse_nd([lev] * len(doms), 8, "SG")["idx1d"]
    meval = [K.matrix_plot_D(deg, lev, a, b)[0] for (a, b) in doms]
    fw = 2.0 * np.arange(idx.shape[0] * deg ** len(doms))
    fr = O.wavelet_to_realspace(meval, idx, deg, fw)
    assert fr.size == (deg * 2 ** lev) ** len(doms)
    assert np.linalg.norm(fr[::stride] - ref) <= 1e-13 * np.linalg.norm(ref)
