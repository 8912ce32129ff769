"""CPU pins of the operator's input provider and callers against the
reference's gold fixtures: FMWT, quadrature, DG coefficient matrices, the
separable-source projection, and whole RK3 / backward-Euler trajectories
computed with the oracle apply (apply_A.m + time_advance.m end to end).

The gold trajectories (gold/generated-inputs/time_advance/*, gold_time_advance.m)
were written by a reference revision whose continuity PDEs used the central
flux (LF = 0); the current pdes/continuity2.m passes LF = -1.  The runs below
therefore set LF = 0; the LF != 0 flux term is pinned by its closed form on
periodic meshes (test_div_flux_term_closed_form) and exercised by the GPU parity
tests through the same builder."""
import numpy as np
import pytest

import oracle as O
from kmx import coeffs as K
from kmx import grid as G


def central_flux(pde):
    for tm in pde.terms:
        for chain in tm:
            for pt in chain:
                if pt.type in ("div", "grad"):
                    pt.LF = 0.0
    return pde


@pytest.mark.parametrize("key", ["2_2", "2_3", "2_6", "4_3", "5_5"])
def test_fmwt(gold, key):
    deg, lev = map(int, key.split("_"))
    F = K.fmwt(deg, lev)
    assert np.max(np.abs(F - gold[f"transformations/operator_two_scale_{key}"])) <= 1e-14
    assert np.max(np.abs(F @ F.T - np.eye(F.shape[0]))) <= 1e-12  # orthogonal (the deg-5 filters reach 2.3e-13)


# gold_matrix_plot_D.m:1-3 (lev, deg); domains pdes/continuity1.m:16, continuity2.m:27-29,
# continuity3.m:24-34
MEVAL_GOLD = [("continuity_1", 7, 7, [(-1, 1)]),
              ("continuity_2", 6, 7, [(-1, 1), (-2, 2)]),
              ("continuity_3", 5, 6, [(-1, 1), (-2, 2), (-3, 3)])]


@pytest.mark.parametrize("name,deg,lev,doms", MEVAL_GOLD)
def test_matrix_plot_D_matches_gold(gold, name, deg, lev, doms):
    """matrix_plot_D.m (default 'quadrature' grid) against the reference's evaluation
    matrices, every 8th row stored; a planted mistake (the first dimension's cell
    width in every dimension) is caught where the domains differ."""
    for d, (a, b) in enumerate(doms):
        key = f"transformations/matrix_plot_D/{name}/matrix_plot_D_{d}"
        stride = int(gold[key + "/stride"][0])
        ref = gold[key]
        M, nodes = K.matrix_plot_D(deg, lev, a, b)
        assert M.shape == (deg * 2 ** lev, deg * 2 ** lev)
        scale = np.max(np.abs(ref))
        assert np.max(np.abs(M[::stride] - ref)) <= 1e-12 * scale  # the deg-6 filters reach 5.6e-13
        assert nodes.min() > a and nodes.max() < b
        if (a, b) != doms[0]:
            bad, _ = K.matrix_plot_D(deg, lev, *doms[0])
            assert np.max(np.abs(bad[::stride] - ref)) > 1e-3 * scale


def test_asgard_matrix_plot_D_mirror():
    """The reference-named entry (matrix_plot_D(pde, opts, dimension)) reads deg
    from opts and lev / domain from the dimension; output grids it does not
    restate raise instead of returning another grid's matrix."""
    from kmx import asgard as A

    opts = A.OPTS(lev=3, deg=3)
    dim = K.Dimension(-2.0, 2.0, 4)
    M, nodes = A.matrix_plot_D(None, opts, dim)
    M2, nodes2 = K.matrix_plot_D(3, 4, -2.0, 2.0)
    assert np.array_equal(M, M2) and np.array_equal(nodes, nodes2)
    opts.output_grid = "uniform"
    with pytest.raises(NotImplementedError):
        A.matrix_plot_D(None, opts, dim)


@pytest.mark.parametrize("deg,lev,a,b", [(2, 3, -1.0, 1.0), (4, 4, -2.0, 5.0), (7, 2, 0.0, 3.0)])
def test_matrix_plot_D_gauss_orthonormality(deg, lev, a, b):
    """The wavelet basis is orthonormal on [a, b], and deg Gauss points per cell
    integrate products of degree < deg polynomials exactly, so with the Gauss
    weights (cell width / 2 times the reference-cell weights) Meval^T W Meval = I."""
    M, _ = K.matrix_plot_D(deg, lev, a, b)
    _, w = K.lgwt(deg, -1.0, 1.0)
    h = (b - a) / 2 ** lev
    W = np.tile(w * h / 2, 2 ** lev)
    assert np.max(np.abs(M.T @ (W[:, None] * M) - np.eye(M.shape[1]))) <= 1e-11  # deg-7 filters: 3.8e-12


def test_lgwt(gold):
    for n, a, b, tag in [(10, -1, 1, "10_neg1_1"), (32, -5, 2, "32_neg5_2")]:
        x, w = K.lgwt(n, a, b)
        assert np.max(np.abs(x - gold[f"quadrature/lgwt_roots_{tag}"].ravel())) <= 1e-14
        assert np.max(np.abs(w - gold[f"quadrature/lgwt_weights_{tag}"].ravel())) <= 1e-14


def test_mass_factors_match_gold(gold):
    """continuity2 lev [4,4] deg 3: the MASS factors (identity + roundoff)."""
    pde = K.continuity2(4)
    for t, d in [(0, 1), (1, 0)]:
        g = gold[f"coefficients/continuity2_coefficients_l4_4_d3_{t + 1}_{d + 1}"]
        m = K.sd_term_coeff_matrix(3, 1.0, pde.dims[d], pde.terms[t][d], pde.params, 4)
        assert np.max(np.abs(m - g)) <= 1e-13


def test_operator_identities():
    """grad = -div^T (coeff_matrix.m:323) and level consistency of the
    hierarchical basis (a matrix built at max_lev restricted to lev equals
    the one built at lev)."""
    dim = K.Dimension(-1, 1, 3)
    pt = K.DIV(K._const(-1.0), -1.0, "P", "P")
    _, div = K.coeff_matrix(3, 0.0, dim, pt, {}, 3)
    g = K.GRAD(K._const(-1.0), 1.0, "P", "P")
    _, grad = K.coeff_matrix(3, 0.0, dim, g, {}, 3)
    assert np.allclose(grad, -div.T, atol=1e-12)
    pde = K.continuity2(3)
    m3 = K.get_coeff_mats(pde, 2, 0.0, None)
    m8 = K.restrict(K.get_coeff_mats(pde, 2, 0.0, 8), pde, 2)
    for a, b in zip(sum(m3, []), sum(m8, [])):
        assert np.max(np.abs(a - b)) <= 1e-11 * max(1.0, np.max(np.abs(a)))


@pytest.mark.parametrize("deg,lev,a,b", [(2, 3, -1.0, 1.0), (3, 4, -2.0, 3.0), (4, 2, 0.0, 1.0)])
def test_div_flux_term_closed_form(deg, lev, a, b):
    """The Lax-Friedrichs part of coeff_matrix.m's DIV (periodic, constant g),
    pinned by what the mathematics fixes rather than by gold files (none use
    LF != 0): the central-flux matrix D0 is skew; S = D(LF) - D0 is symmetric,
    linear in LF and in |g|, has rank = number of cell faces (one jump per face),
    and for the upwind choice (g = -1, LF = -1) is negative semidefinite with its
    most negative eigenvalue -|g| |LF| deg (deg + 1) / h: orthonormal Legendre
    traces r_k = sqrt((2k+1)/h) and l_k = (-1)^k r_k give, in the Fourier mode of
    the periodic mesh that maximises it, |g| |LF| / 2 * (|r|^2 + |l|^2 + 2 |r.l|)."""
    dim = K.Dimension(a, b, lev)

    def D(g, LF):
        return K.coeff_matrix(deg, 0.0, dim, K.DIV(K._const(g), LF, "P", "P"), {}, lev)[0]

    D0 = D(-1.0, 0.0)
    S = D(-1.0, -1.0) - D0
    tol = 1e-12 * np.abs(S).max()
    assert np.abs(D0 + D0.T).max() <= 1e-11
    assert np.abs(S - S.T).max() <= tol
    assert np.abs((D(-1.0, 1.0) - D0) + S).max() <= tol
    assert np.abs((D(-2.0, -1.0) - D(-2.0, 0.0)) - 2 * S).max() <= 2 * tol
    assert np.linalg.matrix_rank(S, tol=1e-9 * np.abs(S).max()) == 2 ** lev
    ev = np.linalg.eigvalsh(S)
    h = (b - a) / 2 ** lev
    assert ev.max() <= 1e-11
    assert abs(ev.min() + deg * (deg + 1) / h) <= 1e-11 * deg * (deg + 1) / h


def _project(u, deg, lev, a, b):
    """Real-space DG coefficients of u: per cell, Gauss quadrature of u times the
    orthonormal Legendre basis sqrt(1/h) P~_k (exact for the polynomials below)."""
    n = 2 ** lev
    h = (b - a) / n
    qx, qw = K.lgwt(deg + 4, -1.0, 1.0)
    phi = K.lin_legendre(qx, deg) * np.sqrt(1.0 / h)
    out = np.zeros(deg * n)
    for i in range(n):
        out[deg * i:deg * (i + 1)] = (qw * u(a + (qx + 1) / 2 * h + i * h)) @ phi * h / 2
    return out


@pytest.mark.parametrize("deg,lev,a,b", [(3, 3, -1.0, 1.0), (4, 2, 0.0, 2.0)])
@pytest.mark.parametrize("LF", [0.0, -1.0])
def test_div_boundary_branches_exact_on_polynomials(deg, lev, a, b, LF):
    """coeff_matrix.m:212-262, the homogeneous parts of the Neumann and Dirichlet
    boundary fluxes (the inhomogeneous values go to boundary_condition_vector.m,
    not restated): with 'N' the boundary face takes the interior trace and with 'D'
    it takes zero, both consistent, so for constant g the DG derivative is exact
    on every polynomial of degree < deg (N) or on one vanishing at both ends (D):
    D P(u) = P(g u').  A periodic matrix on a non-periodic u is not exact (the
    check discriminates the boundary handling)."""
    g = 1.5
    dim = K.Dimension(a, b, lev)
    cases = {"N": (lambda x: x ** 2 - 0.3 * x, lambda x: 2 * x - 0.3),
             "D": (lambda x: (x - a) * (b - x), lambda x: (b - x) - (x - a))}
    for bc, (u, du) in cases.items():
        R = K.coeff_matrix(deg, 0.0, dim, K.DIV(K._const(g), LF, bc, bc), {}, lev)[1]
        ref = _project(lambda x: g * du(x), deg, lev, a, b)
        assert np.abs(R @ _project(u, deg, lev, a, b) - ref).max() <= 1e-12 * np.abs(ref).max(), bc
    u, du = cases["N"]
    Rp = K.coeff_matrix(deg, 0.0, dim, K.DIV(K._const(g), LF, "P", "P"), {}, lev)[1]
    ref = _project(lambda x: g * du(x), deg, lev, a, b)
    assert np.abs(Rp @ _project(u, deg, lev, a, b) - ref).max() > 1e-3 * np.abs(ref).max()


def _setup(pde, deg, grid_type="SG", max_lev=8):
    lev_vec = [d.lev for d in pde.dims]
    tab = G.hash_table_sparse_nd(lev_vec, max_lev, grid_type)
    mats = K.restrict(K.get_coeff_mats(pde, deg, 0.0, max_lev), pde, deg)
    return tab, mats


@pytest.mark.parametrize("name,pfn,lev,deg,grid", [
    ("continuity2_sg_l2_d2", K.continuity2, 2, 2, "SG"),
    ("continuity2_sg_l4_d3", K.continuity2, 4, 3, "SG"),
    ("continuity2_fg_l2_d2", K.continuity2, 2, 2, "FG"),
    ("continuity3_sg_l2_d2", K.continuity3, 2, 2, "SG"),
])
def test_rk3_trajectory_matches_gold(gold, name, pfn, lev, deg, grid):
    """time_advance.m:151-204 with the oracle apply: states after 1..5 steps."""
    pde = central_flux(pfn(lev))
    tab, mats = _setup(pde, deg, grid)
    inp = O.ApplyInputs(mats, tab.idx1d, deg)
    A = lambda v: O.apply_A(inp, v)  # noqa: E731
    src = lambda t: K.md_eval_function(pde, deg, tab.idx1d, pde.sources, t)  # noqa: E731
    f = K.md_eval_function(pde, deg, tab.idx1d, pde.initial_conditions, 0.0)
    dt = pde.set_dt(pde, 0.01)
    for i in range(5):
        t = i * dt
        k1 = A(f) + src(t)
        y1 = f + dt * 0.5 * k1
        k2 = A(y1) + src(t + dt / 2)
        y2 = f + dt * (-1.0 * k1 + 2.0 * k2)
        k3 = A(y2) + src(t + dt)
        f = f + dt * (k1 / 6 + 2 * k2 / 3 + k3 / 6)
        g = gold[f"time_advance/{name}_t{i}"].ravel()
        assert np.linalg.norm(f - g) <= 1e-13 * np.linalg.norm(g), i


def test_backward_euler_matches_gold(gold):
    """time_advance.m:267-349 (direct solve of (I - dt A) f1 = f0 + dt s(t+dt))."""
    pde = central_flux(K.continuity2(2))
    tab, mats = _setup(pde, 2)
    A = O.build_A(O.ApplyInputs(mats, tab.idx1d, 2))
    f = K.md_eval_function(pde, 2, tab.idx1d, pde.initial_conditions, 0.0)
    dt = pde.set_dt(pde, 0.01)
    Id = np.eye(A.shape[0])
    for i in range(5):
        f = np.linalg.solve(Id - dt * A, f + dt * K.md_eval_function(pde, 2, tab.idx1d, pde.sources, (i + 1) * dt))
        g = gold[f"time_advance/continuity2_implicit_l2_d2_t{i}"].ravel()
        assert np.linalg.norm(f - g) <= 1e-13 * np.linalg.norm(g), i


def test_planted_flux_error_is_caught(gold):
    """The upwind flux (LF = -1) does not reproduce the central-flux gold: the
    trajectory check can tell the two apart."""
    pde = K.continuity2(2)  # LF = -1
    tab, mats = _setup(pde, 2)
    inp = O.ApplyInputs(mats, tab.idx1d, 2)
    f = np.zeros(tab.num_elem * 4)
    dt = pde.set_dt(pde, 0.01)
    src = lambda t: K.md_eval_function(pde, 2, tab.idx1d, pde.sources, t)  # noqa: E731
    k1 = O.apply_A(inp, f) + src(0)
    y1 = f + dt * 0.5 * k1
    k2 = O.apply_A(inp, y1) + src(dt / 2)
    y2 = f + dt * (-k1 + 2 * k2)
    k3 = O.apply_A(inp, y2) + src(dt)
    f = f + dt * (k1 / 6 + 2 * k2 / 3 + k3 / 6)
    g = gold["time_advance/continuity2_sg_l2_d2_t0"].ravel()
    assert np.linalg.norm(f - g) > 1e-4 * np.linalg.norm(g)


@pytest.mark.parametrize("lv,ml,gt", [([7], 7, "SG"), ([5, 2], 7, "SG"), ([3, 2, 3], 7, "FG"), ([4, 4, 4], 8, "SG"),
                                      ([2, 3, 2, 3, 3, 2], 4, "SG")])
def test_product_element_tables_match_oracle(lv, ml, gt):
    t = G.hash_table_sparse_nd(lv, ml, gt)
    o = O.hash_table_sparse_nd(lv, ml, gt)
    for k in ("lev", "cell", "idx1d", "ids"):
        assert np.array_equal(getattr(t, k), o[k]), k


@pytest.mark.parametrize("nd,lev,gt", [(2, 3, "SG"), (3, 4, "SG"), (2, 3, "FG"), (3, 7, "SG"), (6, 3, "SG")])
def test_product_connectivity_matches_oracle(nd, lev, gt):
    t = G.hash_table_nd([lev] * nd, gt)
    o = O.hash_table_nd(nd, lev, gt)
    assert np.array_equal(t.idx1d, o["idx1d"])
    rp, cc = G.connect_nd(t, lev, lev, "SG")
    orp, occ = O.connect_nd(nd, lev, gt, lev, lev)
    assert np.array_equal(rp, orp) and np.array_equal(cc, occ)


def test_product_connectivity_gold(gold):
    t = G.hash_table_nd([3, 3], "FG")
    rp, cc = G.connect_nd(t, 6, 3, "SG")
    for i in range(64):
        assert np.array_equal(gold[f"connectivity/connect_n_2_3_FG_{i + 1}"].ravel().astype(int), cc[rp[i]:rp[i + 1]])

