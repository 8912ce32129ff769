"""Discontinuous-Galerkin multiwavelet 1D operators and projections: the CPU
input provider of the operator apply (host side, FP64, numpy).

Restates, in the reference's order of operations:
  lgwt.m (Gauss-Legendre rule), DG_loc/lin_legendre.m, lin_dlegendre.m,
  two_scale/OperatorTwoScale_wavelet2.m (FMWT), coeff_matrix.m (mass / div /
  grad with Lax-Friedrichs flux and P/D/N boundary handling, rotated to
  wavelet space), pterm_coeff_matrix.m (M \\ L), sd_term_coeff_matrix.m
  (chaining), get_coeff_mats.m, forward_wavelet_transform.m,
  combine_dimensions_D.m, md_eval_function.m.
The PDE specifications mirror pdes/continuity{1,2,3}.m and pdes/diffusion2.m;
`diffusion_nd` and `advection_diffusion_varcoeff` generalise them to the
6D LDG diffusion and variable-coefficient cases.
"""
import functools
import math
import os
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

_DATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "two_scale.npz")


# ------------------------------------------------------------- quadrature
@functools.lru_cache(maxsize=None)
def lgwt(N, a, b):
    """Quadrature/lgwt.m:16-59 -- N Gauss-Legendre nodes/weights on [a, b]."""
    N = N - 1
    N1, N2 = N + 1, N + 2
    xu = np.linspace(-1.0, 1.0, N1)
    y = np.cos((2 * np.arange(0, N + 1) + 1) * np.pi / (2 * N + 2)) + (0.27 / N1) * np.sin(np.pi * xu * N / N2)
    L = np.zeros((N1, N2))
    y0 = 2.0
    Lp = None
    for _ in range(200):
        if not np.max(np.abs(y - y0)) > np.finfo(float).eps:
            break
        L[:, 0] = 1.0
        L[:, 1] = y
        for k in range(2, N1 + 1):
            L[:, k] = ((2 * k - 1) * y * L[:, k - 1] - (k - 1) * L[:, k - 2]) / k
        Lp = N1 * (L[:, N1 - 1] - y * L[:, N2 - 1]) / (1 - y ** 2)
        y0 = y
        y = y0 - L[:, N2 - 1] / Lp
    x = (a * (1 - y) + b * (1 + y)) / 2
    w = (b - a) / ((1 - y ** 2) * Lp ** 2)
    return x[::-1].copy(), w[::-1].copy()


def default_quad_number(deg):
    """Quadrature/default_quad_number.m."""
    return max(deg + 1, 10)


def _legendre_table(x, k):
    """P_0..P_{k-1} at x (three-term recurrence)."""
    x = np.asarray(x, dtype=np.float64)
    P = np.zeros(x.shape + (max(k, 1),))
    P[..., 0] = 1.0
    if k > 1:
        P[..., 1] = x
    for n in range(2, k):
        P[..., n] = ((2 * n - 1) * x * P[..., n - 1] - (n - 1) * P[..., n - 2]) / n
    return P[..., :k]


def lin_legendre(x, k):
    """DG_loc/lin_legendre.m -- orthonormal Legendre on [-1,1], zero outside."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    v = _legendre_table(x, k) * np.sqrt(2 * np.arange(k) + 1.0)
    v[(x > 1) | (x < -1), :] = 0.0
    return v


def lin_dlegendre(x, k):
    """DG_loc/lin_dlegendre.m -- derivative of lin_legendre (dp_n = n P_{n-1} + x dp_{n-1})."""
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    P = _legendre_table(x, max(k, 1))
    dP = np.zeros(x.shape + (k,))
    for n in range(1, k):
        dP[..., n] = n * P[..., n - 1] + x * dP[..., n - 1]
    return dP * np.sqrt(2 * np.arange(k) + 1.0)


# ------------------------------------------------------------------- FMWT
@functools.lru_cache(maxsize=None)
def two_scale_filters(deg):
    z = np.load(_DATA)
    return z[f"H0_{deg}"].copy(), z[f"G0_{deg}"].copy()


@functools.lru_cache(maxsize=32)
def fmwt(deg, lev):
    """two_scale/OperatorTwoScale_wavelet2.m:1-144 -- dense forward multiwavelet
    transform (deg*2^lev square)."""
    if lev == 0:
        return np.eye(deg)
    H0, G0 = two_scale_filters(deg)
    tol = 1e4 * np.finfo(float).eps
    H0 = np.where(np.abs(H0) < tol, 0.0, H0)
    G0 = np.where(np.abs(G0) < tol, 0.0, G0)
    i = np.arange(deg)
    sgn = (-1.0) ** (i[:, None] + i[None, :])
    H1 = sgn * H0
    G1 = ((-1.0) ** deg) * sgn * G0
    n = deg * 2 ** lev
    Gmat = np.zeros((deg, n))
    Hmat = np.zeros((deg, n))
    Hmat[:, :deg] = np.eye(deg)
    blocks = {}
    for j in range(lev - 1, -1, -1):
        isize = n // (2 ** j)
        half = isize // 2
        Htmp = Hmat[:, :half].copy()
        Gmat[:, :half] = G0 @ Htmp
        Gmat[:, half:isize] = G1 @ Htmp
        Hmat[:, :half] = H0 @ Htmp
        Hmat[:, half:isize] = H1 @ Htmp
        blocks[2 * j + 2] = Gmat[:, :isize].copy()
        blocks[2 * j + 1] = Hmat[:, : (n if j == 0 else isize)].copy()
    F = np.zeros((n, n))
    F[:deg, :] = blocks[1]
    irow = deg
    for j in range(lev):
        ncells = 2 ** j
        isize = n // ncells
        Fm = blocks[2 + 2 * j]
        for c in range(ncells):
            F[irow + c * deg: irow + (c + 1) * deg, c * isize:(c + 1) * isize] = Fm
        irow += deg * ncells
    return F


def matrix_plot_D(deg, lev, xmin, xmax, output_grid="quadrature"):
    """matrix_plot_D.m:1-172 for its default output grid ('quadrature', OPTS.m:30):
    the evaluation matrix taking a 1D wavelet-space coefficient vector to values
    at the deg Gauss-Legendre points of each of the 2^lev cells of [xmin, xmax]
    (deg * 2^lev rows, points in cell order; columns as the wavelet basis).  Cell i
    holds the orthonormal Legendre values at its points scaled by sqrt(1/h)
    (:138), columns deg*i .. deg*(i+1)-1 of the real-space basis (:148, :168); the
    wavelet columns come from the inverse transform, Meval * FMWT^T (:172).
    Returns (Meval, nodes)."""
    if output_grid != "quadrature":
        raise NotImplementedError(f"matrix_plot_D: output_grid {output_grid!r} (only 'quadrature')")
    n = 2 ** lev
    h = (xmax - xmin) / n
    qx, _ = lgwt(deg, -1.0, 1.0)
    p_val = lin_legendre(qx, deg) * np.sqrt(1.0 / h)
    M = np.zeros((deg * n, deg * n))
    nodes = np.zeros(deg * n)
    for i in range(n):
        M[deg * i:deg * (i + 1), deg * i:deg * (i + 1)] = p_val
        nodes[deg * i:deg * (i + 1)] = ((qx + 1) / 2 + i) * h + xmin
    return M @ fmwt(deg, lev).T, nodes


# ------------------------------------------------------------ term classes
def _one(x, p=None, t=None, dat=None):
    return np.zeros_like(np.asarray(x, dtype=np.float64)) + 1.0


@dataclass
class PTerm:
    """PARTIAL_SD_TERM.m / MASS.m / DIV.m / GRAD.m (DIV_OR_GRAD.m)."""
    type: str
    g: Callable = _one
    LHS_mass_g: Callable = _one
    dV: Callable = _one
    LF: float = 0.0
    BCL: str = "N"
    BCR: str = "N"
    IBCL: str = "N"
    IBCR: str = "N"
    BCL_fList: Optional[list] = None
    BCR_fList: Optional[list] = None


def MASS(g=None, LHS_mass_g=None, dV=None):
    return PTerm("mass", g or _one, LHS_mass_g or _one, dV or _one)


def DIV(g=None, LF=0.0, BCL="N", BCR="N", dV=None, LHS_mass_g=None, BCL_fList=None, BCR_fList=None):
    """DIV.m: IBC = BC."""
    return PTerm("div", g or _one, LHS_mass_g or _one, dV or _one, LF, BCL, BCR, BCL, BCR, BCL_fList, BCR_fList)


def GRAD(g=None, LF=0.0, BCL="N", BCR="N", dV=None, LHS_mass_g=None, BCL_fList=None, BCR_fList=None):
    """GRAD.m: D and N swap for the internal BCs, LF changes sign."""
    sw = {"D": "N", "N": "D"}
    return PTerm("grad", g or _one, LHS_mass_g or _one, dV or _one, -LF, BCL, BCR, sw.get(BCL, BCL),
                 sw.get(BCR, BCR), BCL_fList, BCR_fList)


@dataclass
class Dimension:
    min: float
    max: float
    lev: int = 3
    moment_dV: Callable = _one
    name: str = "x"


@dataclass
class PDE:
    dims: List[Dimension]
    terms: List[List[List[PTerm]]]  # terms[t][d] = chain of pterms (SD_TERM)
    sources: list = field(default_factory=list)  # each: [f_1(x,p,t), ..., f_d, g(t)]
    initial_conditions: list = field(default_factory=list)
    solutions: list = field(default_factory=list)
    params: dict = field(default_factory=dict)
    set_dt: Optional[Callable] = None
    name: str = ""

    @property
    def ndim(self):
        return len(self.dims)


# ---------------------------------------------------------- coeff_matrix.m
def coeff_matrix(deg, t, dim, pterm, params, lev):
    """coeff_matrix.m:1-374.  Returns (wavelet-space matrix, real-space matrix)."""
    assert lev >= dim.lev
    N = 2 ** lev
    h = (dim.max - dim.min) / N
    dof = deg * N
    qx, qw = lgwt(default_quad_number(deg), -1, 1)
    p_L = lin_legendre(-1.0, deg)[0] / np.sqrt(h)
    p_R = lin_legendre(+1.0, deg)[0] / np.sqrt(h)
    p_val = lin_legendre(qx, deg) / np.sqrt(h)
    Dp_val = lin_dlegendre(qx, deg) / np.sqrt(h) * 2 / h
    Jac = h / 2
    F = fmwt(deg, lev)
    dat_R = F.T @ np.ones(dof)  # pterm.dat empty -> ones, to real space (:99-103)
    mass = np.zeros((dof, dof))
    div = np.zeros((dof, dof))
    is_dg = pterm.type in ("div", "grad")
    FluxVal = pterm.LF
    o = np.outer
    for i in range(N):
        xL = dim.min + i * h
        xR = xL + h
        c = np.arange(deg * i, deg * (i + 1))
        qxi = ((qx + 1) / 2 + i) * h + dim.min
        dq = p_val @ dat_R[c]
        G1 = pterm.g(qxi, params, t, dq) * pterm.dV(qxi, params, t)
        val_mass = p_val.T @ (G1[:, None] * p_val * qw[:, None]) * Jac
        val_div = -Dp_val.T @ (G1[:, None] * p_val * qw[:, None]) * Jac
        if np.isnan(np.linalg.norm(G1)):
            raise FloatingPointError("coeff_matrix: NaN coefficient")
        mass[np.ix_(c, c)] += val_mass
        div[np.ix_(c, c)] += val_div
        if not is_dg:
            continue
        FCL = float(np.ravel(pterm.g(np.array([xL]), params, t, dq) * pterm.dV(np.array([xL]), params, t))[0])
        FCR = float(np.ravel(pterm.g(np.array([xR]), params, t, dq) * pterm.dV(np.array([xR]), params, t))[0])
        Z = np.zeros((deg, deg))
        blocks = [
            o(-p_L, p_R) * FCL / 2 + FluxVal * abs(FCL) / 2 * o(-p_L, p_R),
            o(-p_L, p_L) * FCL / 2 + FluxVal * abs(FCL) / 2 * o(-p_L, -p_L),
            o(p_R, p_R) * FCR / 2 + FluxVal * abs(FCR) / 2 * o(p_R, p_R),
            o(p_R, p_L) * FCR / 2 + FluxVal * abs(FCR) / 2 * o(p_R, -p_L),
        ]
        if pterm.IBCL == "D" and i == 0:
            blocks[0] = Z.copy()
            blocks[1] = Z.copy()
        if pterm.IBCR == "D" and i == N - 1:
            blocks[2] = Z.copy()
            blocks[3] = Z.copy()
        if pterm.IBCL == "N" and i == 0:
            blocks[0] = Z.copy()
            blocks[1] = o(-p_L, p_L) * FCL
        if pterm.IBCR == "N" and i == N - 1:
            blocks[2] = o(p_R, p_R) * FCR
            blocks[3] = Z.copy()
        cols = [c - deg, c, c, c + deg]
        use = [0, 1, 2, 3]
        if i == 0:
            use = [1, 2, 3]
        elif i == N - 1:
            use = [0, 1, 2]
        if pterm.IBCL == "P" or pterm.IBCR == "P":
            if lev >= 1:
                if i == 0:
                    cols[0] = np.arange(deg * (N - 1), deg * N)
                if i == N - 1:
                    cols[3] = np.arange(0, deg)
            else:
                cols = [c, c, c, c]
            use = [0, 1, 2, 3]
        for u in use:
            div[np.ix_(c, cols[u])] += blocks[u]
    grad = -div.T
    mats = {"mass": mass, "div": div, "grad": grad}
    unrot = mats[pterm.type]
    return F @ unrot @ F.T, unrot


_PTERM_CACHE = {}


def _params_key(params):
    try:
        return tuple(sorted(params.items()))
    except TypeError:
        return ("id", id(params))


def pterm_coeff_matrix(deg, t, dim, pterm, params, lev):
    """pterm_coeff_matrix.m:1-26: the pterm's RHS matrix L moved to the RHS with
    its own LHS mass matrix M: mat = M \\ L.

    The result is a function of the pterm's fields, the dimension's domain and
    moment_dV, deg, lev, t and params only, so it is built once per distinct
    combination and reused (the functions g / LHS_mass_g / dV / moment_dV are
    keyed by identity and kept alive by the cache, so an id is never reused
    while its entry exists).  Callers get read-only arrays."""
    fns = (pterm.g, pterm.LHS_mass_g, pterm.dV, dim.moment_dV)
    key = (deg, float(t), float(dim.min), float(dim.max), lev, pterm.type, float(pterm.LF), pterm.BCL, pterm.BCR,
           pterm.IBCL, pterm.IBCR, tuple(id(f) for f in fns), _params_key(params))
    hit = _PTERM_CACHE.get(key)
    if hit is not None and all(a is b for a, b in zip(hit[0], fns)):
        return hit[1]
    L, L_un = coeff_matrix(deg, t, dim, pterm, params, lev)
    M, M_un = coeff_matrix(deg, t, dim, MASS(pterm.LHS_mass_g, None, dim.moment_dV), params, lev)
    out = (np.linalg.solve(M, L), np.linalg.solve(M_un, L_un), M)
    for a in out:
        a.setflags(write=False)
    if len(_PTERM_CACHE) > 512:
        _PTERM_CACHE.clear()
    _PTERM_CACHE[key] = (fns, out)
    return out


def sd_term_coeff_matrix(deg, t, dim, pterms, params, lev):
    """sd_term_coeff_matrix.m:1-27: chain the pterm matrices."""
    mat = None
    for pt in pterms:
        m, _, _ = pterm_coeff_matrix(deg, t, dim, pt, params, lev)
        mat = m if mat is None else mat @ m
    return mat


def get_coeff_mats(pde, deg, t=0.0, max_lev=None):
    """get_coeff_mats.m:29-53 with max_lev_coeffs: build at max_lev when given."""
    out = []
    for tm in pde.terms:
        row = []
        for d, dim in enumerate(pde.dims):
            lev = dim.lev if max_lev is None else max_lev
            row.append(sd_term_coeff_matrix(deg, t, dim, tm[d], pde.params, lev))
        out.append(row)
    return out


def restrict(mats, pde, deg):
    """Top-left deg*2^lev_d block of each matrix built at a larger level (the
    hierarchical index of a coarse element is the same at every level)."""
    return [[m[: deg * 2 ** pde.dims[d].lev, : deg * 2 ** pde.dims[d].lev] for d, m in enumerate(tm)] for tm in mats]


# ------------------------------------------------------------ projections
def forward_wavelet_transform(deg, lev, lmin, lmax, foo, dV, params, t):
    """forward_wavelet_transform.m:1-65."""
    qx, qw = lgwt(default_quad_number(deg), -1, 1)
    n = 2 ** lev
    h = (lmax - lmin) / n
    p_val = (lin_legendre(qx, deg) / np.sqrt(h)).T
    f = np.zeros(deg * n)
    for i in range(n):
        x = h * (qx / 2 + 0.5 + i) + lmin
        fx = np.asarray(foo(x, params, t), dtype=np.float64) * np.asarray(dV(x), dtype=np.float64)
        f[i * deg:(i + 1) * deg] = p_val @ (qw * fx)
    f = f * h / 2
    f = fmwt(deg, lev) @ f
    return f * (np.abs(f) > 1e-12)


def _at(fn, x, params, t):
    """fn(x, params, t) at one point, as a float."""
    return float(np.ravel(np.asarray(fn(np.array([x], dtype=np.float64), params, t), dtype=np.float64))[0])


def compute_boundary_condition(g, dV, t, lev, deg, xmin, xmax, bc_func, side, params=None):
    """compute_boundary_condition.m:1-55: the trace term (g bc v)|_boundary of one
    side, deg*2^lev values, nonzero in the boundary cell only: -g bc dV p_L at
    xMin ('L'), +g bc dV p_R at xMax ('R'), with g and dV taken h*1e-7 inside
    the domain when g is not finite there (the reference's 'R' branch reads an
    undefined name in that case; here it uses the same trace value as the finite
    branch)."""
    n = 2 ** lev
    h = (xmax - xmin) / n
    bc = np.zeros(deg * n)
    if side == "L":
        iv = lin_legendre(-1.0, deg)[0] / np.sqrt(h) * _at(bc_func, xmin, params, t)
        x = xmin if np.isfinite(_at(g, xmin, params, t)) else xmin + h * 1e-7
        bc[:deg] = -_at(g, x, params, t) * iv * _at(dV, x, params, t)
    else:
        iv = lin_legendre(1.0, deg)[0] / np.sqrt(h) * _at(bc_func, xmax, params, t)
        x = xmax if np.isfinite(_at(g, xmax, params, t)) else xmax - h * 1e-7
        bc[deg * (n - 1):] = _at(g, x, params, t) * iv * _at(dV, x, params, t)
    return bc


def boundary_condition_terms(pde, deg, idx1d, t=0.0, max_lev=None):
    """boundary_condition_vector.m:1-197 as a list of (time factor, vector): one
    pair per DIV / GRAD pterm side with Dirichlet data, the vector being that
    side's terms combined over the elements with time factor 1, its spatial parts
    evaluated at t.  Per side: in its own dimension the trace term
    (compute_boundary_condition), rotated to wavelet space, moved to the RHS with
    the pterm's LHS mass matrix and multiplied by the matrices of the pterms
    before it in the chain; in every other dimension the projection of the side's
    boundary function times that dimension's p-th pterm g, through that pterm's
    mass solve and the same dimension's preceding pterms.  The 1D matrices are
    those get_coeff_mats builds (at max_lev when given, cut to each dimension's
    level).  A side without boundary functions (the DIV / GRAD default: zero
    functions) has no term."""
    nd = pde.ndim
    params = pde.params
    idx1d = np.asarray(idx1d)
    out = []

    def pmats(d, pterm):
        dim = pde.dims[d]
        lev = dim.lev if max_lev is None else max_lev
        return pterm_coeff_matrix(deg, 0.0, dim, pterm, params, lev)  # (mat, unrotated, LHS mass)

    for tm in pde.terms:
        for d1, dim in enumerate(pde.dims):
            chain = tm[d1]
            dof = deg * 2 ** dim.lev
            for p, pterm in enumerate(chain):
                if pterm.type not in ("grad", "div"):
                    continue
                for side, kind, flist in (("L", pterm.BCL, pterm.BCL_fList), ("R", pterm.BCR, pterm.BCR_fList)):
                    if kind != "D" or flist is None:
                        continue
                    parts = []
                    for d2 in range(nd):
                        if d2 == d1:
                            tmp = compute_boundary_condition(pterm.g, pterm.dV, t, dim.lev, deg, dim.min, dim.max,
                                                             flist[d1], side, params)
                            tmp = fmwt(deg, dim.lev) @ tmp
                            M = pmats(d1, pterm)[2][:dof, :dof]
                            tmp = np.linalg.solve(M, tmp)
                            pre = np.eye(dof)
                            for q in range(p):
                                pre = pre @ pmats(d1, chain[q])[0][:dof, :dof]
                            parts.append(pre @ tmp)
                            continue
                        dim2 = pde.dims[d2]
                        other = tm[d2]
                        mp = other[p]
                        fn = flist[d2]
                        v = forward_wavelet_transform(
                            deg, dim2.lev, dim2.min, dim2.max,
                            lambda x, pp, tt, fn=fn, mp=mp: np.asarray(fn(x, pp, tt)) * np.asarray(mp.g(x, pp, tt)),
                            lambda x, mp=mp: mp.dV(x, params, t), params, t)
                        N = v.size
                        v = np.linalg.solve(pmats(d2, mp)[2][:N, :N], v)
                        for q in range(p):
                            v = pmats(d2, other[q])[0][:N, :N] @ v
                        parts.append(v)
                    out.append((lambda tt, fn=flist[nd]: float(fn(tt, params)),
                                combine_dimensions_D(deg, parts, 1.0, idx1d)))
    return out


def boundary_condition_vector(pde, deg, idx1d, t, max_lev=None):
    """boundary_condition_vector.m:1-197: the inhomogeneous Dirichlet terms of
    every DIV / GRAD pterm with a 'D' side at time t, summed over terms,
    dimensions, pterms and sides (each side's vector times its time factor)."""
    out = np.zeros(np.asarray(idx1d).shape[0] * deg ** pde.ndim)
    for fac, vec in boundary_condition_terms(pde, deg, idx1d, t, max_lev):
        out += vec * fac(t)
    return out


def combine_dimensions_D(deg, fList, time_multiplier, idx1d):
    """combine_dimensions_D.m:1-57: per element, krond of the per-dim blocks."""
    N, nd = idx1d.shape
    B = deg ** nd
    out = np.zeros(N * B)
    for e in range(N):
        v = np.ones(1)
        for d in range(nd):
            I = idx1d[e, d] - 1
            v = np.kron(v, fList[d][I * deg:(I + 1) * deg])
        out[e * B:(e + 1) * B] += v * time_multiplier
    return out


def md_eval_function(pde, deg, idx1d, md_funcs, t):
    """md_eval_function.m:1-50 (unit moment_dV: the dimension mass matrix is
    the identity up to roundoff and is not applied)."""
    nd = pde.ndim
    fval = np.zeros(idx1d.shape[0] * deg ** nd)
    for md in md_funcs:
        fl = [forward_wavelet_transform(deg, pde.dims[d].lev, pde.dims[d].min, pde.dims[d].max, md[d],
                                        lambda x, d=d: pde.dims[d].moment_dV(x), pde.params, t) for d in range(nd)]
        fval += combine_dimensions_D(deg, fl, md[nd](t), idx1d)
    return fval


# ------------------------------------------------------------- PDE specs
@functools.lru_cache(maxsize=None)
def _const(c):
    """The constant function c (one object per value, so equal pterms share their matrices)."""
    return lambda x, p=None, t=None, dat=None: np.zeros_like(np.asarray(x, dtype=np.float64)) + c


def continuity1(lev=2):
    """pdes/continuity1.m (1D advection, periodic)."""
    dims = [Dimension(-1, 1, lev)]
    terms = [[[DIV(_const(-1), -1, "P", "P")]]]
    sources = [[lambda x, p, t: np.cos(2 * np.pi * x), lambda t: -2 * np.pi * np.sin(t)],
               [lambda x, p, t: np.sin(2 * np.pi * x), lambda t: -np.cos(t)]]
    sol = [[lambda x, p, t: np.cos(2 * np.pi * x), lambda t: np.sin(t)]]
    return PDE(dims, terms, sources, sol, sol, {}, lambda pde, cfl: (pde.dims[0].max - pde.dims[0].min) / 2 ** pde.dims[0].lev * cfl, "continuity1")


def continuity2(lev=3):
    """pdes/continuity2.m: df/dt = -df/dx - df/dy + s on [-1,1] x [-2,2], periodic."""
    levs = np.broadcast_to(np.asarray(lev), (2,))
    dims = [Dimension(-1, 1, int(levs[0]), name="x"), Dimension(-2, 2, int(levs[1]), name="y")]
    terms = [
        [[DIV(_const(-1), -1, "P", "P")], [MASS(_const(1))]],
        [[MASS()], [DIV(_const(-1), -1, "P", "P")]],
    ]
    sources = [
        [lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y), lambda t: 2 * np.cos(2 * t)],
        [lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.cos(2 * np.pi * y), lambda t: 2 * np.pi * np.sin(2 * t)],
        [lambda x, p, t: np.sin(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y), lambda t: -np.pi * np.sin(2 * t)],
    ]
    sol = [[lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y), lambda t: np.sin(2 * t)]]
    set_dt = lambda pde, cfl: (pde.dims[0].max - pde.dims[0].min) / 2 ** pde.dims[0].lev * cfl  # noqa: E731
    return PDE(dims, terms, sources, sol, sol, {"parameter1": 0, "parameter2": 1}, set_dt, "continuity2")


def continuity3(lev=3):
    """pdes/continuity3.m: three upwind DIV(g=-1, LF=-1, P/P) terms, one per dim,
    on [-1,1] x [-2,2] x [-3,3]."""
    levs = np.broadcast_to(np.asarray(lev), (3,))
    dims = [Dimension(-1, 1, int(levs[0]), name="x"), Dimension(-2, 2, int(levs[1]), name="y"),
            Dimension(-3, 3, int(levs[2]), name="z")]
    terms = []
    for k in range(3):
        terms.append([[DIV(_const(-1), -1, "P", "P")] if d == k else [MASS()] for d in range(3)])
    c3 = lambda z, p, t: np.cos(2 * np.pi * z / 3)  # noqa: E731
    sources = [
        [lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y), c3, lambda t: 2 * np.cos(2 * t)],
        [lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.cos(2 * np.pi * y), c3, lambda t: 2 * np.pi * np.sin(2 * t)],
        [lambda x, p, t: np.sin(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y), c3, lambda t: -np.pi * np.sin(2 * t)],
        [lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y),
         lambda z, p, t: np.sin(2 * np.pi * z / 3), lambda t: -2 / 3 * np.pi * np.sin(2 * t)],
    ]
    sol = [[lambda x, p, t: np.cos(np.pi * x), lambda y, p, t: np.sin(2 * np.pi * y), c3, lambda t: np.sin(2 * t)]]
    zero = lambda x, p, t: np.zeros_like(x)  # noqa: E731
    ic = [[zero, zero, zero, lambda t: 1.0]]
    set_dt = lambda pde, cfl: (pde.dims[0].max - pde.dims[0].min) / 2 ** pde.dims[0].lev * cfl  # noqa: E731
    return PDE(dims, terms, sources, ic, sol, {"parameter1": 0, "parameter2": 1}, set_dt, "continuity3")


def diffusion_nd(nd=2, lev=3):
    """pdes/diffusion2.m generalised to nd dims: per dim the LDG chain
    DIV(g=1, LF=+1, N/N) * GRAD(g=1, LF=-1, D/D) with the solution's Dirichlet
    data, MASS*MASS in the other dims, on [0,1]^nd; solution prod cos(pi x_i)
    exp(-nd pi^2 t)."""
    levs = np.broadcast_to(np.asarray(lev), (nd,))
    dims = [Dimension(0, 1, int(levs[d]), name=f"x{d + 1}") for d in range(nd)]
    cosf = lambda x, p=None, t=None: np.cos(np.pi * np.asarray(x, dtype=np.float64))  # noqa: E731
    flist = [cosf] * nd + [lambda t, p=None: np.exp(-nd * np.pi ** 2 * t)]  # the solution's boundary values
    terms = []
    for k in range(nd):
        terms.append([[DIV(_const(1), +1, "N", "N"), GRAD(_const(1), -1, "D", "D", BCL_fList=flist, BCR_fList=flist)]
                      if d == k else [MASS(_const(1)), MASS(_const(1))] for d in range(nd)])
    sol = [[cosf] * nd + [lambda t: np.exp(-nd * np.pi ** 2 * t)]]
    set_dt = lambda pde, cfl: cfl * (1.0 / 2 ** pde.dims[0].lev) ** 2  # noqa: E731
    return PDE(dims, terms, [], sol, sol, {"parameter1": 0, "parameter2": 1}, set_dt, f"diffusion{nd}")


@functools.lru_cache(maxsize=None)
def _varcoeff_funcs(nu):
    """g(x) = 1 + 0.5 sin(pi x), -g and nu g (one set of objects per nu, so PDEs
    of any dimension share their 1D matrices)."""
    gvar = lambda x, p=None, t=None, dat=None: 1.0 + 0.5 * np.sin(np.pi * np.asarray(x, dtype=np.float64))  # noqa: E731
    gneg = lambda x, p=None, t=None, dat=None: -gvar(x)  # noqa: E731
    gdif = lambda x, p=None, t=None, dat=None: nu * gvar(x)  # noqa: E731
    return gvar, gneg, gdif


def advection_diffusion_varcoeff(nd=3, lev=3, nu=0.02):
    """Variable-coefficient advection-diffusion on [-1,1]^nd, periodic: per dim
    an upwind DIV(-g_i, LF=-1) term and an LDG diffusion chain
    DIV(nu g_i, LF=+1) * GRAD(1, LF=-1), with g_i(x) = 1 + 0.5 sin(pi x_i) in
    the pterms (coeff_matrix.m:146).  nu keeps the explicit RK3 step of the
    advective CFL stable (|lambda_diff| ~ nu * 2.8e6 at lev 5, deg 3)."""
    levs = np.broadcast_to(np.asarray(lev), (nd,))
    dims = [Dimension(-1, 1, int(levs[d]), name=f"x{d + 1}") for d in range(nd)]
    gvar, gneg, gdif = _varcoeff_funcs(float(nu))
    terms = []
    for k in range(nd):
        terms.append([[DIV(gneg, -1, "P", "P")] if d == k else [MASS()] for d in range(nd)])
    for k in range(nd):
        terms.append([[DIV(gdif, +1, "P", "P"), GRAD(_const(1), -1, "P", "P")] if d == k else [MASS(), MASS()]
                      for d in range(nd)])
    bump = lambda x, p, t: np.exp(-4 * np.asarray(x) ** 2)  # noqa: E731
    ic = [[bump] * nd + [lambda t: 1.0]]
    set_dt = lambda pde, cfl: (pde.dims[0].max - pde.dims[0].min) / 2 ** pde.dims[0].lev * cfl  # noqa: E731
    return PDE(dims, terms, [], ic, [], {}, set_dt, f"advdiff{nd}_varcoeff")


def diffusion1(lev=3, nu=np.pi / 2, lf_div=-1.0, lf_grad=+1.0):
    """pdes/diffusion1.m: df/dt = d2f/dx2 + s on [0, 1] by LDG, DIV(g=1, LF=-1,
    N/N) GRAD(g=1, LF=+1, D/D) with inhomogeneous Dirichlet data cos(nu x)
    exp(-2 nu^2 t), nu = pi/2, plus the two zero-coefficient penalty terms;
    source -nu^2 cos(nu x) exp(-2 nu^2 t).  The reference's gold boundary-vector
    and RK3 fixtures for diffusion1 were written by a revision with nu = 0.01 and
    diffusion2's fluxes, DIV LF = +1, GRAD LF = -1 (`diffusion1_gold`)."""
    soln_x = lambda x, p=None, t=None: np.cos(nu * np.asarray(x, dtype=np.float64))  # noqa: E731
    soln_t = lambda t, p=None: np.exp(-2 * nu ** 2 * t)  # noqa: E731
    flist = [soln_x, soln_t]
    dims = [Dimension(0.0, 1.0, int(np.ravel(lev)[0]), name="x")]
    terms = [
        [[DIV(_const(1.0), lf_div, "N", "N"), GRAD(_const(1.0), lf_grad, "D", "D", BCL_fList=flist, BCR_fList=flist)]],
        [[DIV(_const(0.0), -1, "D", "D")]],
        [[DIV(_const(-0.0), 0, "D", "D")]],
    ]
    sources = [[lambda x, p, t: -nu ** 2 * np.cos(nu * np.asarray(x)), lambda t: np.exp(-2 * nu ** 2 * t)]]
    sol = [[soln_x, lambda t: np.exp(-2 * nu ** 2 * t)]]
    set_dt = lambda pde, cfl: cfl * (1.0 / 2 ** pde.dims[0].lev) ** 2  # noqa: E731
    return PDE(dims, terms, sources, sol, sol, {"parameter1": 0, "parameter2": 1}, set_dt, "diffusion1")


def diffusion1_gold(lev=3):
    """diffusion1 as the reference's gold fixtures were written (nu = 0.01, DIV
    LF = +1, GRAD LF = -1): reproduces gold boundary_conditions/vector_diffusion1_*
    and time_advance/diffusion1_sg_* (operators built at the dimension level)."""
    return diffusion1(lev, nu=0.01, lf_div=+1.0, lf_grad=-1.0)


def diffusion2(lev=3):
    """pdes/diffusion2.m: df/dt = d2f/dx2 + d2f/dy2 on [0, 1]^2 by LDG, per
    dimension DIV(g=1, LF=+1, N/N) GRAD(g=1, LF=-1, D/D) with inhomogeneous
    Dirichlet data cos(pi x) cos(pi y) exp(-2 pi^2 t), MASS MASS in the other
    dimension; no source."""
    levs = np.broadcast_to(np.asarray(lev), (2,))
    cx = lambda x, p=None, t=None: np.cos(np.pi * np.asarray(x, dtype=np.float64))  # noqa: E731
    st = lambda t, p=None: np.exp(-2 * np.pi ** 2 * t)  # noqa: E731
    flist = [cx, cx, st]
    dims = [Dimension(0.0, 1.0, int(levs[0]), name="x"), Dimension(0.0, 1.0, int(levs[1]), name="y")]
    one = _const(1.0)
    ldg = [DIV(one, +1, "N", "N"), GRAD(one, -1, "D", "D", BCL_fList=flist, BCR_fList=flist)]
    mm = [MASS(one), MASS(one)]
    terms = [[ldg, mm], [mm, ldg]]
    sol = [[cx, cx, lambda t: np.exp(-2 * np.pi ** 2 * t)]]
    set_dt = lambda pde, cfl: cfl * (1.0 / 2 ** pde.dims[0].lev) ** 2  # noqa: E731
    return PDE(dims, terms, [], sol, sol, {"parameter1": 0, "parameter2": 1}, set_dt, "diffusion2")


PDES = {"continuity1": continuity1, "continuity2": continuity2, "continuity3": continuity3}


def operator_mats(name, nd, lev, deg):
    """1D matrices of a named PDE at level `lev` (bench / harness helper).
    Returns (mats[t][d], description)."""
    if name == "continuity3":
        pde = continuity3(lev)
    elif name == "continuity2":
        pde = continuity2(lev)
    elif name.startswith("diffusion"):
        pde = diffusion_nd(nd, lev)
    else:
        raise KeyError(name)
    mats = get_coeff_mats(pde, deg)
    return mats, f"{pde.name} DG (deg {deg}, lev {lev}, LF upwind, periodic), wavelet space"
