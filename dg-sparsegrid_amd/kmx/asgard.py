"""Host-side mirror of ASGarD-MATLAB's operator-application interface.

Same names and argument meaning as the reference, executing on the gfx950
library:

    ftmp = apply_A(pde, opts, A_data, f, deg, Vmax, Emax, imex_flag)   apply_A.m:1
    A    = build_A(pde, opts, A_data, deg, imex_flag, sparse)          apply_A.m:170-180, 246-248
    f    = time_advance(pde, opts, A_data, f, t, dt, deg, hash_table)   time_advance.m:1
    Meval, nodes = matrix_plot_D(pde, opts, dimension)                matrix_plot_D.m:1
    fr   = wavelet_to_realspace(pde, opts, Meval_D, f, hash_table)      wavelet_to_realspace.m:1
    err, fval = asgard(pde_fn, **opts)                                  asgard.m:3 (RK3/BE subset)

`pde` is a kmx.coeffs.PDE whose 1D matrices have been built with
get_coeff_mats (stored as pde.mats[t][d]); `A_data` comes from
kmx.grid.global_matrix; connectivity, when used, is pde.connectivity =
(row_ptr, cols) from kmx.grid.connect_nd.  Vectors may be numpy arrays (a
host round trip, like the MEX path) or float64 CUDA tensors (resident).
"""
import hashlib
from dataclasses import dataclass, field

import numpy as np

from . import coeffs as C
from . import grid as G
from . import timestep as TS
from .plan import Plan, RealSpace


@dataclass
class OPTS:
    """OPTS.m:5-41 defaults for the fields this path reads."""
    lev: object = 3
    deg: int = 2
    CFL: float = 0.01
    grid_type: str = "SG"
    timestep_method: str = "RK3"
    use_oldhash: bool = False
    use_connectivity: bool = False
    max_lev: int = 8
    max_lev_coeffs: bool = False  # OPTS.m:25 defaults it on, OPTS.m:168 turns it off without adaptivity
    num_steps: int = 5
    quiet: bool = True
    output_grid: str = "quadrature"  # OPTS.m:30 (matrix_plot_D restates this grid only)
    prec: str = "fp64"  # contraction precision of the device apply (not a reference option)
    extra: dict = field(default_factory=dict)


def _table_key(A_data):
    """A stable key of the element table (its 1D indices), not an object id."""
    idx = np.ascontiguousarray(np.asarray(A_data["element_local_index_D"], dtype=np.int32))
    return idx.shape, hashlib.sha1(idx.tobytes()).hexdigest()


def _plan_for(pde, opts, A_data, deg, imex_flag):
    """The resident plan for this operator, cached on the pde object per (element
    table, deg, imex flag, connectivity).  The cache entry holds the 1D matrices
    it was built from (so their identities stay valid): when pde.mats entries are
    replaced -- the time-dependent refresh of get_coeff_mats -- the changed
    matrices go to the plan through kmx_plan_update_term, or the plan is rebuilt
    if the shapes or the connectivity changed."""
    flags = getattr(pde, "imex", None) or ["N"] * len(pde.mats)
    term_vec = [t for t in range(len(pde.mats)) if flags[t] == imex_flag]  # apply_A.m:23-30
    mats = [list(pde.mats[t]) for t in term_vec]
    conn = pde.connectivity if opts.use_connectivity else None
    key = (_table_key(A_data), deg, imex_flag, bool(opts.use_connectivity))
    cache = pde.__dict__.setdefault("_kmx_plans", {})
    ent = cache.get(key)
    if ent is not None:
        plan, held, held_conn = ent
        if plan is None and not mats:
            return None
        if plan is not None and held_conn is conn and len(held) == len(mats) and all(
                len(a) == len(b) and all(np.shape(x) == np.shape(y) for x, y in zip(a, b)) for a, b in zip(held, mats)):
            for t, (old, new) in enumerate(zip(held, mats)):
                for d, (x, y) in enumerate(zip(old, new)):
                    if x is not y:
                        plan.update_term(t, d, y)
            cache[key] = (plan, mats, conn)
            return plan
    idx = np.asarray(A_data["element_local_index_D"])
    rp = cols = None
    if opts.use_connectivity:
        rp, cols = conn
    plan = Plan(mats, idx, deg, rp, cols, index_base=1, fp32=True, adjoint=True) if mats else None
    cache[key] = (plan, mats, conn)
    return plan


def apply_A(pde, opts, A_data, f, deg, Vmax=None, Emax=None, imex_flag="N"):
    """apply_A.m:1-252, explicit branch: ftmp = A f over the pair list and the
    terms whose imex flag matches.  Vmax/Emax are unused, as in the reference."""
    import torch

    plan = _plan_for(pde, opts, A_data, deg, imex_flag)
    host = isinstance(f, np.ndarray)
    x = torch.from_numpy(np.ascontiguousarray(f, dtype=np.float64)).cuda() if host else f
    if plan is None:
        y = torch.zeros_like(x)
    else:
        y = plan.apply(x, prec=opts.prec)
    return y.cpu().numpy() if host else y


def build_A(pde, opts, A_data, deg, imex_flag="N", sparse=True):
    """apply_A.m's build_A branch (apply_A.m:150-155, 170-180, 246-248): the
    assembled operator of the terms whose imex flag matches, from the device's
    per-pair blocks (kmx_plan_assemble).  sparse=True gives what use_sparse_A
    builds, sparse(A_s2, A_s1, A_s3, dof, dof), as a scipy CSC matrix; False the
    dense A."""
    plan = _plan_for(pde, opts, A_data, deg, imex_flag)
    N = len(np.asarray(A_data["element_local_index_D"]))
    B = deg ** np.asarray(A_data["element_local_index_D"]).reshape(N, -1).shape[1]
    dof = N * B
    if plan is None:
        if sparse:
            from scipy import sparse as sp
            return sp.csc_matrix((dof, dof))
        return np.zeros((dof, dof))
    blocks = plan.assemble().cpu().numpy()  # [q, c, r]
    if opts.use_connectivity:
        rp, cols = pde.connectivity
        rp = np.asarray(rp, dtype=np.int64)
        e = np.repeat(np.arange(N), np.diff(rp))
        j = np.asarray(cols, dtype=np.int64) - 1
    else:
        e = np.repeat(np.arange(N), N)
        j = np.tile(np.arange(N), N)
    r = np.arange(B)
    rows = (e[:, None, None] * B + r[None, None, :]).repeat(B, axis=1)  # [q, c, r] -> e*B + r
    colg = (j[:, None, None] * B + r[None, :, None]).repeat(B, axis=2)  # [q, c, r] -> j*B + c
    if sparse:
        from scipy import sparse as sp
        return sp.csc_matrix((blocks.ravel(), (rows.ravel(), colg.ravel())), shape=(dof, dof))
    A = np.zeros((dof, dof))
    np.add.at(A, (rows.ravel(), colg.ravel()), blocks.ravel())
    return A


def time_advance(pde, opts, A_data, f, t, dt, deg, hash_table=None, Vmax=None, Emax=None):
    """time_advance.m:1-34 dispatch for 'RK3' (default) and 'BE', with the
    PDE's sources and inhomogeneous Dirichlet terms at each stage's time."""
    import torch

    host = isinstance(f, np.ndarray)
    x = torch.from_numpy(np.ascontiguousarray(f, dtype=np.float64)).cuda() if host else f
    # sources and inhomogeneous boundary terms (source_vector.m + boundary_condition_vector.m)
    srcs = pde.__dict__.setdefault("_kmx_sources", {})
    skey = (_table_key(A_data), deg)
    if skey not in srcs:
        s0 = TS.SeparableSource(pde, deg, np.asarray(A_data["element_local_index_D"]))
        srcs[skey] = s0 if len(s0) else None
    src = srcs[skey]
    if opts.timestep_method == "BE":
        plan = _plan_for(pde, opts, A_data, deg, "N")
        rhs = src(t + dt) if src is not None else None  # backward_euler: s(t+dt) + bc(t+dt), time_advance.m:270-271
        y, _, _ = TS.backward_euler(plan, x, dt, rhs, tol=opts.extra.get("be_tol", 1e-12))
    else:
        y = TS.rk3_step(lambda v: apply_A(pde, opts, A_data, v, deg), x, t, dt, src)
    return y.cpu().numpy() if host else y


def matrix_plot_D(pde, opts, dimension):
    """matrix_plot_D.m:1 (no input nodes): the evaluation matrix and points of one
    dimension on opts.output_grid (only 'quadrature', the default), for
    wavelet_to_realspace.  Returns (Meval, nodes)."""
    return C.matrix_plot_D(opts.deg, dimension.lev, dimension.min, dimension.max, opts.output_grid)


def wavelet_to_realspace(pde, opts, Meval_D, f_wSpace, hash_table):
    """wavelet_to_realspace.m:1-79: sum over the elements of hash_table of
    kron_d Meval_D{d}(:, deg-column block of the element's 1D index) times the
    element's coefficients; the device transform is cached on the pde object
    per (Meval, table).  Returns prod(rows of Meval_D{d}) values, dim 1 slowest."""
    import torch

    key = (tuple(id(m) for m in Meval_D), id(hash_table), opts.deg)
    cache = pde.__dict__.setdefault("_kmx_realspace", {})
    if key not in cache:
        cache[key] = RealSpace([np.asarray(m) for m in Meval_D], hash_table.idx1d, opts.deg, index_base=1)
    host = isinstance(f_wSpace, np.ndarray)
    x = torch.from_numpy(np.ascontiguousarray(f_wSpace, dtype=np.float64)).cuda() if host else f_wSpace
    y = cache[key].apply(x)
    return y.cpu().numpy() if host else y


def build(pde_fn, opts):
    """Set-up of asgard_run_pde.m:20-70: element table, 1D matrices, A_data,
    initial condition.  Returns (pde, hash_table, A_data, f0, dt)."""
    pde = pde_fn(opts.lev)
    lev_vec = [d.lev for d in pde.dims]
    if opts.use_oldhash or opts.use_connectivity:
        table = G.hash_table_nd(lev_vec, opts.grid_type)
    else:
        table = G.hash_table_sparse_nd(lev_vec, opts.max_lev, opts.grid_type)
    if opts.use_connectivity:
        L = max(lev_vec)
        pde.connectivity = G.connect_nd(table, L, L, opts.grid_type)
    mats = C.get_coeff_mats(pde, opts.deg, 0.0, opts.max_lev if opts.max_lev_coeffs else None)
    pde.mats = C.restrict(mats, pde, opts.deg) if opts.max_lev_coeffs else mats
    A_data = G.global_matrix(table)
    f0 = C.md_eval_function(pde, opts.deg, table.idx1d, pde.initial_conditions, 0.0)
    dt = pde.set_dt(pde, opts.CFL)
    return pde, table, A_data, f0, dt


def asgard(pde_fn, **kw):
    """asgard.m:3 / asgard_run_pde.m time loop (RK3 or BE), FP64 state on the
    device.  Returns (err, fval) with err the RMS wavelet-space error against
    the analytic solution at the final time (asgard_run_pde.m:456-461) when
    the PDE has one, else None."""
    import torch

    opts = OPTS(**kw)
    pde, table, A_data, f0, dt = build(pde_fn, opts)
    f = torch.from_numpy(f0).cuda()
    t = 0.0
    for _ in range(opts.num_steps):
        f = time_advance(pde, opts, A_data, f, t, dt, opts.deg, table)
        t += dt
    fval = f.cpu().numpy()
    err = None
    if pde.solutions:
        exact = C.md_eval_function(pde, opts.deg, table.idx1d, pde.solutions, t)
        err = float(np.sqrt(np.mean((fval - exact) ** 2)))
    return err, fval
