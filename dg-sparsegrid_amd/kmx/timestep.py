"""Time integration around the operator apply, on the device.

* `SeparableSource`  -- md_eval_function.m sources are sums of separable terms
  prod_d f_{s,d}(x_d) * g_s(t); the spatial parts are projected once
  (forward_wavelet_transform.m, combine_dimensions_D.m) and only the scalar
  g_s(t) is evaluated per stage.
* `rk3_step`         -- time_advance.m:151-204 (Kutta RK3, a21=1/2, a31=-1,
  a32=2, b=[1/6, 2/3, 1/6]); three applies per step.
* `RK3Graph`         -- the same step captured once in a HIP graph and
  replayed (inputs: the state and the stage source coefficients).
* `backward_euler`   -- time_advance.m:267-350 solves (I - dt A) f1 = f0 +
  dt (s + bc) with a direct solve; here matrix-free restarted GMRES in FP64
  (`DeviceGMRES`: graph-captured Arnoldi steps) around mixed-precision inner
  solves, refined until the FP64 residual ||(I - dt A) f1 - b|| / ||b||
  reaches the requested tolerance.
"""
import math
import weakref

import numpy as np

from . import coeffs as _coeffs


def _torch():
    import torch

    return torch


class SeparableSource:
    """sum_s g_s(t) * S_s: the md_eval_function.m sources, S_s =
    combine_dimensions_D(forward transforms), and (boundary=True) the
    inhomogeneous Dirichlet terms of boundary_condition_vector.m, one pair
    (time factor, vector) per pterm side with Dirichlet data.  The boundary
    terms' spatial parts are taken at t = 0 and checked at a later time against
    the vector the reference would build there (time-separable boundary data,
    as every PDE here has); other data raise ValueError."""

    def __init__(self, pde, deg, idx1d, funcs=None, device="cuda", boundary=True, max_lev=None):
        torch = _torch()
        funcs = pde.sources if funcs is None else funcs
        nd = pde.ndim
        self.g = [md[nd] for md in funcs]
        spatial = []
        for md in funcs:
            fl = [_coeffs.forward_wavelet_transform(deg, pde.dims[d].lev, pde.dims[d].min, pde.dims[d].max, md[d],
                                                    lambda x, d=d: pde.dims[d].moment_dV(x), pde.params, 0.0)
                  for d in range(nd)]
            spatial.append(_coeffs.combine_dimensions_D(deg, fl, 1.0, np.asarray(idx1d)))
        self.num_boundary = 0
        if boundary:
            terms = _coeffs.boundary_condition_terms(pde, deg, idx1d, 0.0, max_lev)
            if terms:
                t1 = 0.37
                sep = sum(f(t1) * v for f, v in terms)
                direct = _coeffs.boundary_condition_vector(pde, deg, idx1d, t1, max_lev)
                if np.linalg.norm(sep - direct) > 1e-13 * max(np.linalg.norm(direct), 1e-300):
                    raise ValueError("boundary data are not time-separable")
                for f, v in terms:
                    self.g.append(lambda t, f=f: f(t))
                    spatial.append(v)
                self.num_boundary = len(terms)
        self.S_host = np.stack(spatial) if spatial else np.zeros((0, np.asarray(idx1d).shape[0] * deg ** nd))
        self.S = torch.from_numpy(self.S_host).to(device)
        self.dof = self.S_host.shape[1]

    def __len__(self):
        return len(self.g)

    def coeffs(self, t):
        return np.asarray([float(np.asarray(g(t))) for g in self.g], dtype=np.float64)

    def host(self, t):
        out = np.zeros(self.dof)
        for c, S in zip(self.coeffs(t), self.S_host):
            out = out + S * c
        return out

    def into(self, out, coef):
        """out = sum_s coef[s] * S_s (device; coef a device tensor), same order as host()."""
        out.zero_()
        for s in range(len(self.g)):
            out.add_(self.S[s] * coef[s])
        return out

    def __call__(self, t):
        torch = _torch()
        out = torch.zeros(self.dof, dtype=torch.float64, device=self.S.device)
        if len(self.g):
            self.into(out, torch.from_numpy(self.coeffs(t)).to(self.S.device))
        return out


A21, A31, A32 = 0.5, -1.0, 2.0
B1, B2, B3 = 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0
C2, C3 = 0.5, 1.0


def rk3_step(apply, f, t, dt, source=None, bc=None):
    """time_advance.m:151-204 (no LHS mass matrix).  apply(v) -> A v; source(t)
    and bc(t) return vectors (or None)."""

    def rhs(tt):
        r = 0.0
        if source is not None:
            r = source(tt)
        if bc is not None:
            r = r + bc(tt)
        return r

    s1, s2, s3 = rhs(t), rhs(t + C2 * dt), rhs(t + C3 * dt)
    k1 = apply(f) + s1
    y1 = f + dt * A21 * k1
    k2 = apply(y1) + s2
    y2 = f + dt * (A31 * k1 + A32 * k2)
    k3 = apply(y2) + s3
    return f + dt * (B1 * k1 + B2 * k2 + B3 * k3)


class RK3Graph:
    """One RK3 step of y' = A y + sum_s g_s(t) S_s captured in a HIP graph.

    Static device buffers hold the state and the 3 x n_src stage coefficients;
    `step(t)` writes the coefficients, replays, and returns the new state
    (in place).  The plan's kernels are launched on the capture stream."""

    def __init__(self, plan, f0, dt, source=None, prec="fp64", warmup=2):
        torch = _torch()
        self.plan, self.dt, self.source, self.prec = plan, float(dt), source, prec
        dev = f0.device
        n = plan.dof
        self.f = f0.clone()
        self.k = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3)]
        self.y = torch.empty(n, dtype=torch.float64, device=dev)
        self.s = torch.zeros(n, dtype=torch.float64, device=dev)
        ns = len(source) if source is not None else 0
        self.coef = torch.zeros((3, max(ns, 1)), dtype=torch.float64, device=dev)
        self.graph = None
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):  # warm the kernels (first-launch set-up stays out of the graph)
                self._body(dry=True)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body(dry=False)

    def _stage_rhs(self, i):
        if self.source is not None and len(self.source):
            return self.source.into(self.s, self.coef[i])
        return self.s.zero_()

    def _body(self, dry):
        dt = self.dt
        f, (k1, k2, k3), y = self.f, self.k, self.y
        src = self.f.clone() if dry else f  # a dry run must not advance the state
        self.plan.apply(src, k1, prec=self.prec)
        k1.add_(self._stage_rhs(0))
        torch = _torch()
        torch.add(src, k1, alpha=dt * A21, out=y)
        self.plan.apply(y, k2, prec=self.prec)
        k2.add_(self._stage_rhs(1))
        y.copy_((A31 * k1 + A32 * k2) * dt).add_(src)
        self.plan.apply(y, k3, prec=self.prec)
        k3.add_(self._stage_rhs(2))
        upd = (B1 * k1 + B2 * k2 + B3 * k3) * dt
        if not dry:
            f.add_(upd)

    def step(self, t):
        torch = _torch()
        if self.source is not None and len(self.source):
            c = np.stack([self.source.coeffs(t), self.source.coeffs(t + C2 * self.dt),
                          self.source.coeffs(t + C3 * self.dt)])
            self.coef.copy_(torch.from_numpy(c))
        self.graph.replay()
        return self.f


# ------------------------------------------------------------ implicit (BE)
class HessenbergLSQ:
    """The small least-squares problem of GMRES, min |beta e1 - H y|, with H the
    (k+1) x k Hessenberg matrix of the Arnoldi coefficients, kept as Givens
    rotations applied column by column (plain floats: the columns are short).
    add_column returns |beta e1 - H y| for the columns so far."""

    def __init__(self, beta, m):
        self.R = np.zeros((m, m))  # the rotated columns (upper triangle)
        self.cs, self.sn = [], []
        self.g = [float(beta)]
        self.k = 0

    def add_column(self, col):
        """col: H[0..k+1, k] as a list of k + 2 floats."""
        k, cs, sn, g = self.k, self.cs, self.sn, self.g
        col = list(col)
        for j in range(k):
            p, q, cj, sj = col[j], col[j + 1], cs[j], sn[j]
            col[j] = cj * p + sj * q
            col[j + 1] = cj * q - sj * p
        p, q = col[k], col[k + 1]
        den = math.hypot(p, q)
        ck, sk = (1.0, 0.0) if den == 0 else (p / den, q / den)
        cs.append(ck)
        sn.append(sk)
        col[k] = ck * p + sk * q
        self.R[: k + 1, k] = col[: k + 1]
        g.append(-sk * g[k])
        g[k] = ck * g[k]
        self.k = k + 1
        return abs(g[k + 1])

    def residual(self):
        """|beta e1 - H y| at the minimising y over the columns so far."""
        return abs(self.g[self.k])

    def solve(self):
        """y minimising |beta e1 - H y| over the columns so far."""
        k = self.k
        return np.linalg.solve(np.triu(self.R[:k, :k]), np.asarray(self.g[:k]))


class DeviceGMRES:
    """Restarted GMRES for M x = b with M v = s0 v + s1 (A v), FP64 vectors on the
    device.

    apply_into(x, y) writes y = A x (device, capturable).  Arnoldi step k is the
    product u = A v_k and kmx_arnoldi_step (w = s0 v_k + s1 u, classical
    Gram-Schmidt against v_0 .. v_k twice, v_{k+1} = w / |w|, the coefficients to
    the device); the steps k0 .. k0 + chunk - 1 of each chunk of a restart cycle
    are captured in one HIP graph of their own (k fixed in each, so nothing is
    indexed on the device), whose replay ends by copying the chunk's Hessenberg
    columns to pinned host memory.  The host applies the Givens rotations and
    checks the residual estimate once per chunk, while up to `depth` chunks'
    replays are queued; the next chunk is queued unless the estimates, extrapolated,
    say that the queued ones reach the tolerance.  Steps past convergence, whose
    basis vectors are not used, are the price of never leaving the device idle.  restart must be a multiple of
    chunk and at most KMX_ARNOLDI_MAX_K."""

    def __init__(self, apply_into, n, restart=40, device="cuda", chunk=4, shift=(0.0, 1.0), depth=2):
        torch = _torch()
        from . import _lib

        if restart % chunk or restart > _lib.ARNOLDI_MAX_K:
            raise ValueError("restart must be a multiple of chunk and at most KMX_ARNOLDI_MAX_K")
        self.apply_into, self.n, self.m, self.chunk = apply_into, int(n), int(restart), int(chunk)
        self.depth = max(1, int(depth))
        self.s0, self.s1 = float(shift[0]), float(shift[1])
        m = self.m
        self.V = torch.zeros((m + 1, self.n), dtype=torch.float64, device=device)
        self.u = torch.empty(self.n, dtype=torch.float64, device=device)
        self.Mx = torch.empty(self.n, dtype=torch.float64, device=device)
        self.work = torch.empty(_lib.arnoldi_work(self.n), dtype=torch.float64, device=device)
        self.hdev = torch.zeros((m, m + 2), dtype=torch.float64, device=device)
        self.hhost = torch.zeros((m, m + 2), dtype=torch.float64).pin_memory()
        self.ystage = torch.zeros(m, dtype=torch.float64).pin_memory()  # least-squares solution, host side
        self.graphs = [None] * (m // self.chunk)

    def _chunk(self, c0):
        from .plan import arnoldi_step

        for k in range(c0, c0 + self.chunk):
            self.apply_into(self.V[k], self.u)
            arnoldi_step(self.n, k, self.V, self.u, self.s0, self.s1, self.hdev[k], self.work)
        self.hhost[c0:c0 + self.chunk].copy_(self.hdev[c0:c0 + self.chunk], non_blocking=True)

    def _graph(self, ci):
        torch = _torch()
        if self.graphs[ci] is None:
            c0 = ci * self.chunk
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._chunk(c0)  # warm the kernels outside the capture (the same steps the replay repeats)
            torch.cuda.current_stream().wait_stream(side)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._chunk(c0)
            torch.cuda.synchronize()
            self.graphs[ci] = g
        return self.graphs[ci]

    def _converges_in_flight(self, rs, target):
        """Whether the residual estimates rs (one per step of the cycle so far),
        extrapolated at the rate of their last chunk of steps, reach target within
        the steps already queued ((depth - 1) chunks): then the next chunk is held
        back until a later chunk shows it is needed.  Timing only: a wrong guess
        costs one wait for the device, never a different result."""
        span = min(self.chunk, len(rs) - 1)
        if span <= 0 or target <= 0.0 or rs[-1] <= 0.0 or rs[-1 - span] <= 0.0:
            return False
        rate = (rs[-1] / rs[-1 - span]) ** (1.0 / span)
        if rate >= 1.0:
            return False
        return math.log(target / rs[-1]) / math.log(rate) <= (self.depth - 1) * self.chunk

    def matvec(self, x, out):
        """out = M x."""
        self.apply_into(x, out)
        out.mul_(self.s1).add_(x, alpha=self.s0)
        return out

    def solve(self, b, tol=1e-6, maxiter=400, bnorm=None):
        """x with ||b - M x|| <= tol ||b|| (estimate), from x = 0 (bnorm: ||b|| when
        the caller has it).  Returns (x, relative residual estimate, iterations
        used).  A cycle that meets tol by the least-squares estimate returns without
        forming b - M x (callers that need the true residual form it themselves)."""
        torch = _torch()
        m = self.m
        x = torch.zeros_like(b)
        bnorm = float(torch.linalg.vector_norm(b)) if bnorm is None else float(bnorm)
        if bnorm == 0.0:
            return x, 0.0, 0
        it, rel, r, beta = 0, 1.0, b, bnorm
        while it < maxiter:
            if r is not b:
                beta = float(torch.linalg.vector_norm(r))
            rel = beta / bnorm
            if rel <= tol:
                break
            torch.div(r, beta, out=self.V[0])
            ls = HessenbergLSQ(beta, m)
            nch = m // self.chunk
            done_ev = {}

            def launch(ci):
                self._graph(ci).replay()
                done_ev[ci] = torch.cuda.Event()
                done_ev[ci].record()

            for ci in range(min(self.depth, nch)):
                if it + ci * self.chunk < maxiter:
                    launch(ci)
            rs = [beta]  # the residual estimate after each step of this cycle
            for ci in range(nch):
                if ci not in done_ev:  # held back by the estimate below, needed after all
                    launch(ci)
                done_ev.pop(ci).synchronize()  # chunk ci done: its Hessenberg columns are on the host
                hv_all = self.hhost.numpy()
                done = False
                for c in range(self.chunk):
                    k = ci * self.chunk + c
                    res = ls.add_column(hv_all[k, : k + 2].tolist())
                    rs.append(res)
                    it += 1
                    if res / bnorm <= tol or hv_all[k, k + 1] == 0.0 or it >= maxiter:
                        done = True
                        break
                if done or it >= maxiter:
                    break
                nxt = ci + self.depth
                if nxt < nch and it + (self.depth - 1) * self.chunk < maxiter and not self._converges_in_flight(
                        rs, tol * bnorm):
                    # queued behind the chunks in flight, so the device never waits on the
                    # host's rotations (steps past convergence are wasted work whose basis
                    # vectors go unused)
                    launch(nxt)
            y = self.ystage[: ls.k]
            y.copy_(torch.from_numpy(ls.solve()))
            x.addmv_(self.V[: ls.k].t(), y.to(b.device, non_blocking=True))
            est = ls.residual() / bnorm
            if est <= tol or it >= maxiter:
                return x, est, it
            r = b - self.matvec(x, self.Mx)
        return x, rel, it


def gmres(matvec, b, x0=None, tol=1e-6, restart=40, maxiter=400):
    """Restarted GMRES for y = matvec(v) (allocating form) on the device: x0 is
    added to a DeviceGMRES solution of the residual system.  Returns (x, relative
    residual, iterations)."""
    torch = _torch()
    x0 = torch.zeros_like(b) if x0 is None else x0
    r0 = b - matvec(x0)
    solver = DeviceGMRES(lambda x, y: y.copy_(matvec(x)), b.numel(), restart, b.device)  # M = A
    d, _, it = solver.solve(r0, tol=tol * float(torch.linalg.vector_norm(b)) / max(
        float(torch.linalg.vector_norm(r0)), 1e-300), maxiter=maxiter)
    x = x0 + d
    return x, float(torch.linalg.vector_norm(b - matvec(x))) / max(float(torch.linalg.vector_norm(b)), 1e-300), it


_SOLVERS = weakref.WeakKeyDictionary()  # plan -> {(dt, inner precision, restart, n): DeviceGMRES}


def backward_euler(plan, f0, dt, rhs=None, tol=1e-10, inner_prec="fp32", inner_tol=1e-5, restart=40,
                   max_refine=20, max_inner=400):
    """Solve (I - dt A) f1 = f0 + dt*rhs  (time_advance.m:328-349's system).

    Iterative refinement: the FP64 residual r = b - (I - dt A) f1 (plain FP64
    apply) drives a correction solved by GMRES with `inner_prec` applies to
    relative accuracy inner_tol, until ||r|| / ||b|| <= tol.  The GMRES solver
    (its captured Arnoldi graphs) is kept per plan and step size for later steps.
    Returns (f1, history of FP64 relative residuals, inner iterations)."""
    torch = _torch()
    b = f0.clone() if rhs is None else f0 + dt * rhs
    bnorm = float(torch.linalg.vector_norm(b))
    af = torch.empty_like(b)  # A f, FP64
    r = torch.empty_like(b)

    key = (float(dt), inner_prec, int(restart), f0.numel())
    cache = _SOLVERS.setdefault(plan, {})
    solver = cache.get(key)
    if solver is None:
        pref = weakref.ref(plan)  # the cache must not keep its own key alive
        solver = DeviceGMRES(lambda x, y: pref().apply(x, y, prec=inner_prec), f0.numel(), restart, f0.device,
                             shift=(1.0, -float(dt)))
        cache[key] = solver
    f = f0.clone()  # f0 is a good first guess for small dt
    hist, inner = [], 0
    for _ in range(max_refine):
        plan.apply(f, af, prec="fp64")
        torch.sub(b, f, out=r).add_(af, alpha=dt)  # r = b - (f - dt A f)
        rnorm = float(torch.linalg.vector_norm(r))
        rel = rnorm / bnorm if bnorm else 0.0
        hist.append(rel)
        if rel <= tol:
            break
        d, _, its = solver.solve(r, tol=inner_tol, maxiter=max_inner, bnorm=rnorm)
        inner += its
        f.add_(d)
    return f, hist, inner
