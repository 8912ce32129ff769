"""Many independent PDE instances advanced together.

Each instance has its own adapted sparse grid (levels, deg, element count),
its own 1D operators and term scales; all instances share one persistent
batched apply launch (kmx_batch_apply) per RK3 stage, and the whole RK3 step
(3 batched applies + the stage updates on the concatenated state) is captured
once in a HIP graph and replayed.

`make_instances` draws the instances from a seed: d in {2, 3}, deg in {2, 3, 4},
level uniform in [4, 8], term scales uniform in [0.5, 2], and an adapted grid obtained from the sparse grid by one refinement
sweep that keeps each method-1 child (get_child_elements.m) with probability
0.6.  Adapted grids use all element pairs (the reference does not combine
adaptivity with connectivity, OPTS.m:175-177).
"""
import numpy as np

from . import coeffs as C
from . import grid as G
from .plan import Batch, Plan
from .timestep import A21, A31, A32, B1, B2, B3

_MATS = {}


def _operator(nd, lev_build, deg, nu):
    key = (nd, lev_build, deg, nu)
    if key not in _MATS:
        pde = C.advection_diffusion_varcoeff(nd, lev_build, nu)
        _MATS[key] = C.get_coeff_mats(pde, deg, 0.0, None)
    return _MATS[key]


def make_instances(count=64, seed=0, keep_prob=0.6, lev2d=(4, 8), lev3d=(4, 8), degs=(2, 3, 4), nu=0.02,
                   cfl=0.01):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        nd = int(rng.choice([2, 3]))
        lo, hi = lev2d if nd == 2 else lev3d
        lev = int(rng.integers(lo, hi + 1))
        deg = int(rng.choice(degs))
        max_lev = lev + 2  # children up to lev + 1 (get_child_elements.m: lev + 1 < max_lev)
        base = G.hash_table_sparse_nd([lev] * nd, max_lev, "SG")
        table = G.refine_random(base, max_lev, keep_prob, rng)
        lev_build = int(table.lev.max())
        mats = _operator(nd, lev_build, deg, nu)
        scales = rng.uniform(0.5, 2.0, size=2 * nd)
        h = 2.0 / 2 ** lev_build
        x0 = rng.standard_normal(table.num_elem * deg ** nd)
        out.append(dict(nd=nd, lev=lev, deg=deg, table=table, mats=mats, scales=scales, dt=cfl * h * h, f0=x0))
    return out


class BatchedSystem:
    """Plans + concatenated device state for a list of instances.

    sweep: None (default) applies every instance pair by pair (the persistent
    launch); a tuple of precisions, e.g. ("fp32",), builds level-block plans and
    runs every instance that has a dimension-sweep path as one merged sweep
    (one launch per phase for all of them) in those precisions ("bf16" runs on
    the FP32 merged sweep's tables, with every instance separable)."""

    def __init__(self, instances, device="cuda", sweep=None):
        import torch

        self.instances = instances
        self.plans = []
        sweep = tuple(sweep or ())
        for inst in instances:
            p = Plan(inst["mats"], inst["table"].idx1d, inst["deg"], None, None, index_base=1, fp32=True,
                     level_blocks=bool(sweep))
            for t, s in enumerate(inst["scales"]):
                p.set_term_scale(t, float(s))
            self.plans.append(p)
        # blocks above 27 values run outside the persistent launch, each with
        # the kernel variant timed fastest on its own plan (the autotuner of a
        # level-block plan may also pick its own dimension sweep)
        for p in self.plans:
            if p.B > 27 and not (sweep and p.sweep()["available"]):
                p.autotune()
        self.batch = Batch(self.plans)
        for prec in sweep:
            self.batch.set_sweep(prec, True)
        self.sweep = {prec: self.batch.sweep(prec) for prec in ("fp64", "fp32", "bf16")}
        sizes = [p.dof for p in self.plans]
        self.offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
        self.dof = int(self.offsets[-1])
        self.f = torch.from_numpy(np.concatenate([inst["f0"] for inst in instances])).to(device)
        self.dt = torch.from_numpy(np.concatenate([np.full(p.dof, inst["dt"]) for p, inst in
                                                   zip(self.plans, instances)])).to(device)
        self.pairs = sum(p.P for p in self.plans)
        self.flops = sum(p.P * p.T * 2 * p.ndim * p.deg ** (p.ndim + 1) for p in self.plans)
        # flops one apply performs per precision (merged sweeps + the pair path of the rest)
        self.flops_by_prec = {}
        for prec in ("fp64", "fp32", "bf16"):
            on = self.sweep[prec]["active"]
            self.flops_by_prec[prec] = sum(
                p.sweep()["flops"] if on and p.sweep()["available"] else p.P * p.T * 2 * p.ndim * p.deg ** (p.ndim + 1)
                for p in self.plans)

    def views(self, v):
        return [v[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.plans))]

    def apply(self, x, y, prec="fp32"):
        self.batch.apply(self.views(x), self.views(y), prec=prec)
        return y

    def rk3(self, nsteps=1, prec="fp32", f=None):
        """nsteps fused RK3 steps of every instance (kmx_batch_rk3) on the
        concatenated state f (default: the system's own), in place."""
        f = self.f if f is None else f
        self.batch.rk3(self.views(f), [inst["dt"] for inst in self.instances], nsteps, prec=prec)
        return f


class BatchRK3Graph:
    """One RK3 step (time_advance.m:151-204, no sources) of every instance,
    with per-instance dt, captured in a HIP graph."""

    def __init__(self, system, prec="fp32", warmup=2):
        import torch

        self.s, self.prec = system, prec
        n = system.dof
        dev = system.f.device
        self.k = [torch.empty(n, dtype=torch.float64, device=dev) for _ in range(3)]
        self.y = torch.empty(n, dtype=torch.float64, device=dev)
        # per-dof stage weights dt * a_ij and dt * b_i: every stage update is then
        # one or two fused multiply-adds over the state
        dt = system.dt
        self.c21, self.c31, self.c32 = dt * A21, dt * A31, dt * A32
        self.b1, self.b2, self.b3 = dt * B1, dt * B2, dt * B3
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            for _ in range(warmup):
                self._body(dry=True)
        torch.cuda.current_stream().wait_stream(side)
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._body(dry=False)

    def _stages(self, f, k1, k2, k3, y):
        import torch

        s = self.s
        s.apply(f, k1, self.prec)
        torch.addcmul(f, self.c21, k1, out=y)
        s.apply(y, k2, self.prec)
        torch.addcmul(f, self.c31, k1, out=y).addcmul_(self.c32, k2)
        s.apply(y, k3, self.prec)

    def _body(self, dry):
        s = self.s
        f = s.f.clone() if dry else s.f
        k1, k2, k3 = self.k
        self._stages(f, k1, k2, k3, self.y)
        f.addcmul_(self.b1, k1).addcmul_(self.b2, k2).addcmul_(self.b3, k3)

    def step(self):
        self.graph.replay()
        return self.s.f

    def eager_step(self, f):
        """The same update without the graph (for checking replays)."""
        import torch

        k1, k2, k3 = (torch.empty_like(f) for _ in range(3))
        y = torch.empty_like(f)
        self._stages(f, k1, k2, k3, y)
        return f.clone().addcmul_(self.b1, k1).addcmul_(self.b2, k2).addcmul_(self.b3, k3)


def balanced_groups(plans, k, prec="fp32"):
    """The plans' indices in k groups of about equal work (sweep flops plus the
    stage launches' per-value traffic), heaviest first to the lightest group."""
    w = [p.sweep(prec)["flops"] + 4.0 * p.dof for p in plans]
    load, groups = [0.0] * k, [[] for _ in range(k)]
    for i in sorted(range(len(plans)), key=lambda i: -w[i]):
        j = min(range(k), key=lambda j: load[j])
        groups[j].append(i)
        load[j] += w[i]
    return [sorted(g) for g in groups if g]


class BatchRK3Fused:
    """steps_per_replay fused RK3 steps of every instance (kmx_batch_rk3: per
    stage one merged-sweep launch and one stage launch) captured in a HIP graph
    and replayed on the system's state.  Needs every instance on the merged
    separable sweep for prec.  split > 1 runs the instances as that many batches
    of about equal work on streams of their own inside the graph, so one group's
    stage launch (bound by HBM) can overlap another's op launch (bound by its
    texture path); each instance's arithmetic is the same, so the states are
    identical to split = 1.  autotune: time each batch's load batches per wave
    first (kmx_batch_autotune)."""

    def __init__(self, system, prec="fp32", steps_per_replay=1, split=1, autotune=False):
        import torch

        self.s, self.prec, self.n = system, prec, int(steps_per_replay)
        dts = [inst["dt"] for inst in system.instances]
        if split > 1:
            self.groups = balanced_groups(system.plans, int(split), prec)
            self.batches = []
            for g in self.groups:
                b = Batch([system.plans[i] for i in g])
                b.set_sweep(prec, True)
                self.batches.append(b)
        else:
            self.groups, self.batches = [list(range(len(system.plans)))], [system.batch]
        # each batch's load batches per wave timed on itself (kmx_batch_autotune)
        self.wave_batches = [b.autotune(prec) for b in self.batches] if autotune else None
        self.streams = [torch.cuda.Stream() for _ in self.batches]

        def run(f, n):
            views = system.views(f)
            cur = torch.cuda.current_stream()
            for b, g, st in zip(self.batches, self.groups, self.streams):
                st.wait_stream(cur)
                with torch.cuda.stream(st):
                    b.rk3([views[i] for i in g], [dts[i] for i in g], n, prec=prec)
            for st in self.streams:
                cur.wait_stream(st)

        scratch = system.f.clone()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            run(scratch, 1)  # allocates the stage workspaces outside the capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            run(system.f, self.n)

    def step(self):
        self.graph.replay()
        return self.s.f

