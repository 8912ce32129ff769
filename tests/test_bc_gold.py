"""Inhomogeneous Dirichlet boundary terms (compute_boundary_condition.m,
boundary_condition_vector.m) and the RK3 trajectories that need them, against
the reference's gold fixtures (gold_compute_boundary.m,
gold_boundary_condition_vector.m, gold_time_advance.m), with the oracle's
apply_A in the steps.

The fixtures were written by earlier revisions of the reference: the trace
fixtures by a compute_boundary_condition without the dV argument (dV = 1 here),
the diffusion1 vector and trajectory fixtures with nu = 0.01 and diffusion2's
fluxes (DIV LF = +1, GRAD LF = -1; kmx.coeffs.diffusion1_gold).  The
trajectories are reproduced with the 1D operators built at each dimension's
level (OPTS.m:168 turns max_lev_coeffs off without adaptivity; the LDG chain
DIV * GRAD built at a finer level and cut is a different operator).
"""
import numpy as np
import pytest

import oracle as O
from kmx import coeffs as K
from kmx import grid as G


@pytest.mark.parametrize("deg,lev", [(2, 2), (4, 4), (5, 5)])
def test_compute_boundary_condition_gold(gold, deg, lev):
    pde = K.diffusion1_gold(lev)
    pt = pde.terms[0][0][1]  # the GRAD pterm: Dirichlet on both sides
    for side, fl in (("L", pt.BCL_fList), ("R", pt.BCR_fList)):
        got = K.compute_boundary_condition(pt.g, pt.dV, 0.0, lev, deg, 0.0, 1.0, fl[0], side, pde.params)
        want = gold[f"boundary_conditions/compute_diffusion1_bc{side}_{deg}d_{lev}l_0t_0dim_1p"].ravel()
        assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max()), side


@pytest.mark.parametrize("deg,lev", [(2, 2), (4, 4), (5, 5)])
def test_boundary_condition_vector_gold(gold, deg, lev):
    pde = K.diffusion1_gold(lev)
    tab = G.hash_table_sparse_nd([lev], 8, "SG")
    got = K.boundary_condition_vector(pde, deg, tab.idx1d, 0.0)
    want = gold[f"boundary_conditions/vector_diffusion1_l{lev}_d{deg}"].ravel()
    assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)
    # the current file's upwind DIV flux (LF = -1) does not reproduce it: the check sees flux signs
    bad = K.boundary_condition_vector(K.diffusion1(lev, nu=0.01), deg, tab.idx1d, 0.0)
    assert np.linalg.norm(bad - want) > 1e-2 * np.linalg.norm(want)


def rk3_with_bc(pde, deg, name, gold, steps=5, bc=True):
    """time_advance.m:151-204 with sources and boundary vectors at t, t + dt/2,
    t + dt (oracle apply); the relative error after each step."""
    lev_vec = [d.lev for d in pde.dims]
    tab = G.hash_table_sparse_nd(lev_vec, 8, "SG")
    mats = K.get_coeff_mats(pde, deg)  # at each dimension's level
    inp = O.ApplyInputs(mats, tab.idx1d, deg)

    def rhs(t):
        r = K.md_eval_function(pde, deg, tab.idx1d, pde.sources, t) if pde.sources else 0.0
        return r + (K.boundary_condition_vector(pde, deg, tab.idx1d, t) if bc else 0.0)

    f = K.md_eval_function(pde, deg, tab.idx1d, pde.initial_conditions, 0.0)
    dt = pde.set_dt(pde, 0.01)
    errs = []
    for i in range(steps):
        t = i * dt
        k1 = O.apply_A(inp, f) + rhs(t)
        y1 = f + dt * 0.5 * k1
        k2 = O.apply_A(inp, y1) + rhs(t + dt / 2)
        y2 = f + dt * (-1.0 * k1 + 2.0 * k2)
        k3 = O.apply_A(inp, y2) + rhs(t + dt)
        f = f + dt * (k1 / 6 + 2 * k2 / 3 + k3 / 6)
        g = gold[f"time_advance/{name}_t{i}"].ravel()
        errs.append(float(np.linalg.norm(f - g) / np.linalg.norm(g)))
    return errs


@pytest.mark.parametrize("name,pfn,lev,deg", [
    ("diffusion1_sg_l2_d2", K.diffusion1_gold, 2, 2),
    ("diffusion1_sg_l3_d3", K.diffusion1_gold, 3, 3),
    ("diffusion2_sg_l2_d2", K.diffusion2, 2, 2),
    ("diffusion2_sg_l3_d3", K.diffusion2, 3, 3),
])
def test_rk3_with_dirichlet_matches_gold(gold, name, pfn, lev, deg):
    errs = rk3_with_bc(pfn(lev), deg, name, gold)
    print(name, ["%.1e" % e for e in errs])
    assert max(errs) <= 1e-13, errs


def test_dropped_boundary_terms_are_caught(gold):
    errs = rk3_with_bc(K.diffusion2(2), 2, "diffusion2_sg_l2_d2", gold, steps=1, bc=False)
    assert errs[0] > 1e-4


@pytest.mark.parametrize("pfn,lev,deg", [(K.diffusion2, 3, 3), (lambda lev: K.diffusion_nd(3, lev), 2, 2)])
def test_boundary_terms_are_time_separable(pfn, lev, deg):
    """The spatial parts of these PDEs' boundary data do not depend on t, so the
    vector at any t is the t = 0 vectors times their time factors."""
    pde = pfn(lev)
    tab = G.hash_table_sparse_nd([d.lev for d in pde.dims], 8, "SG")
    terms = K.boundary_condition_terms(pde, deg, tab.idx1d, 0.0)
    assert len(terms) == 2 * pde.ndim
    for t in (0.013, 0.2):
        direct = K.boundary_condition_vector(pde, deg, tab.idx1d, t)
        sep = sum(f(t) * v for f, v in terms)
        assert np.linalg.norm(direct - sep) <= 1e-14 * np.linalg.norm(direct)
