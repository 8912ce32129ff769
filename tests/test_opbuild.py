"""The 1D operators built by the library (kmx_fmwt_matrix, kmx_quad_points,
kmx_sd_term_matrix): the transform against the reference's gold
operator_two_scale fixtures and the CPU restatement, the quadrature points
against lgwt.m's, and the device-built wavelet-space matrices of the PDEs here
against kmx.coeffs (coeff_matrix.m / pterm_coeff_matrix.m /
sd_term_coeff_matrix.m restated on the CPU and pinned by the gold coefficient,
transform and trajectory fixtures).

The device path solves M \\ L cell by cell in the real-space basis and rotates
once, F (M \\ L) F^T; the reference (and the CPU path) rotates first and solves,
(F M F^T) \\ (F L F^T) = F^-T (M \\ L) F^T.  The two agree exactly for an
orthogonal F; the transform built from the two-scale filter data is orthogonal
only to ||F F^T - I|| (1e-14 at deg 2-3, 6e-14 at deg 4, 2.3e-13 at deg 5-6), so
the bar is 1e-13 plus four times that defect, relative to the largest entry.
MASS g with LHS mass g is F F^T on the device and I on the CPU: the defect itself."""
import numpy as np
import pytest

from kmx import coeffs as K
from kmx import opbuild as OB


@pytest.mark.parametrize("key", ["2_2", "2_3", "2_6", "4_3", "5_5"])
def test_library_fmwt_matches_gold(gold, key):
    deg, lev = map(int, key.split("_"))
    F = OB.fmwt_matrix(deg, lev)
    assert np.max(np.abs(F - gold[f"transformations/operator_two_scale_{key}"])) <= 1e-14


@pytest.mark.parametrize("deg", [1, 2, 3, 4, 5, 6, 8])
def test_library_fmwt_matches_restatement(deg):
    for lev in range(0, 7):
        assert np.max(np.abs(OB.fmwt_matrix(deg, lev) - K.fmwt(deg, lev))) <= 1e-14, lev


@pytest.mark.parametrize("deg,lev", [(2, 3), (4, 5), (7, 2)])
def test_library_quadrature_points(deg, lev):
    x = OB.quad_points(deg, lev, -1.0, 2.0)
    qx, _ = K.lgwt(K.default_quad_number(deg), -1, 1)
    h = 3.0 / 2 ** lev
    want = np.array([((qx + 1) / 2 + i) * h - 1.0 for i in range(2 ** lev)])
    assert x.shape == want.shape and np.max(np.abs(x - want)) <= 1e-15


def _cases():
    """(name, dimension, pterm chain, params, deg, lev): every chain kind the PDEs
    here use -- periodic upwind DIV with variable g, the LDG chain with variable g,
    MASS with g != 1, Neumann / Dirichlet DIV * GRAD with LF, central flux, and a
    variable LHS mass."""
    out = []
    ad = K.advection_diffusion_varcoeff(3, 5, 0.02)
    out.append(("advdiff varcoeff DIV", ad.dims[0], ad.terms[0][0], ad.params, 3, 5))
    out.append(("advdiff varcoeff LDG", ad.dims[0], ad.terms[3][0], ad.params, 3, 5))
    d2 = K.diffusion2(4)
    out.append(("diffusion2 LDG D/N", d2.dims[0], d2.terms[0][0], d2.params, 4, 4))
    out.append(("diffusion2 MASS MASS", d2.dims[1], d2.terms[0][1], d2.params, 4, 4))
    c2 = K.continuity2(6)
    out.append(("continuity2 DIV upwind", c2.dims[1], c2.terms[1][1], c2.params, 2, 6))
    g = lambda x, p=None, t=None, dat=None: 2.0 + np.cos(3 * np.asarray(x, dtype=np.float64))  # noqa: E731
    out.append(("MASS g, LHS g", K.Dimension(-1.0, 1.0, 4), [K.MASS(g, LHS_mass_g=g)], {}, 5, 4))
    out.append(("central DIV with LHS g", K.Dimension(0.0, 2.0, 5),
                [K.DIV(K._const(1.0), 0.0, "P", "P", LHS_mass_g=g)], {}, 3, 5))
    out.append(("one cell", K.Dimension(0.0, 1.0, 0), [K.MASS()], {}, 3, 0))
    return out


CASES = _cases()


@pytest.mark.gpu
@pytest.mark.parametrize("i", range(len(CASES)), ids=[c[0] for c in CASES])
def test_device_sd_term_matches_restatement(kmx_gpu, i):
    name, dim, chain, params, deg, lev = CASES[i]
    want = K.sd_term_coeff_matrix(deg, 0.0, dim, chain, params, lev)
    got = OB.sd_term_matrix(deg, 0.0, dim, chain, params, lev).cpu().numpy()
    err = float(np.max(np.abs(got - want)) / np.max(np.abs(want)))
    F = K.fmwt(deg, lev)
    defect = float(np.max(np.abs(F @ F.T - np.eye(F.shape[0]))))
    print(f"{name}: deg {deg} lev {lev}: max |device - CPU| / max |CPU| = {err:.2e} "
          f"(transform defect {defect:.1e})")
    assert err <= 1e-13 + 4 * defect


@pytest.mark.gpu
def test_device_get_coeff_mats_drive_the_gold_trajectory(kmx_gpu, gold):
    """diffusion2 lev 2 deg 2 with device-built operators: the RK3 states (oracle
    apply, boundary terms from kmx.coeffs) stay within 1e-12 of the gold trajectory."""
    import oracle as O
    from kmx import grid as G

    pde = K.diffusion2(2)
    mats = [[m.cpu().numpy() for m in row] for row in OB.get_coeff_mats(pde, 2)]
    tab = G.hash_table_sparse_nd([2, 2], 8, "SG")
    inp = O.ApplyInputs(mats, tab.idx1d, 2)
    bc = lambda t: K.boundary_condition_vector(pde, 2, tab.idx1d, t)  # noqa: E731
    f = K.md_eval_function(pde, 2, tab.idx1d, pde.initial_conditions, 0.0)
    dt = pde.set_dt(pde, 0.01)
    for i in range(5):
        t = i * dt
        k1 = O.apply_A(inp, f) + bc(t)
        y1 = f + dt * 0.5 * k1
        k2 = O.apply_A(inp, y1) + bc(t + dt / 2)
        y2 = f + dt * (-k1 + 2 * k2)
        k3 = O.apply_A(inp, y2) + bc(t + dt)
        f = f + dt * (k1 / 6 + 2 * k2 / 3 + k3 / 6)
        g = gold[f"time_advance/diffusion2_sg_l2_d2_t{i}"].ravel()
        assert np.linalg.norm(f - g) <= 1e-12 * np.linalg.norm(g), i
