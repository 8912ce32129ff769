"""GPU wavelet -> real space transform (wavelet_to_realspace.m) against the
oracle's restatement (per element kron_multd with rectangular column blocks,
summed).  FP64 bar 1e-13 relative (2-norm)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def run_case(kmx, meval, idx, deg, seed):
    import torch

    rng = np.random.default_rng(seed)
    fw = rng.standard_normal(idx.shape[0] * deg ** len(meval))
    ref = O.wavelet_to_realspace(meval, idx, deg, fw)
    rs = kmx.RealSpace(meval, idx, deg)
    fr = rs.apply(torch.from_numpy(fw).cuda()).cpu().numpy()
    rs.destroy()
    return rel(fr, ref)


@pytest.mark.parametrize(
    "nd,lev,deg,oldhash",
    [(1, 8, 7, False), (2, 4, 3, False), (2, 5, 2, True), (3, 3, 4, False), (3, 4, 2, True), (4, 3, 2, False)],
)
def test_realspace_matches_oracle(kmx_gpu, nd, lev, deg, oldhash):
    rng = np.random.default_rng(nd * 100 + lev)
    h = O.hash_table_nd(nd, lev, "SG") if oldhash else O.hash_table_sparse_nd([lev] * nd, lev, "SG")
    n = deg * 2 ** lev
    npts = [max(3, n // (1 + d)) + d for d in range(nd)]  # rectangular, uneven per dim
    meval = [rng.standard_normal((npts[d], n)) for d in range(nd)]
    assert run_case(kmx_gpu, meval, h["idx1d"], deg, seed=lev) <= 1e-13


def test_realspace_adapted_grid(kmx_gpu):
    """Refined grid: level groups with absent cells."""
    from kmx import grid as G

    rng = np.random.default_rng(3)
    base = G.hash_table_sparse_nd([3, 3], 5, "SG")
    table = G.refine_random(base, 5, 0.6, rng)
    lev = int(table.lev.max())
    n = 3 * 2 ** lev
    meval = [rng.standard_normal((40, n)), rng.standard_normal((33, n))]
    assert run_case(kmx_gpu, meval, table.idx1d, 3, seed=4) <= 1e-13


def test_realspace_errors(kmx_gpu):
    rng = np.random.default_rng(0)
    h = O.hash_table_sparse_nd([3, 3], 3, "SG")
    meval = [rng.standard_normal((10, 2 * 8)), rng.standard_normal((10, 2 * 4))]  # dim 2 too narrow for lev 3
    with pytest.raises(kmx_gpu.KmxError):
        kmx_gpu.RealSpace(meval, h["idx1d"], 2)


def test_asgard_mirror_wavelet_to_realspace(kmx_gpu):
    """The reference-named entry (kmx.asgard.wavelet_to_realspace) on a
    continuity2 grid with host vectors."""
    from kmx import asgard as A
    from kmx import coeffs as K
    from kmx import grid as G

    opts = A.OPTS(lev=3, deg=2)
    pde = K.continuity2(3)
    table = G.hash_table_sparse_nd([3, 3], 8, "SG")
    rng = np.random.default_rng(8)
    meval = [rng.standard_normal((12, 16)), rng.standard_normal((9, 16))]
    fw = rng.standard_normal(table.num_elem * 4)
    fr = A.wavelet_to_realspace(pde, opts, meval, fw, table)
    assert rel(fr, O.wavelet_to_realspace(meval, table.idx1d, 2, fw)) <= 1e-13


# gold_wavelet_to_realspace.m:1,3 (lev, deg); max_lev 8 (OPTS.m:24); domains
# pdes/continuity1.m:16, continuity3.m:24-34
@pytest.mark.parametrize("name,deg,lev,doms", [("continuity_1", 7, 8, [(-1, 1)]),
                                               ("continuity_3", 4, 3, [(-1, 1), (-2, 2), (-3, 3)])])
def test_asgard_realspace_matches_gold(kmx_gpu, gold, name, deg, lev, doms):
    """The reference-named entries end to end on the GPU: matrix_plot_D's evaluation
    matrices, then wavelet_to_realspace of f = 2 * (0, 1, ..., N deg^d - 1)
    (gold_wavelet_to_realspace.m:36-42) against the reference's gold vector (the
    3D one stored as every 8th value)."""
    from kmx import asgard as A
    from kmx import coeffs as K
    from kmx import grid as G

    opts = A.OPTS(lev=lev, deg=deg)
    pde = K.PDE(dims=[K.Dimension(a, b, lev) for (a, b) in doms], terms=[])
    table = G.hash_table_sparse_nd([lev] * len(doms), 8, "SG")
    meval = [A.matrix_plot_D(pde, opts, dim)[0] for dim in pde.dims]
    fw = 2.0 * np.arange(table.num_elem * deg ** len(doms))
    fr = A.wavelet_to_realspace(pde, opts, meval, fw, table)
    key = f"transformations/wavelet_to_realspace_{name}"
    ref = gold[key].ravel()
    assert rel(fr[:: int(gold[key + "/stride"][0])], ref) <= 1e-13
