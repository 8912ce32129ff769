"""Host-side pieces of the time-step harness that need no GPU."""
import numpy as np


def test_gmres_least_squares_update_matches_lstsq():
    """The host side of the device GMRES: Givens rotations applied column by column
    give the minimiser of |beta e1 - H y| and its residual, as numpy's lstsq does."""
    from kmx.timestep import HessenbergLSQ

    rng = np.random.default_rng(0)
    for m in (1, 5, 40):
        H = np.triu(rng.standard_normal((m + 1, m)), -1)
        beta = 2.5
        ls = HessenbergLSQ(beta, m)
        e1 = np.zeros(m + 1)
        e1[0] = beta
        for k in range(m):
            res = ls.add_column(H[: k + 2, k].tolist())
            yk = np.linalg.lstsq(H[: k + 2, : k + 1], e1[: k + 2], rcond=None)[0]
            assert abs(res - np.linalg.norm(H[: k + 2, : k + 1] @ yk - e1[: k + 2])) <= 1e-12 * beta
        assert np.max(np.abs(ls.solve() - yk)) <= 1e-10 * max(1.0, np.abs(yk).max())
