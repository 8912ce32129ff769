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
            assert ls.residual() == res
        assert np.max(np.abs(ls.solve() - yk)) <= 1e-10 * max(1.0, np.abs(yk).max())


def test_gmres_holds_back_chunks_only_when_the_queued_steps_converge():
    """The device GMRES holds the next chunk back when the residual estimates,
    extrapolated at their last chunk's rate, reach the target within the queued
    steps ((depth - 1) chunks), and never for stalling or growing residuals."""
    from types import SimpleNamespace

    from kmx.timestep import DeviceGMRES

    g = SimpleNamespace(chunk=4, depth=2)
    held = DeviceGMRES._converges_in_flight
    rs = [10.0 ** -i for i in range(9)]  # a decade per step, at 1e-8 after 8 steps
    assert held(g, rs, 1e-11)      # 3 more steps: within the 4 queued
    assert held(g, rs, 1e-12)      # exactly 4
    assert not held(g, rs, 1e-13)  # 5: the next chunk is needed
    assert not held(g, [1.0, 1.0, 1.0, 1.0, 1.0], 1e-3)  # stalled
    assert not held(g, [1.0, 2.0, 3.0], 1e-3)            # growing
    assert not held(g, [1.0], 1e-3)                      # no step yet
    assert not held(SimpleNamespace(chunk=4, depth=1), rs, 1e-11)  # nothing queued
    assert not held(g, rs, 0.0)  # tol = 0: never converges by estimate
