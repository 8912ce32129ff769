"""Many instances in one persistent launch (cfg4 shape, reduced size):
batched apply == per-plan applies == oracle, graph-replayed RK3 == eager RK3,
and the FP32 batched trajectory stays within 1e-6 of the FP64 one."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def system(kmx_gpu):
    from kmx import batch as BT

    inst = BT.make_instances(count=12, seed=3, lev2d=(4, 6), lev3d=(3, 4))
    return BT.BatchedSystem(inst)


def test_batched_apply_matches_plans_and_oracle(system):
    import torch

    x = torch.randn(system.dof, dtype=torch.float64, device="cuda")
    for prec, tol in (("fp64", 1e-13), ("fp32", 5e-6)):
        y = torch.empty_like(x)
        system.apply(x, y, prec)
        torch.cuda.synchronize()
        for i, (p, inst) in enumerate(zip(system.plans, system.instances)):
            xi = x[system.offsets[i]:system.offsets[i + 1]].contiguous()
            yi = y[system.offsets[i]:system.offsets[i + 1]].cpu().numpy()
            assert rel(yi, p.apply(xi, prec=prec).cpu().numpy()) <= tol
            if i < 3 and prec == "fp64":
                mats = [[inst["scales"][t] * inst["mats"][t][0]] + list(inst["mats"][t][1:])
                        for t in range(len(inst["mats"]))]
                yo = O.apply_A(O.ApplyInputs(mats, inst["table"].idx1d, inst["deg"]), xi.cpu().numpy(), nthreads=8)
                assert rel(yi, yo) <= tol


def test_graph_rk3_matches_eager_and_fp32_error(kmx_gpu):
    import torch

    from kmx import batch as BT

    inst = BT.make_instances(count=12, seed=4, lev2d=(4, 6), lev3d=(3, 4))
    s64 = BT.BatchedSystem(inst)
    s32 = BT.BatchedSystem(inst)
    g64 = BT.BatchRK3Graph(s64, prec="fp64")
    g32 = BT.BatchRK3Graph(s32, prec="fp32")
    fe = s64.f.clone()
    for _ in range(10):
        g64.step()
        g32.step()
        fe = g64.eager_step(fe)
        torch.cuda.synchronize()
        assert rel(s64.f.cpu().numpy(), fe.cpu().numpy()) <= 1e-14
    err = rel(s32.f.cpu().numpy(), s64.f.cpu().numpy())
    print(f"fp32 batched RK3 vs fp64 after 10 steps: {err:.3e}")
    assert err <= 1e-6
