"""The headline workload, pinned on the CPU: the 64 adapted advection-diffusion
instances bench.py times (make_instances(count=64, seed=0)) span the stated
configuration -- 2D and 3D, base levels 4..8 in both, deg 2..4, all element pairs
-- and have the pair and dof totals the bench line and its profiles quote."""
import numpy as np


def test_headline_instances_span_the_stated_configuration():
    from kmx import batch as BT

    inst = BT.make_instances(count=64, seed=0)
    assert len(inst) == 64
    nds = {i["nd"] for i in inst}
    assert nds == {2, 3}
    for nd in (2, 3):
        levs = sorted({i["lev"] for i in inst if i["nd"] == nd})
        assert levs[0] >= 4 and levs[-1] <= 8, (nd, levs)
        assert levs[-1] >= 7, (nd, levs)  # the upper end of 4..8 is reached in both
    assert {i["deg"] for i in inst} <= {2, 3, 4} and len({i["deg"] for i in inst}) == 3
    pairs = sum(i["table"].num_elem ** 2 for i in inst)  # all element pairs
    dof = sum(i["table"].num_elem * i["deg"] ** i["nd"] for i in inst)
    assert pairs == 608196071
    assert dof == 3456277
    for i in inst:
        assert np.all((i["scales"] >= 0.5) & (i["scales"] <= 2.0))
        assert i["f0"].size == i["table"].num_elem * i["deg"] ** i["nd"]
