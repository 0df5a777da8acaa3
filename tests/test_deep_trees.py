"""Deep operator chains on the CPU: the library takes trees up to 64
operator levels and rejects deeper ones before touching a device; the
oracle, run with stacks large enough not to overflow (2D + 1 entries), agrees
with the FP64 interval classifier at those depths; and at the reference's
own 32-entry stacks it reports overflows on chains of 40 levels and more,
where the reference has no defined behaviour to match."""
import numpy as np
import pytest

import csgray as C
import oracle as O
from common import setups
from deep_trees import SHAPES, left_deep_difference, mixed_unbalanced, right_deep_union


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_depth_64_is_accepted_and_65_rejected(shape):
    gen = SHAPES[shape]
    s = C.Scene(gen(64), upload=False)
    assert s.info().op_depth == 64
    s.close()
    with pytest.raises(C.CsgError) as e:
        C.Scene(gen(65), upload=False)
    assert e.value.status == C.E_DEPTH


@pytest.mark.parametrize("depth", [16, 32, 64])
@pytest.mark.parametrize("gen", [left_deep_difference, right_deep_union,
                                 lambda d: mixed_unbalanced(d, cylinders=False)])
def test_oracle_at_depth_matches_interval_classifier(gen, depth):
    text = gen(depth)
    o = O.OracleScene(text)
    assert o.info()["op_depth"] == depth
    for setup in (0, 1):
        cam6 = setups()[setup][4]
        l3, ph8 = setups()[setup][5], setups()[setup][6]
        w, h = 96, 54
        k = o.render(cam6, l3, ph8, w, h, bits=64, degen=True, stack_cap=2 * depth + 1)
        assert ((k["degen"] & O.DG_OVERFLOW) == 0).all()
        c = o.classify(cam6, w, h)
        use = (c["status"] >= 0) & (k["degen"] == 0)
        assert use.sum() > 0.5 * w * h
        khit, chit = k["prim_id"] >= 0, c["status"] == 1
        same = (khit == chit) & np.where(khit & chit, k["prim_id"] == c["prim_id"], True)
        assert (~same & use).sum() == 0
        both = khit & chit & use
        assert (np.abs(k["depth"][both] - c["t"][both]) <= 1e-5 * c["t"][both]).all()
        assert both.sum() >= 30


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_reference_stack_overflows_on_deep_chains(shape):
    """The reference's 32-entry stacks hold the action stack's 2D + 1 entries
    of a full tree only up to D = 15; a chain pushes about one entry a level
    (its primitive children are tested in place), so chains of 16 levels stay
    within them and chains of 40 and 64 overflow them (the reference would
    drop pushes there; the library's stacks are sized to the tree)."""
    gen = SHAPES[shape]
    cam6, l3, ph8 = setups()[1][4], setups()[1][5], setups()[1][6]
    for depth, overflow in ((16, False), (40, True), (64, True)):
        r = O.OracleScene(gen(depth)).render(cam6, l3, ph8, 96, 54, bits=64, degen=True, stack_cap=32)
        assert bool(((r["degen"] & O.DG_OVERFLOW) != 0).any()) == overflow, depth
