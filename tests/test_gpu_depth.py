"""Trees of 16, 32 and 64 operator levels on the GPU (the traversal stack
holds one LDS slot per level, up to 64): left-deep Difference chains,
right-deep Union chains and mixed-operator unbalanced chains, with the scene
read from global memory and staged in LDS, with the cooperative kernel off,
taking every ray after one step (coop_when=1) and taking the frame's tail.

Exact mode equals the FP32 oracle run with stacks that cannot overflow
(2D + 1 entries) bit for bit, event counters included; fast mode agrees with
the FP64 oracle on unflagged rays with depth within 1e-4.  (The reference's
own 32-entry stacks overflow on chains of 40 levels and more; the oracle's
interval-classifier check at these depths is a CPU test.)"""
import numpy as np
import pytest
import torch

import csgray as C
import oracle as O
from common import setups
from deep_trees import SHAPES

pytestmark = pytest.mark.gpu

W, H = 96, 54
_ORACLE = {}


def _oracle(shape, depth, setup, bits):
    key = (shape, depth, setup, bits)
    if key not in _ORACLE:
        cam6, l3, ph8 = setups()[setup][4:7]
        _ORACLE[key] = O.OracleScene(SHAPES[shape](depth)).render(
            cam6, l3, ph8, W, H, bits=bits, stack_cap=2 * depth + 1, counters=bits == 32, degen=bits == 64)
    return _ORACLE[key]


@pytest.mark.parametrize("coop,when", [(-1, 0), (1, 1), (0, 0)], ids=["coop-off", "coop-every-ray", "coop-tail"])
@pytest.mark.parametrize("placement", [C.PLACE_GLOBAL, C.PLACE_LDS], ids=["global", "lds"])
@pytest.mark.parametrize("depth", [16, 32, 64])
@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_exact_matches_fp32_oracle_at_depth(gpu, shape, depth, placement, coop, when):
    print(f"\n[case] {shape} depth={depth} placement={placement} coop={coop} when={when}", flush=True)
    sname, cam, light, phong, *_ = setups()[1]
    s = C.Scene(SHAPES[shape](depth))
    assert s.info().op_depth == depth
    g = s.render(cam, light, phong, W, H, mode=C.MODE_EXACT, counters=True, placement=placement, coop=coop,
                 coop_when=when)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in g.items()}
    s.close()
    o = _oracle(shape, depth, 1, 32)
    assert (o["prim_id"] >= 0).sum() > 100
    assert np.array_equal(g["prim_id"], o["prim_id"])
    assert np.array_equal(g["flags"], o["flags"])
    assert np.array_equal(g["depth"].view(np.uint32), o["depth"].view(np.uint32))
    assert np.array_equal(g["counters"], o["counters"])
    assert np.abs(g["rgba"] - o["rgba"]).max() <= 2e-6


@pytest.mark.parametrize("setup", [0, 1])
@pytest.mark.parametrize("depth", [16, 32, 64])
@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_fast_matches_fp64_oracle_at_depth(gpu, shape, depth, setup):
    print(f"\n[case] fast {shape} depth={depth} setup={setup}", flush=True)
    sname, cam, light, phong, *_ = setups()[setup]
    s = C.Scene(SHAPES[shape](depth))
    g = s.render(cam, light, phong, W, H, mode=C.MODE_FAST)
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy() for k, v in g.items()}
    s.close()
    o = _oracle(shape, depth, setup, 64)
    use = o["degen"] == 0
    same = (g["prim_id"] == o["prim_id"]) & ((g["flags"] & 14) == (o["flags"] & 14))
    assert same[use].all()
    both = use & (o["prim_id"] >= 0) & same
    rel = np.abs(g["depth"][both].astype(np.float64) - o["depth"][both]) / o["depth"][both]
    assert rel.max(initial=0.0) <= 1e-4
