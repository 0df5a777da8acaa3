"""Writes events_cheese512_1080p.json: the oracle's mean per-ray event counts
(GoTo, Compute, box tests, sphere/cylinder/cube tests, hit fraction) for the
benchmark frame — cheese scene of 512 spheres (seed 512) at 1920x1080 from
(3, 2.5, 4).  bench.py turns them into algorithmic FP32 operations per ray."""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "cuda-csg-tree-raycasting_amd"))

import numpy as np  # noqa: E402

import csgray as C  # noqa: E402
import oracle as O  # noqa: E402
from csgray import scenes as S  # noqa: E402


def main():
    w, h = 1920, 1080
    text = S.cheese(512, seed=512)
    cam = C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    o = O.OracleScene(text)
    t0 = time.time()
    r = o.render(C.cam_tuple(cam), S.LIGHT_TOWARD, O.phong8(*S.PHONG), w, h, bits=32, counters=True)
    dt = time.time() - t0
    c = r["counters"].reshape(-1, 6).astype(np.float64)
    mean = dict(zip(O.COUNTER_NAMES, (float(v) for v in c.mean(0))))
    mean["hit"] = float((r["prim_id"] >= 0).mean())
    out = {"workload": "cheese 512 spheres seed 512, 1920x1080, eye (3,2.5,4) -> origin, vfov 60",
           "rays": w * h, "source": "oracle FP32 restatement, per-ray counters", "mean_events_per_ray": mean,
           "max_events_per_ray": dict(zip(O.COUNTER_NAMES, (int(v) for v in c.max(0))))}
    with open(os.path.join(HERE, "events_cheese512_1080p.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out), f"({dt:.1f} s)")


if __name__ == "__main__":
    main()
