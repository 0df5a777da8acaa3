"""Benchmark: shaded primary rays per second on one MI355X.

Workload (one step = one frame): the cheese scene — a cube of edge 2 at the
origin minus a balanced Union of 512 spheres (centres U[-1,1]^3, radii
U[0.05, 0.2], seed 512) — at 1920x1080, camera at (3, 2.5, 4) looking at the
origin with a 60 degree vertical fov, light toward (1,1,1), Phong ka 0.1,
kd 0.7, ks 0.2, n 32.  Every ray is traced through the CSG state machine,
shaded, and writes RGBA (16 B), depth (4 B), primitive id (4 B) and flags
(1 B).  Inputs are resident in HBM before timing starts.

Prints one JSON line (rank 0).  --gpus above 1 is refused: this build runs on
one GPU.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-csg-tree-raycasting_amd"))

FLOP_FIXTURE = os.path.join(ROOT, "tests", "golden", "events_cheese512_1080p.json")
# HBM bytes per render call from the FETCH_SIZE / WRITE_SIZE counter runs,
# scaled by the same counters on a copy of known size (line
# "hbm_bytes_per_launch: N" of this summary).
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_render_cheese512_1080p.md")
COPY_MIB = 256

# FP32 VALU operations per event, counted on the reference lines the event
# executes: ray generation (RaycastingKernels.cu:11-27, Ray.cuh:12-18), box test
# (:718-757), sphere (:135-181), cylinder (:202-336), cube (:375-434), a
# Compute decision (:620-653), and per hit the details (:183-200 etc.) plus
# Phong (:71-105).
FLOPS_RAYGEN, FLOPS_BOX, FLOPS_SPHERE, FLOPS_CYL, FLOPS_CUBE, FLOPS_COMPUTE = 50, 26, 37, 90, 59, 2
FLOPS_HIT = 25 + 87
PEAK_FP32_TFLOPS = 157.3


def flops_per_ray(ev):
    return (FLOPS_RAYGEN + FLOPS_BOX * ev["box"] + FLOPS_SPHERE * ev["sphere"] + FLOPS_CYL * ev["cylinder"]
            + FLOPS_CUBE * ev["cube"] + FLOPS_COMPUTE * ev["compute"] + FLOPS_HIT * ev["hit"])


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


FLAG_BITS = {"tie": 1, "box": 2, "root": 4, "cyl_end": 8, "cube_normal": 16, "tangent": 32, "overflow": 64}


def accuracy(g, o64):
    """GPU outputs `g` against the FP64 oracle `o64` over the same rays (flat
    arrays): hit-id agreement on all and on unflagged rays, inside/outside
    agreement, the flagged share with each flag bit's share, and relative
    depth error on agreeing hits (all, and unflagged)."""
    import numpy as np
    pid, opid = g["prim_id"].reshape(-1), o64["prim_id"].reshape(-1)
    dg = o64["degen"].reshape(-1)
    unf = dg == 0
    same = pid == opid
    hit = same & (opid >= 0)
    rel = np.abs(g["depth"].reshape(-1)[hit].astype(np.float64) - o64["depth"].reshape(-1)[hit]) / \
        np.abs(o64["depth"].reshape(-1)[hit].astype(np.float64))
    rel_unf = rel[unf[hit]]
    return {"rays": int(pid.size), "id_agree_all": round(float(same.mean()), 7),
            "id_agree_unflagged": round(float(same[unf].mean()), 7),
            "inside_outside_agree": round(float(((pid >= 0) == (opid >= 0)).mean()), 7),
            "flagged": round(float(1 - unf.mean()), 6),
            "flag_bits": {k: round(float(((dg & b) != 0).mean()), 6) for k, b in FLAG_BITS.items()},
            "depth_rel_max": float(f"{rel.max() if rel.size else 0:.3e}"),
            "depth_rel_p999": float(f"{np.quantile(rel, 0.999) if rel.size else 0:.3e}"),
            "depth_rel_max_unflagged": float(f"{rel_unf.max() if rel_unf.size else 0:.3e}")}


def cpu_baseline(cases, threads, gpu_c2=None, gpu_c5=None, mode="fast"):
    """The oracle (the CPU restatement of the reference path, scalar C++,
    -O2 -ffp-contract=off) timed on this host: FP32 (the reference's
    arithmetic) and FP64, on one core and on `threads` cores, for the
    Wikipedia 640x360 and cheese 1920x1080 frames in full and the 256-tree
    batch and the stress frame on a 1/64 ray subsample (every 64th row).
    The FP64 runs carry the degenerate-decision flagger and are the checker
    of the GPU's own cheese and stress outputs (accuracy block)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    legs, acc = [], {}

    def leg(config, bits, nthreads, rays, sample, fn):
        t0 = time.perf_counter()
        res = fn(nthreads)
        dt = time.perf_counter() - t0
        legs.append({"config": config, "precision": f"fp{bits}", "threads": nthreads, "rays": int(rays),
                     "seconds": round(dt, 3), "mrays_s": round(rays / dt / 1e6, 4), "sample": sample})
        return res

    c1, c2, c3, c5 = cases["c1"], cases["c2"], cases["c3"], cases["c5"]
    # C1, the reference's own CPU configuration, in full
    o1 = O.OracleScene(c1["text"])
    for bits in (32, 64):
        for n in (1, threads):
            leg("C1 Wikipedia 640x360", bits, n, 640 * 360, "full frame",
                lambda n, b=bits: o1.render(c1["cam6"], c1["light3"], c1["phong8"], 640, 360, bits=b, threads=n))
    # C2, the headline frame, in full on all threads; one core on every 16th row
    o2 = O.OracleScene(c2["text"])
    w, h = 1920, 1080
    leg("C2 cheese 1920x1080", 32, threads, w * h, "full frame",
        lambda n: o2.render(c2["cam6"], c2["light3"], c2["phong8"], w, h, bits=32, threads=n))
    r64 = leg("C2 cheese 1920x1080", 64, threads, w * h, "full frame (with the degenerate-decision flagger)",
              lambda n: o2.render(c2["cam6"], c2["light3"], c2["phong8"], w, h, bits=64, threads=n, degen=True))
    if gpu_c2 is not None:
        acc[f"C2 cheese 1920x1080 {mode}, every ray"] = accuracy(gpu_c2, r64)
    rows16 = len(range(0, h, 16)) * w
    for bits in (32, 64):
        leg("C2 cheese 1920x1080", bits, 1, rows16, "every 16th row",
            lambda n, b=bits: o2.render(c2["cam6"], c2["light3"], c2["phong8"], w, h, bits=b, threads=n,
                                        rows=(0, h, 16)))
    # C3, the 256-tree batch, every 64th row of every job
    scenes3 = [(O.OracleScene(t), cam6, size) for t, cam6, size in c3["jobs"]]
    rays3 = sum(len(range(0, size, 64)) * size for _, _, size in scenes3)
    for bits in (32, 64):
        leg("C3 256-tree batch", bits, threads, rays3, "every 64th row of every job (1/64)",
            lambda n, b=bits: [o.render(cam6, c3["light3"], c3["phong8"], size, size, bits=b, threads=n,
                                        rows=(0, size, 64)) for o, cam6, size in scenes3])
    # C5, the stress frame (4 jittered samples a pixel), every 64th row
    o5 = O.OracleScene(c5["text"])
    w5, h5, spp, seed = 3840, 2160, 4, c5["seed"]
    rays5 = len(range(0, h5, 64)) * w5 * spp
    leg("C5 stress 3840x2160x4", 32, threads, rays5, "every 64th row, all samples (1/64)",
        lambda n: o5.render_samples(c5["cam6"], c5["light3"], c5["phong8"], w5, h5, spp, seed, bits=32, threads=n,
                                    rows=(0, h5, 64)))
    r5 = leg("C5 stress 3840x2160x4", 64, threads, rays5,
             "every 64th row, all samples (1/64; with the degenerate-decision flagger)",
             lambda n: o5.render_samples(c5["cam6"], c5["light3"], c5["phong8"], w5, h5, spp, seed, bits=64,
                                         threads=n, rows=(0, h5, 64), degen=True))
    if gpu_c5 is not None:
        acc["C5 stress 3840x2160x4 fast, every 64th row"] = accuracy(
            gpu_c5, {k: v[0:h5:64] for k, v in r5.items()})
    headline = next(l for l in legs if l["config"].startswith("C2") and l["precision"] == "fp32"
                    and l["threads"] == threads)
    single = next(l for l in legs if l["config"].startswith("C2") and l["precision"] == "fp32" and l["threads"] == 1)
    base = {"value": headline["mrays_s"], "unit": "Mrays/s", "cores": threads, "kind": "port",
            "single_core_mrays_s": single["mrays_s"], "cpu_model": _cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"the headline 1920x1080 frame in full ({w * h} rays), FP32 oracle, {threads} threads, "
                      f"{headline['seconds']} s; single core: every 16th row ({rows16} rays)",
            "legs": legs}
    return base, acc


def _time_ms(torch, fn, reps):
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def bench_batched(torch, C, S, reps, coop):
    """The 256-tree mix: per-scene launches vs one batched launch vs the
    captured (hipGraph) batch after autotuning."""
    mix = S.batch_mix(256, seed=256)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    specs = []
    for text, eye, size in mix:
        sc = C.Scene(text)
        specs.append((sc, C.camera_look_at(eye, (0, 0, 0), 60), light, ph, size, size, C.alloc_outputs(size, size)))
    rays = sum(sp[4] * sp[5] for sp in specs)

    def plain():
        for sc, cam, l, p, w, h, o in specs:
            sc.render(cam, l, p, w, h, outputs=o, **coop)

    jobs = C.make_jobs(specs)
    plain_ms = _time_ms(torch, plain, reps)
    batch_ms = _time_ms(torch, lambda: C.render_batch(jobs, **coop), reps)
    batch = C.Batch(jobs, **coop)
    tune = batch.autotune(reps=2)
    graph_ms = _time_ms(torch, batch.launch, reps)
    batch.close()
    return {"workload": "256 random trees (2-1024 leaves, balanced, ops uniform), 128^2..1024^2, orbit cameras",
            "rays": rays, "plain_per_scene_mrays_s": round(rays / plain_ms / 1e3, 1),
            "batched_mrays_s": round(rays / batch_ms / 1e3, 1), "graph_mrays_s": round(rays / graph_ms / 1e3, 1),
            "batched_vs_plain": round(plain_ms / min(batch_ms, graph_ms), 2),
            "autotune": {"block": tune.block, "tile_walk": tune.tile_walk, "heaviest_first": tune.heaviest_first,
                         "depth_split": tune.depth_split},
            "ms": {"plain": round(plain_ms, 2), "batched": round(batch_ms, 2), "graph": round(graph_ms, 2)}}


def bench_wikipedia(torch, C, S, reps, coop):
    """The Wikipedia scene (Difference of an Intersection of cube and sphere
    with a Union of three cylinders) at 640x360 from (3, 2.5, 4): the
    reference's own CPU-runnable configuration, on the GPU."""
    w, h = 640, 360
    scene = C.Scene(S.WIKIPEDIA)
    cam = C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    outs = C.alloc_outputs(w, h)
    ms = _time_ms(torch, lambda: scene.render(cam, light, ph, w, h, outputs=outs, **coop), reps)
    return {"workload": "Wikipedia scene, 640x360, 1 ray/pixel", "rays": w * h, "ms": round(ms, 4),
            "mrays_s": round(w * h / ms / 1e3, 1)}


def bench_stress(torch, C, S, reps, coop):
    """Cube minus a balanced Union of 4095 small spheres / cylinders (8191
    nodes, 13 operator levels, scene in global memory), 3840x2160 with 4
    jittered samples per pixel: the plain launch against the captured batch
    of the same frame after autotuning its tiling; both must give the same
    bits."""
    import numpy as np
    w, h, spp, seed = 3840, 2160, 4, 4095
    scene = C.Scene(S.stress(4095, seed=4095))
    cam = C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    outs = C.alloc_outputs(w, h, spp)
    before = C.coop_rays()
    plain_ms = _time_ms(torch, lambda: scene.render(cam, light, ph, w, h, outputs=outs, samples=spp, seed=seed,
                                                    **coop), reps)
    handed = (C.coop_rays() - before) / (reps + 1)
    plain_img = {k: v.cpu().numpy() for k, v in outs.items()}
    bouts = C.alloc_outputs(w, h, spp)
    jobs = C.make_jobs([(scene, cam, light, ph, w, h, bouts)])
    batch = C.Batch(jobs, samples=spp, seed=seed, **coop)
    tune = batch.autotune(reps=1)
    graph_ms = _time_ms(torch, batch.launch, reps)
    same = all(np.array_equal(plain_img[k].view(np.uint8), bouts[k].cpu().numpy().view(np.uint8))
               for k in ("rgba", "depth", "prim_id", "flags"))
    rows64 = {k: plain_img[k][0:h:64].copy() for k in ("depth", "prim_id", "flags")}
    del plain_img
    batch.close()
    rays = w * h * spp
    info = scene.info()
    plan = scene.plan(coop=coop["coop"])
    return {"workload": "stress: cube minus Union of 4095 spheres/cylinders (8191 nodes), 3840x2160, 4 spp",
            "rays": rays, "nodes": info.nodes, "op_depth": info.op_depth, "scene_in_lds": bool(plan.scene_in_lds),
            "plain_ms": round(plain_ms, 2), "plain_mrays_s": round(rays / plain_ms / 1e3, 1),
            "graph_ms": round(graph_ms, 2), "graph_mrays_s": round(rays / graph_ms / 1e3, 1),
            "autotune": {"block": tune.block, "tile_walk": tune.tile_walk, "heaviest_first": tune.heaviest_first,
                         "depth_split": tune.depth_split},
            "coop_rays_per_frame": round(handed), "plain_equals_batched": same, "_gpu_rows64": rows64}


def bench_serving(C, S, lambdas, seconds, coop):
    """Poisson arrivals of (tree, orbit camera, size) requests against 64
    resident trees; latency = submit -> result ready."""
    import threading

    import numpy as np
    light, ph = C.light(S.LIGHT_TOWARD), C.phong(*S.PHONG)
    sizes = [(256, 256), (512, 512), (1024, 1024), (1920, 1080)]
    out = []
    for lam in lambdas:
        server = C.Server(budget_bytes=16 << 30, streams=3, max_batch_jobs=256, **coop)
        mix = S.batch_mix(64, seed=64)
        scenes = [C.Scene(t) for t, _, _ in mix]
        ids = [server.add_scene(s) for s in scenes]
        rng = np.random.default_rng(int(lam))
        n = max(1, int(lam * seconds))
        gaps = rng.exponential(1.0 / lam, size=n)
        arrivals = np.cumsum(gaps)
        reqs = []
        for _ in range(n):
            k = int(rng.integers(0, len(ids)))
            w, h = sizes[int(rng.integers(0, len(sizes)))]
            reqs.append((ids[k], C.camera_look_at(S.orbit_eye(rng, 4.0, 85.0), (0, 0, 0), 60), w, h))
        tickets = []
        stats = []
        lock = threading.Lock()
        ready = threading.Semaphore(0)

        def reaper():
            for i in range(n):
                ready.acquire()
                with lock:
                    t = tickets[i]
                st = server.wait(t)
                stats.append((st.submit_s, st.done_s, st.status))
                server.release(t)

        th = threading.Thread(target=reaper)
        th.start()
        t0 = time.perf_counter()
        for i, (sid, cam, w, h) in enumerate(reqs):
            # sleep until just before the arrival, then spin: a pure spin
            # holds the interpreter lock and starves the thread that
            # releases finished outputs back to the server's budget
            gap = arrivals[i] - (time.perf_counter() - t0)
            if gap > 3e-4:
                time.sleep(gap - 2e-4)
            while time.perf_counter() - t0 < arrivals[i]:
                pass
            t = server.submit(sid, cam, light, ph, w, h)
            with lock:
                tickets.append(t)
            ready.release()
        th.join()
        lat = np.array([(d - s) * 1e3 for s, d, _ in stats])
        span = max(d for _, d, _ in stats) - min(s for s, _, _ in stats)
        rays = sum(w * h for _, _, w, h in reqs)
        ss = server.stats()
        server.close()
        out.append({"lambda_per_s": lam, "requests": n, "peak_gb_in_use": round(ss.peak_bytes_in_use / 2**30, 2),
                    "p50_ms": round(float(np.percentile(lat, 50)), 2),
                    "p99_ms": round(float(np.percentile(lat, 99)), 2), "mrays_s": round(rays / span / 1e6, 1),
                    "offered_mrays_s": round(rays / arrivals[-1] / 1e6, 1), "batches": ss.batches,
                    "failed": int(sum(1 for *_, st in stats if st != 0))})
    return out


def calibrate_copy():
    """A plain device copy of COPY_MIB MiB, five times: each dispatch reads and
    writes exactly COPY_MIB MiB, which scales the FETCH_SIZE / WRITE_SIZE
    counters of the render's profile."""
    import torch
    n = COPY_MIB * (1 << 20) // 4
    a = torch.arange(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_mib": COPY_MIB, "dispatches": 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", choices=("fast", "exact"), default="fast")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the batched and serving legs")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--coop-budget", type=int, default=0,
                    help="steps before a ray may go to the cooperative kernel (0 default, <0 never)")
    ap.add_argument("--coop-when", type=int, choices=(0, 1), default=0,
                    help="0: hand rays over only once every tile is taken; 1: as soon as the budget is spent")
    ap.add_argument("--coop-grace", type=int, default=0,
                    help="tail hand-over waits this percent of the tile hand-out time (0 default, <0 none)")
    ap.add_argument("--only", choices=("plain", "wikipedia", "batched", "serving", "stress"), default=None,
                    help="run one leg (plain still sets the headline value)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help=f"only copy a {COPY_MIB} MiB buffer 5 times (reference for the HBM byte counters)")
    args = ap.parse_args()
    coop = {"coop": args.coop_budget, "coop_when": args.coop_when, "coop_grace": args.coop_grace}
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    import torch
    import csgray as C
    from csgray import scenes as S

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    w, h = 1920, 1080
    text = S.cheese(512, seed=512)
    cam = C.camera_look_at(S.VIEW_EYE, S.VIEW_TARGET, S.VIEW_VFOV_DEG)
    light = C.light(S.LIGHT_TOWARD)
    ph = C.phong(*S.PHONG)
    mode = C.MODE_FAST if args.mode == "fast" else C.MODE_EXACT
    scene = C.Scene(text)
    plan = scene.plan(mode=mode, coop=args.coop_budget)
    outs = {
        "rgba": torch.empty((h, w, 4), dtype=torch.float32, device=dev),
        "depth": torch.empty((h, w), dtype=torch.float32, device=dev),
        "prim_id": torch.empty((h, w), dtype=torch.int32, device=dev),
        "flags": torch.empty((h, w), dtype=torch.uint8, device=dev),
    }
    stream = torch.cuda.current_stream(dev)
    for _ in range(args.warmup):
        scene.render(cam, light, ph, w, h, mode=mode, outputs=outs, **coop)
    torch.cuda.synchronize()

    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record(stream)
        scene.render(cam, light, ph, w, h, mode=mode, outputs=outs, **coop)
        ends[i].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = sorted(s.elapsed_time(e) for s, e in zip(starts, ends))  # ms, kernel-only (one launch per step)
    total_ms = starts[0].elapsed_time(ends[-1])
    rays = w * h
    value = rays * args.steps / (total_ms / 1e3) / 1e6
    p50 = per[len(per) // 2]
    p99 = per[min(len(per) - 1, int(round(0.99 * (len(per) - 1))))]
    avg_kernel_ms = sum(per) / len(per)

    roofline = None
    if os.path.exists(FLOP_FIXTURE):
        with open(FLOP_FIXTURE) as f:
            ev = json.load(f)["mean_events_per_ray"]
        fpr = flops_per_ray(ev)
        achieved = fpr * rays / (avg_kernel_ms / 1e3) / 1e12
        traffic = None
        if os.path.exists(TRAFFIC_SUMMARY):
            with open(TRAFFIC_SUMMARY) as f:
                for ln in f:
                    if ln.startswith("hbm_bytes_per_launch:"):
                        traffic = int(ln.split(":")[1])
        roofline = {"bound": "valu_fp32", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS,
                    "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": traffic,
                    "flops_per_ray": round(fpr, 1), "kernel": "render call (RenderKernel + CoopKernel)",
                    "avg_kernel_ms": round(avg_kernel_ms, 4)}

    extra = {}
    if not args.no_extra:
        if args.only in (None, "batched"):
            extra["batched"] = bench_batched(torch, C, S, 3, coop)
        if args.only in (None, "wikipedia"):
            extra["wikipedia"] = bench_wikipedia(torch, C, S, 200, coop)
        if args.only in (None, "stress"):
            extra["stress"] = bench_stress(torch, C, S, 2, coop)
        if args.only in (None, "serving"):
            extra["serving"] = bench_serving(C, S, (200, 1000, 5000), args.serve_seconds, coop)

    gpu_c5 = extra.get("stress", {}).pop("_gpu_rows64", None)
    cpu, acc = None, None
    if not args.no_cpu:
        import numpy as np  # noqa: F401
        gpu_c2 = {k: outs[k].cpu().numpy() for k in ("depth", "prim_id", "flags")}
        ph8 = S.PHONG + S.BACKGROUND
        view6 = C.cam_tuple(cam)
        cases = {
            "c1": {"text": S.WIKIPEDIA, "cam6": view6, "light3": S.LIGHT_TOWARD, "phong8": ph8},
            "c2": {"text": text, "cam6": view6, "light3": S.LIGHT_TOWARD, "phong8": ph8},
            "c3": {"jobs": [(t, C.cam_tuple(C.camera_look_at(eye, (0, 0, 0), 60)), size)
                            for t, eye, size in S.batch_mix(256, seed=256)],
                   "light3": S.LIGHT_TOWARD, "phong8": ph8},
            "c5": {"text": S.stress(4095, seed=4095), "cam6": view6, "light3": S.LIGHT_TOWARD, "phong8": ph8,
                   "seed": 4095},
        }
        cpu, acc = cpu_baseline(cases, min(os.cpu_count() or 1, 16), gpu_c2, gpu_c5, args.mode)

    line = {
        "metric": "Mrays/s shaded primary rays (plain vs batched) + p50/p99 frame latency",
        "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scene)",
        "config": {"workload": "plain: cheese 512 spheres, 1920x1080, 1 ray/pixel", "mode": args.mode,
                   "rays_per_step": rays, "p50_ms": round(p50, 4), "p99_ms": round(p99, 4),
                   "wall_s": round(wall, 3), "block": plan.block, "grid": plan.grid,
                   "scene_in_lds": bool(plan.scene_in_lds), "lds_bytes": plan.lds_bytes,
                   "coop_budget": plan.coop_budget, "coop_when": args.coop_when,
                   "coop_grace": args.coop_grace if args.coop_grace else plan.coop_grace,
                   "coop_grid": plan.coop_grid, "coop_block": plan.coop_block},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "accuracy": acc,
    }
    line.update(extra)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
