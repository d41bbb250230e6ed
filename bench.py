"""Benchmark: RK4 ray-steps/s of the FP64 generic-metric path on one MI355X.

Workload (one "step" = one pass over this batch, inputs generated on the
device from the camera, outputs left in HBM):
  * Alcubierre R=2, σ=8 (thickness 1/8), v_s=0.5, camera at x=-10 (literal:
    the reference camera looks toward −x, away from the bubble), 1024²
  * the same bubble seen from x=+10 (the bubble in view), max 40,000 steps
  * Morris–Thorne b0=1, camera at l=15, 1024², max 20,000 steps (reference
    defaults: nearly every ray runs out of steps)
all at 60° FOV, step 0.05, FD step 1e-5, stars on. The curvature field
(Alcubierre T_μν on a 128³ grid over [-4,4]³, Morris–Thorne on a 4096-point
line) is timed once after the ray workload and reported as points/s.

Timing uses HIP events (torch.cuda.Event) recorded on the stream every
kernel is launched on. The CPU baseline is the oracle (the reference
algorithm restated in C, -O3, no FMA) on a bounded sample of the same rays,
timed on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6

# Algorithmic FP64 operations per RK4 ray-step of the kernels' formulation
# (reference operation order with structural zeros removed; add/sub/mul/div,
# sqrt and each transcendental count 1).
OPS_PER_STEP = {"morris_thorne": 418, "alcubierre_flat": 115, "schwarzschild": 272}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    return ap.parse_args()


def workload():
    import geon_amd as geon
    alc = geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5, ship_x=0.0)
    mt = geon.MorrisThorne(throat_radius=1.0)
    return [
        ("alcubierre_x-10", alc, geon.Camera((0.0, -10.0, 0.0, 0.0), 60.0, 1024, 1024),
         geon.RenderConfig(step=0.05, max_steps=20_000)),
        ("alcubierre_x+10", alc, geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 1024, 1024),
         geon.RenderConfig(step=0.05, max_steps=40_000)),
        ("morris_thorne_l15", mt, geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, 1024, 1024),
         geon.RenderConfig(step=0.05, max_steps=20_000)),
    ]


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import numpy as np
    import torch

    import geon_amd as geon

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    ctx = geon.Context(0)
    jobs = workload()
    bufs = {}
    for name, m, cam, cfg in jobs:
        n = cam.width * cam.height
        bufs[name] = (torch.empty(n * 3, dtype=torch.float32, device=dev),
                      torch.empty(n, dtype=torch.int32, device=dev),
                      torch.empty(n * 3, dtype=torch.float64, device=dev))

    def run_once(want_stats):
        out = {}
        for name, m, cam, cfg in jobs:
            rgb, cls, d = bufs[name]
            out[name] = geon.render_dev(m, cam, cfg, rgb.data_ptr(), cls.data_ptr(), d.data_ptr(),
                                        stream_ptr=stream.cuda_stream, want_stats=want_stats,
                                        ctx=ctx)
        return out

    # warmup (also records the deterministic per-render ray-step counts)
    stats = None
    for _ in range(max(1, args.warmup)):
        stats = run_once(True)
    torch.cuda.synchronize()
    per_render_ms = {}
    for name, m, cam, cfg in jobs:   # per-render device time, for the breakdown
        rgb, cls, d = bufs[name]
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(stream):
            e0.record(stream)
        geon.render_dev(m, cam, cfg, rgb.data_ptr(), cls.data_ptr(), d.data_ptr(),
                        stream_ptr=stream.cuda_stream, want_stats=False, ctx=ctx)
        e1.record(stream)
        e1.synchronize()
        per_render_ms[name] = e0.elapsed_time(e1)

    ray_steps_per_step = sum(int(s["ray_steps"]) for s in stats.values())
    torch.cuda.synchronize()
    t_start = torch.cuda.Event(enable_timing=True)
    t_stop = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    t_start.record(stream)
    for _ in range(args.steps):
        run_once(False)
    t_stop.record(stream)
    t_stop.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_total = t_start.elapsed_time(t_stop)
    ms_per_step = ms_total / args.steps
    value = ray_steps_per_step * args.steps / (ms_total * 1e-3)

    # check the timed renders still produce the warmup statistics
    final = run_once(True)
    for k in stats:
        assert final[k]["ray_steps"] == stats[k]["ray_steps"], k

    # dominant kernel: the Morris–Thorne render (every step is a full FD step)
    mt_ms = final["morris_thorne_l15"]["kernel_ms"]
    mt_steps = int(final["morris_thorne_l15"]["ray_steps"])
    achieved = mt_steps * OPS_PER_STEP["morris_thorne"] / (mt_ms * 1e-3) / 1e12
    roofline = {"bound": "valu_fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4), "traffic": None,
                "kernel": "geon_render_rays_mt",
                "ops_per_ray_step": OPS_PER_STEP["morris_thorne"]}

    # curvature field (timed separately, device events)
    curv = {}
    g = torch.linspace(-4.0, 4.0, 128, dtype=torch.float64, device=dev)
    X, Y, Z = torch.meshgrid(g, g, g, indexing="ij")
    xs = torch.stack([torch.zeros_like(X).ravel(), X.ravel(), Y.ravel(), Z.ravel()]).contiguous()
    out = torch.empty(xs.shape[1] * 16, dtype=torch.float64, device=dev)
    lline = torch.linspace(-4.0, 4.0, 4096, dtype=torch.float64, device=dev)
    xl = torch.stack([torch.zeros_like(lline), lline, torch.full_like(lline, math.pi / 2),
                      torch.zeros_like(lline)]).contiguous()
    outl = torch.empty(4096 * 16, dtype=torch.float64, device=dev)
    alc = jobs[0][1]
    mt = jobs[2][1]
    for label, m, x, o, n in [("alcubierre_128^3", alc, xs, out, xs.shape[1]),
                              ("morris_thorne_4096", mt, xl, outl, 4096)]:
        for rep in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            geon._abi.check(geon.lib().geon_curvature_dev(
                ctx.handle, geon.C.byref(geon._cmetric(m)), geon.C.c_void_p(x.data_ptr()), n,
                geon.CURV_STRESS, geon.C.c_void_p(o.data_ptr()), geon.C.c_void_p(stream.cuda_stream)),
                "geon_curvature_dev")
            e1.record(stream)
            e1.synchronize()
        ms = e0.elapsed_time(e1)
        curv[label] = {"ms": round(ms, 3), "points_per_s": n / (ms * 1e-3)}

    result = {
        "metric": "RK4 ray-steps/s on 1 GPU",
        "value": value,
        "unit": "ray-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (camera rays generated on device)",
        "config": {"workload": "generic-metric FP64 renders: Alcubierre 1024² (x=-10 literal, "
                               "x=+10 bubble in view), Morris-Thorne 1024² (l=15); step 0.05",
                   "ray_steps_per_step": ray_steps_per_step,
                   "per_render_ms": {k: round(v, 3) for k, v in per_render_ms.items()},
                   "per_render_ray_steps": {k: int(v["ray_steps"]) for k, v in stats.items()},
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "curvature": curv,
        "wall_s": round(wall, 3),
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(jobs)
    print(json.dumps(result))


def cpu_baseline(jobs):
    """Oracle (reference algorithm in C) on a bounded sample of the Morris–Thorne rays."""
    import numpy as np

    from oracle import oracle as O
    name, m, cam, cfg = jobs[2]
    om = O.metric(O.MORRIS_THORNE, throat_radius=m.throat_radius)
    ocam = O.camera(cam.observer, cam.fov_y_deg, cam.width, cam.height)
    st, tet = O.tetrad(om, ocam.observer)
    rng = np.random.default_rng(0)
    idx = rng.choice(cam.width * cam.height, size=256, replace=False)
    y0 = np.array([O.ray_state_sub(ocam, tet, int(i) % cam.width, int(i) // cam.width)
                   for i in idx])
    threads = O.THREADS
    res = O.trace_rays(om, y0, cfg.step, cfg.max_steps, nthreads=threads)
    steps = int(res["steps"].sum())
    return {"value": steps / res["seconds"], "unit": "ray-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(idx)} random pixels of the Morris-Thorne 1024² render "
                      f"({steps} ray-steps, {res['seconds']:.1f} s)"}


if __name__ == "__main__":
    main()
