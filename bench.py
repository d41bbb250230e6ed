"""Benchmark of the geodesic path on one MI355X.

Headline (`value`): RK4 ray-steps/s of the FP64 generic-metric renders —
Alcubierre R=2, σ=8 (thickness 1/8), v_s=0.5, 1024² from x=-10 (as given:
the reference camera looks toward −x, away from the bubble) and from x=+10
(bubble in view, max 40,000 steps); Morris–Thorne b0=1, l=15, 1024², max
20,000 (reference defaults: nearly every ray runs out of steps). One bench
step renders all three; rays are generated on the device from the camera and
outputs stay in HBM. Timing: HIP events recorded on the stream every kernel
is launched on.

Also reported on the same run:
  precision_split  Schwarzschild 4096², r=30M, 60°, capture r<2.01M, escape
                   100M, max 4000: plain FP64 kernel vs FP32-mixed scheduler
                   (FP64 for |b−3√3M|<1e-3M and flagged rays); pixel-class
                   mismatch and escape-angle error vs the FP64 frame, plus a
                   CPU-oracle check on a pixel sample
  ragged_batch     256 frames, sides log-uniform 64..2048, random metrics, one
                   persistent launch per kind, FP64 and FP32-mixed
  serving          Poisson arrivals, 80% renders 32²..2048² / 20% bisections
                   64×40, p50/p99 latency and sustained ray-steps/s of the
                   FP32-mixed server, next to the plain FP64 kernel on the same
                   renders
  curvature        Alcubierre T_μν on 128³ over [-4,4]³, Morris–Thorne on 4096 points
  adaptive         Cash–Karp RK4(5) traces of the Morris–Thorne camera rays at
                   512², next to the fixed-step trace of the same rays (both on
                   device-resident states)
  cpu_baseline     the oracle (reference algorithm in C, -O3, no FMA) on a sample
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
FP32_PEAK_TFLOPS = 157.3


def measured_traffic():
    """HBM bytes per launch of geon_render_rays_mt on this workload, measured by
    rocprofv3 FETCH_SIZE/WRITE_SIZE passes scaled by a 1 GiB copy (the summary
    and its commands are in profiles/render_mt_fp64_hbm_traffic.txt)."""
    path = os.path.join(ROOT, "profiles", "render_mt_fp64_hbm_traffic.txt")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def op_counts():
    """FP64 operations per RK4 ray-step and per curvature point of the kernels'
    own formulation, counted by geon_amd/opcount (add/sub/mul/div/sqrt/sin/cos/
    tanh count 1; structural zeros are not executed and not counted)."""
    with open(os.path.join(ROOT, "geon_amd", "opcount", "op_counts.json")) as f:
        return json.load(f)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--quick", action="store_true", help="headline only")
    ap.add_argument("--sections", default="split,ragged,serving,curvature,adaptive",
                    help="comma-separated extra sections to run after the headline")
    ap.add_argument("--copy-probe", action="store_true",
                    help="one 1 GiB device-to-device copy (HBM counter calibration) and exit")
    ap.add_argument("--serving-seconds", type=float, default=2.0)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)
    import torch
    if args.copy_probe:
        n = 1 << 28                                   # 2^28 floats = 1 GiB each way
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": 4 * n, "copy_probe_bytes_written": 4 * n}))
        return

    import geon_amd as geon

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)
    ctx = geon.Context(0)
    ev = lambda: torch.cuda.Event(enable_timing=True)  # noqa: E731

    def outs(n):
        return (torch.empty(3 * n, dtype=torch.float32, device=dev),
                torch.empty(n, dtype=torch.int32, device=dev),
                torch.empty(3 * n, dtype=torch.float64, device=dev))

    # ---------------- headline: FP64 generic-metric renders ----------------
    alc = geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5, ship_x=0.0)
    mt = geon.MorrisThorne(throat_radius=1.0)
    jobs = [
        ("alcubierre_x-10", alc, geon.Camera((0.0, -10.0, 0.0, 0.0), 60.0, 1024, 1024),
         geon.RenderConfig(step=0.05, max_steps=20_000)),
        ("alcubierre_x+10", alc, geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 1024, 1024),
         geon.RenderConfig(step=0.05, max_steps=40_000)),
        ("morris_thorne_l15", mt, geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, 1024, 1024),
         geon.RenderConfig(step=0.05, max_steps=20_000)),
    ]
    bufs = {name: outs(cam.width * cam.height) for name, _, cam, _ in jobs}

    def run_once(want_stats):
        res = {}
        for name, m, cam, cfg in jobs:
            rgb, cls, d = bufs[name]
            res[name] = geon.render_dev(m, cam, cfg, rgb.data_ptr(), cls.data_ptr(), d.data_ptr(),
                                        stream_ptr=stream.cuda_stream, want_stats=want_stats,
                                        ctx=ctx)
        return res

    stats = None
    for _ in range(max(1, args.warmup)):
        stats = run_once(True)
    ray_steps_per_step = sum(int(s["ray_steps"]) for s in stats.values())
    torch.cuda.synchronize()
    t0, t1 = ev(), ev()
    wall0 = time.perf_counter()
    t0.record(stream)
    for _ in range(args.steps):
        run_once(False)
    t1.record(stream)
    t1.synchronize()
    wall = time.perf_counter() - wall0
    ms_total = t0.elapsed_time(t1)
    ms_per_step = ms_total / args.steps
    value = ray_steps_per_step * args.steps / (ms_total * 1e-3)
    final = run_once(True)
    for k in stats:
        assert final[k]["ray_steps"] == stats[k]["ray_steps"], k
    mt_ms = final["morris_thorne_l15"]["kernel_ms"]
    mt_steps = int(final["morris_thorne_l15"]["ray_steps"])
    mt_ops = op_counts()["ray_step"]["morris_thorne"]["ops"]
    achieved = mt_steps * mt_ops / (mt_ms * 1e-3) / 1e12
    roofline = {"bound": "valu_fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                "traffic": measured_traffic(), "algorithmic_bytes": 40 * 1024 * 1024,
                "kernel": "geon_render_rays_mt", "ops_per_ray_step": mt_ops,
                "ray_steps_per_launch": mt_steps, "kernel_ms": round(mt_ms, 3)}
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
        "config": {"workload": "generic-metric FP64 renders: Alcubierre 1024² (x=-10 as given, "
                               "x=+10 bubble in view), Morris-Thorne 1024² (l=15); step 0.05",
                   "ray_steps_per_step": ray_steps_per_step,
                   "per_render_ray_steps": {k: int(v["ray_steps"]) for k, v in stats.items()},
                   "per_render_kernel_ms": {k: round(v["kernel_ms"], 3) for k, v in final.items()},
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "wall_s": round(wall, 3),
    }
    log(f"headline: {value:.4e} ray-steps/s, {ms_per_step:.1f} ms per step")
    sections = set() if args.quick else set(args.sections.split(","))
    if "split" in sections:
        result["precision_split"] = precision_split(geon, ctx, torch, stream, outs, ev)
        log(f"precision split: {json.dumps(result['precision_split'])}")
    if "ragged" in sections:
        result["ragged_batch"] = ragged_batch(geon, ctx, torch, stream, ev)
        log(f"ragged batch: {json.dumps(result['ragged_batch'])}")
    if "serving" in sections:
        result["serving"] = serving(geon, ctx, torch, stream, outs, ev, args.serving_seconds)
        log(f"serving: {json.dumps(result['serving'])}")
    if "curvature" in sections:
        result["curvature"] = curvature(geon, ctx, torch, stream, ev, alc, mt)
        log(f"curvature: {json.dumps(result['curvature'])}")
    if "adaptive" in sections:
        result["adaptive"] = adaptive(geon, ctx, torch, stream, ev, jobs)
        log(f"adaptive: {json.dumps(result['adaptive'])}")
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(jobs)
        log(f"cpu baseline: {json.dumps(result['cpu_baseline'])}")
    print(json.dumps(result))


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _angle(np, a, b):
    cross = np.linalg.norm(np.cross(a, b), axis=-1)
    return np.arctan2(cross, (a * b).sum(axis=-1))


def precision_split(geon, ctx, torch, stream, outs, ev):
    import numpy as np
    metric = geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    cam = geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 4096, 4096)
    n = cam.width * cam.height
    o64 = outs(n)
    cfg64 = geon.RenderConfig(step=0.05, max_steps=4000)
    geon.render_dev(metric, cam, cfg64, o64[0].data_ptr(), o64[1].data_ptr(), o64[2].data_ptr(),
                    stream_ptr=stream.cuda_stream, ctx=ctx)
    a, b = ev(), ev()
    a.record(stream)
    st64 = geon.render_dev(metric, cam, cfg64, o64[0].data_ptr(), o64[1].data_ptr(),
                           o64[2].data_ptr(), stream_ptr=stream.cuda_stream, ctx=ctx)
    b.record(stream)
    b.synchronize()
    ms64 = a.elapsed_time(b)
    tuned = geon.autotune(geon.SCHWARZSCHILD, geon.PREC_MIXED, ctx=ctx)
    om = outs(n)
    cfgm = geon.RenderConfig(step=0.05, max_steps=4000, precision=geon.PREC_MIXED)
    best = None
    for _ in range(2):
        bt = geon.Batch(ctx)
        fid = bt.add_frame(metric, cam, cfgm, om[0].data_ptr(), om[1].data_ptr(), om[2].data_ptr())
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record(stream)
        bt.run(stream_ptr=stream.cuda_stream)
        b.record(stream)
        b.synchronize()
        ms = a.elapsed_time(b)
        stm = bt.frame_stats(fid)
        executed = bt.totals()["executed_ray_steps"]
        bt.close()
        best = (ms, stm, executed) if best is None or ms < best[0] else best
    msm, stm, executed = best
    f32_ops = op_counts()["ray_step_fp32"]["schwarzschild"]["ops"]
    f32_tflops = executed * f32_ops / (msm * 1e-3) / 1e12
    cls64, clsm = o64[1].cpu().numpy(), om[1].cpu().numpy()
    d64 = o64[2].cpu().numpy().reshape(3, -1).T
    dm = om[2].cpu().numpy().reshape(3, -1).T
    same = cls64 == clsm
    esc = same & (cls64 == 2) & (np.abs(d64).max(axis=1) > 0) & (np.abs(dm).max(axis=1) > 0)
    ang = _angle(np, dm[esc], d64[esc])
    # oracle spot check of the mixed frame on a pixel sample
    from oracle import oracle as O
    rng = np.random.default_rng(3)
    idx = rng.choice(n, size=1024, replace=False)
    omet = O.metric(O.SCHWARZSCHILD, mass=1.0, horizon_pad=1.005, escape_radius=100.0)
    ocam = O.camera(cam.observer, 60.0, 4096, 4096)
    st, tet = O.tetrad(omet, ocam.observer)
    y0 = np.array([O.ray_state_sub(ocam, tet, int(i) % 4096, int(i) // 4096) for i in idx])
    ref = O.trace_rays(omet, y0, 0.05, 4000)
    o_mis = float((ref["cls"] != clsm[idx]).mean())
    oesc = (ref["cls"] == clsm[idx]) & (ref["cls"] == 2) & (ref["y"][:, 1] < 1000.0)
    o_ang = _angle(np, dm[idx][oesc], ref["dir"][oesc])
    return {
        "workload": "Schwarzschild 4096², r=30M, 60°, step 0.05, max 4000, capture r<2.01M, "
                    "escape 100M",
        "fp64_plain_ms": round(ms64, 3),
        "fp64_plain_ray_steps_per_s": st64["ray_steps"] / (ms64 * 1e-3),
        "mixed_ms": round(msm, 3),
        "mixed_ray_steps_per_s": stm["ray_steps"] / (msm * 1e-3),
        "speedup_vs_plain_fp64": ms64 / msm * stm["ray_steps"] / st64["ray_steps"],
        "mixed_fp64_rays": int(stm["fp64_rays"]),
        "pixel_class_mismatch_vs_fp64": float((~same).mean()),
        "escape_angle_err_rad": {"max": float(ang.max()), "mean": float(ang.mean()),
                                 "p99.99": float(np.percentile(ang, 99.99)),
                                 "above_1e-3": int((ang > 1e-3).sum()), "escaped": int(esc.sum())},
        "oracle_sample": {"pixels": 1024, "class_mismatch": o_mis,
                          "escape_angle_err_max": float(o_ang.max()) if o_ang.size else 0.0},
        "tuned_launch": tuned,
        # every executed step (FP32 and the FP64 re-traces, ~3 %) at the FP32
        # step's counted operations, against the published FP32 vector peak
        "roofline_fp32": {"achieved": round(f32_tflops, 3), "peak": FP32_PEAK_TFLOPS,
                          "unit": "TFLOP/s", "frac": round(f32_tflops / FP32_PEAK_TFLOPS, 4),
                          "ops_per_ray_step": f32_ops, "executed_ray_steps": executed},
    }


def ragged_batch(geon, ctx, torch, stream, ev):
    from geon_amd import workloads
    res = {}
    for label, prec in (("fp64", geon.PREC_FP64), ("mixed", geon.PREC_MIXED)):
        geon.autotune(geon.SCHWARZSCHILD, prec, ctx=ctx)
        geon.autotune(geon.ALCUBIERRE, prec, ctx=ctx)
        geon.autotune(geon.MORRIS_THORNE, prec, ctx=ctx)
        frames = workloads.ragged_frames(precision=prec)
        bt = geon.Batch(ctx)
        ids = [bt.add_frame(m, cam, cfg) for m, cam, cfg in frames]
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record(stream)
        bt.run(stream_ptr=stream.cuda_stream)
        b.record(stream)
        b.synchronize()
        ms = a.elapsed_time(b)
        steps = sum(int(bt.frame_stats(i)["ray_steps"]) for i in ids)
        tot = bt.totals()
        bt.close()
        res[label] = {"ms": round(ms, 2), "ray_steps": steps,
                      "ray_steps_per_s": steps / (ms * 1e-3), "fp64_rays": tot["fp64_rays"]}
    res["workload"] = "256 frames, side log-uniform 64..2048, metrics uniform over " \
                      "{Schwarzschild M∈[0.5,2], Alcubierre v∈[0.1,0.9], Morris-Thorne b0∈[0.5,2]}, " \
                      "escape radius 80, max 5000, seed 1234"
    res["speedup_mixed_vs_fp64"] = res["mixed"]["ray_steps_per_s"] / res["fp64"]["ray_steps_per_s"]
    return res


def serving(geon, ctx, torch, stream, outs, ev, seconds):
    """Poisson request streams of `seconds` of arrivals at each rate."""
    import numpy as np

    from geon_amd import workloads
    out = {"workload": f"Poisson arrivals for {seconds} s at each rate: 80% renders side "
                       "log-uniform 32..2048 (random metric, camera r∈[10,50], escape 100, max 5000, "
                       "FP32-mixed), 20% bisections 64 rays × 40 halvings (M=1, r_obs∈[10,50], "
                       "photon-sphere decisions); seed 7; 8 GiB admission budget, 2048-step quanta",
           "rates": {}}
    plain_rate = None
    for rate in (50, 100, 200, 400):
        reqs = workloads.serving_stream(rate, int(rate * seconds))
        srv = geon.Server(ctx, hbm_budget=8 << 30, quantum_steps=2048)
        tickets = []
        t_start = time.perf_counter()
        i = 0
        while i < len(reqs):
            now = time.perf_counter() - t_start
            while i < len(reqs) and reqs[i]["t"] <= now:
                r = reqs[i]
                if r["kind"] == "render":
                    tickets.append(("render", srv.submit_render(r["metric"], r["camera"], r["cfg"])))
                else:
                    tickets.append(("bisect", srv.submit_bisect(r["mass"], r["r_obs"],
                                                                iters=r["iters"])))
                i += 1
            time.sleep(0.0005)
        srv.drain()
        wall = time.perf_counter() - t_start
        lat = {"render": [], "bisect": []}
        steps = 0
        for kind, t in tickets:
            res = srv.poll(t)
            assert res["status"] == 0, res
            lat[kind].append(res["latency_ms"])
            steps += res["ray_steps"]
        c = srv.counters()
        srv.close()
        pr = lambda v, q: round(float(np.percentile(v, q)), 2) if v else None  # noqa: E731
        out["rates"][str(rate)] = {
            "requests": len(tickets), "wall_s": round(wall, 3),
            "render_latency_ms": {"p50": pr(lat["render"], 50), "p99": pr(lat["render"], 99)},
            "bisect_latency_ms": {"p50": pr(lat["bisect"], 50), "p99": pr(lat["bisect"], 99)},
            "offered_ray_steps_per_s": steps / reqs[-1]["t"],
            "served_ray_steps_per_s": steps / wall, "quanta": c["quanta"],
            "device_busy_frac": round(c["busy_ms"] * 1e-3 / wall, 3),
        }
        log(f"serving rate {rate}/s: {json.dumps(out['rates'][str(rate)])}")
        if plain_rate is None:
            # the same renders through the plain FP64 kernel, one after another
            tot_steps, tot_ms = 0, 0.0
            for r in reqs:
                if r["kind"] != "render":
                    continue
                cam = r["camera"]
                o = outs(cam.width * cam.height)
                cfg = geon.RenderConfig(step=0.05, max_steps=5000)
                a, b = ev(), ev()
                a.record(stream)
                st = geon.render_dev(r["metric"], cam, cfg, o[0].data_ptr(), o[1].data_ptr(),
                                     o[2].data_ptr(), stream_ptr=stream.cuda_stream, ctx=ctx)
                b.record(stream)
                b.synchronize()
                tot_ms += a.elapsed_time(b)
                tot_steps += st["ray_steps"]
            plain_rate = tot_steps / (tot_ms * 1e-3)
            out["plain_fp64_same_renders_ray_steps_per_s"] = plain_rate
    # capacity: the highest sustained rate (arrivals beyond it queue up)
    best = max(v["served_ray_steps_per_s"] for v in out["rates"].values())
    out["served_vs_plain_fp64"] = best / plain_rate
    return out


def curvature(geon, ctx, torch, stream, ev, alc, mt):
    res = {}
    g = torch.linspace(-4.0, 4.0, 128, dtype=torch.float64, device="cuda")
    X, Y, Z = torch.meshgrid(g, g, g, indexing="ij")
    xs = torch.stack([torch.zeros_like(X).ravel(), X.ravel(), Y.ravel(), Z.ravel()]).contiguous()
    lline = torch.linspace(-4.0, 4.0, 4096, dtype=torch.float64, device="cuda")
    xl = torch.stack([torch.zeros_like(lline), lline, torch.full_like(lline, math.pi / 2),
                      torch.zeros_like(lline)]).contiguous()
    ops = op_counts()["curvature_stress"]
    for label, key, m, x in (("alcubierre_128^3", "alcubierre", alc, xs),
                             ("morris_thorne_4096", "morris_thorne", mt, xl)):
        n = x.shape[1]
        o = torch.empty(n * 16, dtype=torch.float64, device="cuda")
        times = []
        for _ in range(5):
            a, b = ev(), ev()
            a.record(stream)
            geon._abi.check(geon.lib().geon_curvature_dev(
                ctx.handle, geon.C.byref(geon._cmetric(m)), geon.C.c_void_p(x.data_ptr()), n,
                geon.CURV_STRESS, geon.C.c_void_p(o.data_ptr()),
                geon.C.c_void_p(stream.cuda_stream)), "geon_curvature_dev")
            b.record(stream)
            b.synchronize()
            times.append(a.elapsed_time(b))
        ms = min(times[1:])
        tf = n * ops[key]["ops"] / (ms * 1e-3) / 1e12
        res[label] = {"ms": round(ms, 4), "points_per_s": n / (ms * 1e-3),
                      "ops_per_point": ops[key]["ops"], "achieved_tflops": round(tf, 3),
                      "frac_fp64_peak": round(tf / FP64_PEAK_TFLOPS, 4)}
    return res


def adaptive(geon, ctx, torch, stream, ev, jobs):
    """geon_trace_rays_adaptive_dev on the Morris–Thorne frame's camera rays
    at 512² (h0 = the render's step, at most its max_steps accepted steps,
    two tolerances), and geon_trace_rays_dev (fixed RK4 steps) on the same
    rays; states resident in HBM, HIP events on the launch stream. A
    Cash–Karp step evaluates the right-hand side 6 times and an RK4 step 4
    times: rhs_per_s compares the kernels."""
    import numpy as np
    _, m, cam, cfg = jobs[2]
    y0 = geon.ray_states(m, geon.Camera(cam.observer, cam.fov_y_deg, 512, 512))
    n = len(y0)
    dy0 = torch.from_numpy(np.ascontiguousarray(y0.T)).to("cuda")
    cls = torch.empty(n, dtype=torch.int32, device="cuda")
    steps = torch.empty(n, dtype=torch.int32, device="cuda")
    rej = torch.empty(n, dtype=torch.int32, device="cuda")
    yf = torch.empty(8 * n, dtype=torch.float64, device="cuda")
    dirs = torch.empty(3 * n, dtype=torch.float64, device="cuda")
    cm = geon.C.byref(geon._cmetric(m))
    vp = geon.C.c_void_p
    torch.cuda.synchronize()   # the states' copy ran on torch's stream
    res = {"workload": "Morris-Thorne b0=1 camera rays, l=15, 60°, 512²", "rays": n,
           "h0": cfg.step, "max_steps": cfg.max_steps}

    def timed(launch):
        times = []
        for _ in range(3):
            a, b = ev(), ev()
            a.record(stream)
            launch()
            b.record(stream)
            b.synchronize()
            times.append(a.elapsed_time(b))
        return min(times)

    def fixed():
        geon._abi.check(geon.lib().geon_trace_rays_dev(
            ctx.handle, cm, vp(dy0.data_ptr()), n, cfg.step, cfg.max_steps, vp(cls.data_ptr()),
            vp(steps.data_ptr()), vp(yf.data_ptr()), None, vp(dirs.data_ptr()),
            vp(stream.cuda_stream)), "geon_trace_rays_dev")

    ms = timed(fixed)
    st = int(steps.to(torch.int64).sum())
    res["fixed_rk4"] = {"ms": round(ms, 3), "steps": st, "steps_per_s": st / (ms * 1e-3),
                        "rhs_per_s": 4 * st / (ms * 1e-3),
                        "classes": torch.bincount(cls, minlength=4).tolist()}
    for rtol, atol in ((1e-6, 1e-9), (1e-9, 1e-12)):
        def run():
            geon._abi.check(geon.lib().geon_trace_rays_adaptive_dev(
                ctx.handle, cm, vp(dy0.data_ptr()), n, cfg.step, rtol, atol, cfg.max_steps,
                vp(cls.data_ptr()), vp(steps.data_ptr()), vp(rej.data_ptr()), vp(yf.data_ptr()),
                vp(dirs.data_ptr()), vp(stream.cuda_stream)), "geon_trace_rays_adaptive_dev")
        ms = timed(run)
        acc, rj = int(steps.to(torch.int64).sum()), int(rej.to(torch.int64).sum())
        res[f"cash_karp_rtol{rtol:g}"] = {
            "ms": round(ms, 3), "accepted": acc, "rejected": rj,
            "steps_per_s": (acc + rj) / (ms * 1e-3), "rhs_per_s": 6 * (acc + rj) / (ms * 1e-3),
            "classes": torch.bincount(cls, minlength=4).tolist()}
    return res


def cpu_baseline(jobs):
    """Oracle (reference algorithm in C) on a bounded sample of the Morris–Thorne rays:
    1024 random pixels on all the threads it may use (~25 CPU-seconds), and 64
    of them on one thread."""
    import platform

    import numpy as np

    from oracle import oracle as O
    name, m, cam, cfg = jobs[2]
    om = O.metric(O.MORRIS_THORNE, throat_radius=m.throat_radius)
    ocam = O.camera(cam.observer, cam.fov_y_deg, cam.width, cam.height)
    st, tet = O.tetrad(om, ocam.observer)
    rng = np.random.default_rng(0)
    idx = rng.choice(cam.width * cam.height, size=1024, replace=False)
    y0 = np.array([O.ray_state_sub(ocam, tet, int(i) % cam.width, int(i) // cam.width)
                   for i in idx])
    threads = O.THREADS
    res = O.trace_rays(om, y0, cfg.step, cfg.max_steps, nthreads=threads)
    steps = int(res["steps"].sum())
    one = O.trace_rays(om, y0[:64], cfg.step, cfg.max_steps, nthreads=1)
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": steps / res["seconds"], "unit": "ray-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"{len(idx)} random pixels of the Morris-Thorne 1024² render "
                      f"({steps} ray-steps, {res['seconds']:.2f} s on {threads} threads)",
            "single_thread_ray_steps_per_s": int(one["steps"].sum()) / one["seconds"],
            "host_cpus": os.cpu_count(), "cpu_model": model}


if __name__ == "__main__":
    main()
