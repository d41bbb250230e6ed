"""Benchmark of the geodesic path on one MI355X.

Headline (`value`): RK4 ray-steps/s of the FP64 generic-metric renders,
counting the steps that evaluate the Christoffel symbols at their four stages
(`gamma_steps`) — Alcubierre R=2, σ=8 (thickness 1/8), v_s=0.5, 1024² from
x=-10 (as given: the reference camera looks toward −x, away from the bubble)
and from x=+10 (bubble in view, max 40,000 steps); Morris–Thorne b0=1, l=15,
1024², max 20,000 (reference defaults: nearly every ray runs out of steps).
Steps through regions where Γ is exactly zero (beyond the Alcubierre wall),
which the kernels take as spans of additions, are reported separately
(`all_ray_steps_per_s`). One bench step renders all three; rays are generated
on the device from the camera and outputs stay in HBM. Timing: HIP events
recorded on the stream every kernel is launched on.

Also reported on the same run:
  precision_split  Schwarzschild 4096², r=30M, 60°, capture r<2.01M, escape
                   100M, max 4000: plain FP64 kernel vs FP32-mixed scheduler
                   (FP64 for |b−3√3M|<1e-3M and flagged rays); pixel-class
                   mismatch and escape-angle error vs the FP64 frame, plus a
                   CPU-oracle check on a pixel sample
  ragged_batch     256 frames, sides log-uniform 64..2048, random metrics, one
                   persistent launch per kind, FP64 and FP32-mixed, with the
                   mixed frames' class mismatch and escape-angle error against
                   the FP64 frames
  serving          Poisson arrivals, 80% renders 32²..2048² / 20% bisections
                   64×40, p50/p99 latency of the FP32-mixed server; at each
                   rate the plain FP64 kernel runs the same renders one after
                   another, and a seeded sample of served renders is compared
                   with those FP64 frames
  curvature        Alcubierre T_μν on 128³ over [-4,4]³, Morris–Thorne on 4096 points
  adaptive         Cash–Karp RK4(5) traces of the Morris–Thorne camera rays at
                   512², next to the fixed-step trace of the same rays (both on
                   device-resident states)
  cpu_baseline     the oracle (reference algorithm in C, -O3, no FMA) on a sample
                   of the Morris–Thorne frame's pixels, checked against the
                   GPU frame in HBM
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
METRIC = "RK4 ray-steps/s on 1 GPU; p99 frame latency; FP32-mixed vs FP64 at \u22641e-3 mismatch"
# Minimal FP64 operations per Morris-Thorne RK4 ray-step (51 add/mul + 5 div +
# 3 sin per right-hand side, four per step, plus the RK4 combination); the
# kernels' own formulation, which keeps the reference's rounding, executes 442.
MT_OPS_MINIMAL = 340


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
    ap.add_argument("--quantum-steps", type=int, default=16384,
                    help="serving: steps per lane in each quantum")
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
    gamma_per_step = sum(int(s["gamma_steps"]) for s in stats.values())
    all_per_step = sum(int(s["ray_steps"]) for s in stats.values())
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
    value = gamma_per_step * args.steps / (ms_total * 1e-3)
    final = run_once(True)
    for k in stats:
        assert final[k]["ray_steps"] == stats[k]["ray_steps"], k
        assert final[k]["gamma_steps"] == stats[k]["gamma_steps"], k
    mt_ms = final["morris_thorne_l15"]["kernel_ms"]
    mt_steps = int(final["morris_thorne_l15"]["ray_steps"])
    mt_counted = op_counts()["ray_step"]["morris_thorne"]["ops"]
    achieved = mt_steps * MT_OPS_MINIMAL / (mt_ms * 1e-3) / 1e12
    achieved_counted = mt_steps * mt_counted / (mt_ms * 1e-3) / 1e12
    roofline = {"bound": "valu_fp64", "achieved": round(achieved, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
                "traffic": measured_traffic(), "algorithmic_bytes": 40 * 1024 * 1024,
                "kernel": "geon_render_rays_mt", "ops_per_ray_step": MT_OPS_MINIMAL,
                "ops_per_ray_step_executed_formulation": mt_counted,
                "frac_executed_formulation": round(achieved_counted / FP64_PEAK_TFLOPS, 4),
                "ray_steps_per_launch": mt_steps, "kernel_ms": round(mt_ms, 3)}
    result = {
        "metric": METRIC,
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
                   "value_counts": "RK4 ray-steps that evaluate Γ at their four stages",
                   "ray_steps_per_step": gamma_per_step,
                   "per_render_gamma_ray_steps": {k: int(v["gamma_steps"]) for k, v in stats.items()},
                   "per_render_ray_steps": {k: int(v["ray_steps"]) for k, v in stats.items()},
                   "per_render_kernel_ms": {k: round(v["kernel_ms"], 3) for k, v in final.items()},
                   "parallelism": "single GPU"},
        # every ray-step, including those through exactly flat regions (Γ = 0)
        # that the kernels take as spans of additions
        "all_ray_steps_per_s": all_per_step * args.steps / (ms_total * 1e-3),
        "roofline": roofline,
        "wall_s": round(wall, 3),
    }
    log(f"headline: {value:.4e} Γ-evaluating ray-steps/s ({result['all_ray_steps_per_s']:.4e} "
        f"with flat-span steps), {ms_per_step:.1f} ms per step")
    sections = set() if args.quick else set(args.sections.split(","))
    if "split" in sections:
        result["precision_split"] = precision_split(geon, ctx, torch, stream, outs, ev)
        log(f"precision split: {json.dumps(result['precision_split'])}")
    if "ragged" in sections:
        result["ragged_batch"] = ragged_batch(geon, ctx, torch, stream, ev)
        log(f"ragged batch: {json.dumps(result['ragged_batch'])}")
    if "serving" in sections:
        result["serving"] = serving(geon, ctx, torch, stream, outs, ev, args.serving_seconds,
                                    args.quantum_steps)
        log(f"serving: {json.dumps(result['serving'])}")
    if "curvature" in sections:
        result["curvature"] = curvature(geon, ctx, torch, stream, ev, alc, mt)
        log(f"curvature: {json.dumps(result['curvature'])}")
    if "adaptive" in sections:
        result["adaptive"] = adaptive(geon, ctx, torch, stream, ev, jobs)
        log(f"adaptive: {json.dumps(result['adaptive'])}")
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(jobs, bufs["morris_thorne_l15"])
        log(f"cpu baseline: {json.dumps(result['cpu_baseline'])}")
    if "serving" in result:
        r100 = result["serving"]["rates"].get("100") or next(iter(result["serving"]["rates"].values()))
        result["p99_frame_latency_ms"] = {"rate_req_per_s": r100["rate"],
                                          "render_p99": r100["render_latency_ms"]["p99"],
                                          "render_p50": r100["render_latency_ms"]["p50"]}
    if "precision_split" in result:
        ps = result["precision_split"]
        result["fp32_mixed_vs_fp64"] = {"speedup": ps["speedup_vs_plain_fp64"],
                                        "pixel_class_mismatch": ps["pixel_class_mismatch_vs_fp64"],
                                        "workload": "Schwarzschild 4096² precision split"}
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


def _mixed_vs_fp64(torch, cls64, dir64, clsm, dirm, acc):
    """Accumulate class mismatches and escape-angle errors of one FP32-mixed
    frame against its FP64 frame (device tensors: classes n, directions SoA 3n)."""
    acc["pixels"] += int(cls64.numel())
    acc["class_mismatch"] += int((cls64 != clsm).sum())
    d64, dm = dir64.view(3, -1).T, dirm.view(3, -1).T
    both = (cls64 == 2) & (clsm == 2) & (d64.abs().amax(dim=1) > 0) & (dm.abs().amax(dim=1) > 0)
    if int(both.sum()):
        a, b = d64[both], dm[both]
        ang = torch.atan2(torch.linalg.cross(a, b).norm(dim=1), (a * b).sum(dim=1))
        acc["angle_max"] = max(acc["angle_max"], float(ang.max()))
        acc["angle_sum"] += float(ang.sum())
        acc["escaped"] += int(both.sum())
        acc["above_1e-3"] += int((ang > 1e-3).sum())


def _acc():
    return {"pixels": 0, "class_mismatch": 0, "escaped": 0, "angle_max": 0.0, "angle_sum": 0.0,
            "above_1e-3": 0}


def _acc_report(a):
    return {"pixel_class_mismatch_vs_fp64": a["class_mismatch"] / max(1, a["pixels"]),
            "escape_angle_err_rad": {"max": a["angle_max"],
                                     "mean": a["angle_sum"] / max(1, a["escaped"]),
                                     "above_1e-3": a["above_1e-3"], "escaped": a["escaped"]},
            "pixels": a["pixels"]}


_TUNED = set()


def tune_kinds(geon, ctx):
    """Autotune the scheduler's launch for every metric kind of the batch and
    serving workloads in both precisions, once per context."""
    if id(ctx) in _TUNED:
        return
    for prec in (geon.PREC_FP64, geon.PREC_MIXED):
        for kind in (geon.SCHWARZSCHILD, geon.ALCUBIERRE, geon.MORRIS_THORNE):
            geon.autotune(kind, prec, ctx=ctx)
    _TUNED.add(id(ctx))


def _run_batch(geon, ctx, torch, stream, ev, frames):
    """One batch launch over `frames`; returns (device outputs, stats)."""
    bt = geon.Batch(ctx)
    o = []
    for m, cam, cfg in frames:
        n = cam.width * cam.height
        cls = torch.empty(n, dtype=torch.int32, device="cuda")
        d = torch.empty(3 * n, dtype=torch.float64, device="cuda")
        o.append((type(m).__name__, cls, d))
    ids = [bt.add_frame(m, cam, cfg, 0, c.data_ptr(), d.data_ptr())
           for (m, cam, cfg), (_, c, d) in zip(frames, o)]
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
    return o, {"ms": round(ms, 2), "ray_steps": steps,
               "ray_steps_per_s": steps / (ms * 1e-3), "fp64_rays": tot["fp64_rays"]}


def ragged_batch(geon, ctx, torch, stream, ev):
    from geon_amd import workloads
    res = {}
    outs = {}
    tune_kinds(geon, ctx)
    for label, prec in (("fp64", geon.PREC_FP64), ("mixed", geon.PREC_MIXED)):
        outs[label], res[label] = _run_batch(geon, ctx, torch, stream, ev,
                                             workloads.ragged_frames(precision=prec))
    # each metric kind's frames alone, FP64 and FP32-mixed
    per_kind = {}
    for kind in ("Schwarzschild", "MorrisThorne", "Alcubierre"):
        row = {}
        for label, prec in (("fp64", geon.PREC_FP64), ("mixed", geon.PREC_MIXED)):
            fr = [f for f in workloads.ragged_frames(precision=prec) if type(f[0]).__name__ == kind]
            _, st = _run_batch(geon, ctx, torch, stream, ev, fr)
            row["frames"] = len(fr)
            row[label + "_ms"] = st["ms"]
            row[label + "_fp64_rays"] = st["fp64_rays"]
        row["speedup_mixed_vs_fp64"] = row["fp64_ms"] / row["mixed_ms"]
        per_kind[kind] = row
    res["per_kind_alone"] = per_kind
    per, allk = {}, _acc()
    for (kind, c64, d64), (_, cm, dm) in zip(outs["fp64"], outs["mixed"]):
        _mixed_vs_fp64(torch, c64, d64, cm, dm, per.setdefault(kind, _acc()))
        _mixed_vs_fp64(torch, c64, d64, cm, dm, allk)
    res["mixed_vs_fp64"] = _acc_report(allk)
    res["mixed_vs_fp64_per_metric"] = {k: _acc_report(v) for k, v in per.items()}
    res["workload"] = "256 frames, side log-uniform 64..2048, metrics uniform over " \
                      "{Schwarzschild M∈[0.5,2], Alcubierre v∈[0.1,0.9], Morris-Thorne b0∈[0.5,2]}, " \
                      "escape radius 80, max 5000, seed 1234"
    res["speedup_mixed_vs_fp64"] = res["fp64"]["ms"] / res["mixed"]["ms"]
    return res


def serving(geon, ctx, torch, stream, outs, ev, seconds, quantum=16384):
    """Poisson request streams of `seconds` of arrivals at each rate. At each rate
    the plain FP64 kernel then renders the same requests one after another:
    `speedup_vs_plain_fp64` is that FP64 device time over the server's wall
    time for the whole stream (bisections included on the server's side only),
    and a seeded sample of the served renders is compared with the FP64 frames."""
    import numpy as np

    from geon_amd import workloads
    out = {"workload": f"Poisson arrivals for {seconds} s at each rate: 80% renders side "
                       "log-uniform 32..2048 (random metric, camera r∈[10,50], escape 100, max 5000, "
                       "FP32-mixed), 20% bisections 64 rays × 40 halvings (M=1, r_obs∈[10,50], "
                       "photon-sphere decisions); seed 7; 8 GiB admission budget, "
                       f"{quantum}-step quanta",
           "rates": {}}
    tune_kinds(geon, ctx)
    for rate in (50, 100, 200, 400):
        reqs = workloads.serving_stream(rate, int(rate * seconds))
        rng = np.random.default_rng(rate)
        renders = [i for i, r in enumerate(reqs) if r["kind"] == "render"]
        keep = set(rng.choice(renders, min(16, len(renders)), replace=False).tolist())
        srv = geon.Server(ctx, hbm_budget=8 << 30, quantum_steps=quantum)
        tickets = []
        t_start = time.perf_counter()
        i = 0
        while i < len(reqs):
            now = time.perf_counter() - t_start
            while i < len(reqs) and reqs[i]["t"] <= now:
                r = reqs[i]
                if r["kind"] == "render":
                    tickets.append(("render", srv.submit_render(r["metric"], r["camera"], r["cfg"],
                                                                keep=i in keep)))
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
        served = {k: srv.fetch_render(tickets[k][1], reqs[k]["camera"].width,
                                      reqs[k]["camera"].height) for k in sorted(keep)}
        c = srv.counters()
        srv.close()
        # the same renders through the plain FP64 kernel, one after another
        fp64_ms, fp64_steps, fp64_gamma, acc = 0.0, 0, 0, _acc()
        for k, r in enumerate(reqs):
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
            fp64_ms += a.elapsed_time(b)
            fp64_steps += st["ray_steps"]
            fp64_gamma += st["gamma_steps"]
            if k in keep:
                _, cls_m, dir_m = served[k]
                _mixed_vs_fp64(torch, o[1], o[2], torch.from_numpy(cls_m).cuda(),
                               torch.from_numpy(np.ascontiguousarray(dir_m.T).reshape(-1)).cuda(),
                               acc)
        pr = lambda v, q: round(float(np.percentile(v, q)), 2) if v else None  # noqa: E731
        out["rates"][str(rate)] = {
            "rate": rate, "requests": len(tickets), "wall_s": round(wall, 3),
            "render_latency_ms": {"p50": pr(lat["render"], 50), "p99": pr(lat["render"], 99)},
            "bisect_latency_ms": {"p50": pr(lat["bisect"], 50), "p99": pr(lat["bisect"], 99)},
            "offered_ray_steps_per_s": steps / reqs[-1]["t"],
            "served_ray_steps_per_s": steps / wall, "quanta": c["quanta"],
            "device_busy_frac": round(c["busy_ms"] * 1e-3 / wall, 3),
            "plain_fp64_same_renders_ms": round(fp64_ms, 2),
            "plain_fp64_same_renders_ray_steps": fp64_steps,
            "plain_fp64_same_renders_gamma_ray_steps": fp64_gamma,
            "speedup_vs_plain_fp64": fp64_ms * 1e-3 / wall,
            "served_sample_vs_fp64": _acc_report(acc),
        }
        log(f"serving rate {rate}/s: {json.dumps(out['rates'][str(rate)])}")
    best = max(out["rates"].values(), key=lambda v: v["speedup_vs_plain_fp64"])
    out["served_vs_plain_fp64"] = best["speedup_vs_plain_fp64"]
    out["served_vs_plain_fp64_at_rate"] = best["rate"]
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


def cpu_baseline(jobs, frame):
    """Oracle (reference algorithm in C) on a bounded sample of the Morris–Thorne rays:
    1024 random pixels on all the threads it may use (~25 CPU-seconds), and 64
    of them on one thread; the sample's classes and escape directions are
    checked against the GPU frame already in HBM."""
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
    gcls = frame[1].cpu().numpy()[idx]
    gdir = frame[2].view(3, -1).T.cpu().numpy()[idx]
    esc = (gcls == res["cls"]) & (res["cls"] == O.ESCAPED)
    # rays the reference's fixed step flings out at the coordinate pole end far
    # past the escape sphere with a direction of amplified rounding noise
    flung = esc & (np.abs(res["y"][:, 1]) > 10.0 * m.escape_radius)
    real = esc & ~flung
    ang = _angle(np, gdir[real], res["dir"][real]) if real.any() else np.zeros(1)
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
            "gpu_frame_check": {"class_mismatches": int((gcls != res["cls"]).sum()),
                                "escaped": int(real.sum()), "flung_at_pole": int(flung.sum()),
                                "max_angle_rad": float(ang.max())},
            "host_cpus": os.cpu_count(), "cpu_model": model}


if __name__ == "__main__":
    main()
