"""The batched and served workloads at their configured scale.

C4: the 256-frame ragged bundle (sides log-uniform 64..2048, metrics uniform
over Schwarzschild M∈[0.5,2], Alcubierre v∈[0.1,0.9], Morris–Thorne
b0∈[0.5,2], escape radius 80, max 5000, seed 1234) in one FP64 batch and one
FP32-mixed batch:
  * FP32-mixed against FP64, per frame: pixel-class mismatch ≤ 1e-3, and the
    escape-angle error stated (max and mean over pixels escaped in both);
  * 16 seeded frames rendered alone equal their pixels in the batch bit for
    bit, in both precisions;
  * 256 seeded pixels against the oracle: the FP64 batch at the FP64 bar,
    the FP32-mixed batch reported.
C5: a Poisson request stream at 200 req/s (seed 7; 80 % FP32-mixed renders
32²..2048², 20 % bisections 64 × 40) submitted at its arrival times; a
seeded subset of 16 renders and 4 bisections keeps its outputs, and each
equals the same request served alone, bit for bit.
"""
import math
import time

import numpy as np
import pytest

from parity_util import compare_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import geon_amd as geon
    c = geon.Context(0)
    yield c
    c.close()


def _outputs(torch, cam):
    n = cam.width * cam.height
    return (torch.zeros(3 * n, dtype=torch.float32, device="cuda"),
            torch.full((n,), -7, dtype=torch.int32, device="cuda"),
            torch.zeros(3 * n, dtype=torch.float64, device="cuda"))


def _run_batch(geon, ctx, torch, frames):
    b = geon.Batch(ctx)
    outs = []
    for m, cam, cfg in frames:
        o = _outputs(torch, cam)
        b.add_frame(m, cam, cfg, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr())
        outs.append(o)
    b.run()
    tot = b.totals()
    b.close()
    torch.cuda.synchronize()
    return outs, tot


def _ocam(O, cam):
    return O.camera(cam.observer, cam.fov_y_deg, cam.width, cam.height)


def _ometric(O, m):
    import geon_amd as geon
    if isinstance(m, geon.Schwarzschild):
        return O.metric(O.SCHWARZSCHILD, mass=m.mass, horizon_pad=m.horizon_pad,
                        escape_radius=m.escape_radius)
    if isinstance(m, geon.Alcubierre):
        return O.metric(O.ALCUBIERRE, bubble_radius=m.bubble_radius,
                        bubble_thickness=m.bubble_thickness, ship_velocity=m.ship_velocity,
                        ship_x=m.ship_x, escape_radius=m.escape_radius)
    return O.metric(O.MORRIS_THORNE, throat_radius=m.throat_radius, escape_radius=m.escape_radius)


def test_c4_ragged_256_frames(ctx, O):
    import torch

    import geon_amd as geon
    from geon_amd import workloads
    f64 = workloads.ragged_frames(precision=geon.PREC_FP64)
    fmx = workloads.ragged_frames(precision=geon.PREC_MIXED)
    assert len(f64) == 256
    out64, tot64 = _run_batch(geon, ctx, torch, f64)
    outmx, totmx = _run_batch(geon, ctx, torch, fmx)
    print(f"C4: {sum(c.width * c.height for _, c, _ in f64)} pixels; FP64 "
          f"{tot64['executed_ray_steps']} ray-steps in {tot64['device_ms']:.0f} ms, mixed "
          f"{totmx['executed_ray_steps']} in {totmx['device_ms']:.0f} ms "
          f"({totmx['fp64_rays']} rays re-traced in FP64)", flush=True)
    worst_mis = 0.0
    per = {}       # kind → [max angle, Σ angle, escaped pixels compared, above 1e-3 rad]
    for k, ((m, cam, _), o64, omx) in enumerate(zip(f64, out64, outmx)):
        c64, cmx = o64[1], omx[1]
        assert int((c64 < 0).sum()) == 0 and int((cmx < 0).sum()) == 0, k
        mis = float((c64 != cmx).double().mean())
        d64, dmx = o64[2].view(3, -1).T, omx[2].view(3, -1).T
        both = (c64 == 2) & (cmx == 2) & (d64.abs().amax(dim=1) > 0) & (dmx.abs().amax(dim=1) > 0)
        st = per.setdefault(type(m).__name__, [0.0, 0.0, 0, 0])
        if int(both.sum()):
            a, b = d64[both], dmx[both]
            ang = torch.atan2(torch.linalg.cross(a, b).norm(dim=1), (a * b).sum(dim=1))
            st[0] = max(st[0], float(ang.max()))
            st[1] += float(ang.sum())
            st[2] += int(both.sum())
            st[3] += int((ang > 1e-3).sum())
        worst_mis = max(worst_mis, mis)
        assert mis <= 1e-3, (k, type(m).__name__, cam.width, mis)
    print(f"C4 FP32-mixed vs FP64: worst per-frame class mismatch {worst_mis:.2e}", flush=True)
    for kind, (mx, sm, n, above) in per.items():
        print(f"C4 {kind}: escape angle max {mx:.3e} rad, mean {sm / max(1, n):.3e} rad, "
              f"{above} of {n} above 1e-3 rad", flush=True)
    above = sum(v[3] for v in per.values())
    assert above <= 1e-5 * sum(v[2] for v in per.values())
    assert max(v[0] for v in per.values()) < 0.05

    # 16 seeded frames alone equal their place in the batch
    rng = np.random.default_rng(11)
    pick = sorted(rng.choice(256, 16, replace=False).tolist())
    for frames, outs in ((f64, out64), (fmx, outmx)):
        alone, _ = _run_batch(geon, ctx, torch, [frames[i] for i in pick])
        for i, a in zip(pick, alone):
            for x, y in zip(a, outs[i]):
                assert torch.equal(x, y), i

    # seeded oracle pixels
    checked = 0
    for i in sorted(rng.choice(256, 8, replace=False).tolist()):
        m, cam, cfg = f64[i]
        om, ocam = _ometric(O, m), _ocam(O, cam)
        st, tet = O.tetrad(om, ocam.observer)
        idx = rng.choice(cam.width * cam.height, 32, replace=False)
        y0 = np.array([O.ray_state_sub(ocam, tet, int(p) % cam.width, int(p) // cam.width)
                       for p in idx])
        ref = O.trace_rays(om, y0, cfg.step, cfg.max_steps)
        tr = geon.trace_rays(m, y0, cfg.step, cfg.max_steps, ctx=ctx)
        cls64 = out64[i][1].cpu().numpy()[idx]
        dir64 = out64[i][2].view(3, -1).T.cpu().numpy()[idx]
        assert np.array_equal(cls64, tr["cls"]) and np.array_equal(dir64, tr["dir"])
        compare_rays(O, om, y0, tr, ref, cfg.step, cfg.max_steps, f"C4 frame {i} ({type(m).__name__})",
                     esc_r=m.escape_radius)
        clsmx = outmx[i][1].cpu().numpy()[idx]
        print(f"C4 frame {i}: FP32-mixed vs oracle class mismatches {(clsmx != ref['cls']).sum()}"
              f"/32", flush=True)
        checked += 32
    assert checked == 256


def _wait(srv, tickets, timeout=300.0):
    t0 = time.time()
    out = {}
    while len(out) < len(tickets):
        for t in tickets:
            if t not in out:
                r = srv.poll(t)
                if r is not None:
                    out[t] = r
        assert time.time() - t0 < timeout, "server did not finish"
        time.sleep(0.002)
    return out


def test_c5_poisson_stream_sampled_requests_equal_alone(ctx):
    import geon_amd as geon
    from geon_amd import workloads
    reqs = workloads.serving_stream(200, 400)
    rng = np.random.default_rng(3)
    renders = [i for i, r in enumerate(reqs) if r["kind"] == "render"]
    bisects = [i for i, r in enumerate(reqs) if r["kind"] == "bisect"]
    keep = set(rng.choice(renders, 16, replace=False).tolist()) | \
        set(rng.choice(bisects, 4, replace=False).tolist())
    srv = geon.Server(ctx, hbm_budget=8 << 30, quantum_steps=2048)
    tickets = {}
    t_start = time.perf_counter()
    i = 0
    while i < len(reqs):
        now = time.perf_counter() - t_start
        while i < len(reqs) and reqs[i]["t"] <= now:
            r = reqs[i]
            if r["kind"] == "render":
                tickets[i] = srv.submit_render(r["metric"], r["camera"], r["cfg"], keep=i in keep)
            else:
                tickets[i] = srv.submit_bisect(r["mass"], r["r_obs"], iters=r["iters"])
            i += 1
        time.sleep(0.0005)
    res = _wait(srv, list(tickets.values()))
    wall = time.perf_counter() - t_start
    assert all(v["status"] == 0 for v in res.values())
    lat = sorted(v["latency_ms"] for k, v in res.items())
    print(f"C5: {len(reqs)} requests at 200/s served in {wall:.2f} s; latency p50 "
          f"{np.percentile(lat, 50):.1f} ms, p99 {np.percentile(lat, 99):.1f} ms", flush=True)
    mixed = {k: (srv.fetch_render(tickets[k], reqs[k]["camera"].width, reqs[k]["camera"].height)
                 if reqs[k]["kind"] == "render" else srv.fetch_bisect(tickets[k], 64))
             for k in sorted(keep)}
    stats = {k: res[tickets[k]] for k in keep}
    srv.close()
    for k in sorted(keep):
        r = reqs[k]
        alone = geon.Server(ctx, hbm_budget=8 << 30, quantum_steps=2048)
        if r["kind"] == "render":
            t = alone.submit_render(r["metric"], r["camera"], r["cfg"], keep=True)
            ra = _wait(alone, [t])[t]
            out = alone.fetch_render(t, r["camera"].width, r["camera"].height)
            for x, y in zip(out, mixed[k]):
                assert np.array_equal(x, y), k
            for key in ("captured", "escaped", "diverged", "ray_steps", "fp64_rays"):
                assert ra["stats"][key] == stats[k]["stats"][key], (k, key)
        else:
            t = alone.submit_bisect(r["mass"], r["r_obs"], iters=r["iters"])
            _wait(alone, [t])
            a, b = alone.fetch_bisect(t, 64)
            assert np.array_equal(a, mixed[k][0]) and np.array_equal(b, mixed[k][1]), k
            assert np.abs(b / (3 * math.sqrt(3)) - 1.0).max() < 1e-6
        alone.close()
