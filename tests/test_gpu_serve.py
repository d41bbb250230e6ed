"""Continuous-batching server on the GPU.

- Every request's result is the same alone and in a mix of concurrently
  admitted requests of other metrics, sizes and precisions (bit for bit:
  per-ray arithmetic never depends on slots or neighbours).
- Shadow-bisection requests run as k-section rounds: with photon-sphere
  decisions they converge on b_c = 3√3 M; with full traces they reproduce
  validate_shadow's own fixed-step result (validate_shadow.rs:45-94).
- HBM admission: with a small budget requests wait for memory and still all
  complete; a request larger than the budget is refused with an error.
"""
import math
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import geon_amd as geon
    c = geon.Context(0)
    yield c
    c.close()


def _requests():
    import geon_amd as geon
    return [
        (geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0),
         geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 128, 96),
         geon.RenderConfig(step=0.05, max_steps=4000, precision=geon.PREC_MIXED)),
        (geon.MorrisThorne(throat_radius=1.3, escape_radius=60.0),
         geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, 80, 80),
         geon.RenderConfig(step=0.05, max_steps=3000)),
        (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.7,
                         escape_radius=60.0),
         geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 64, 64),
         geon.RenderConfig(step=0.05, max_steps=3000, precision=geon.PREC_MIXED)),
        (geon.Schwarzschild(mass=1.4, horizon_pad=1.005, escape_radius=90.0),
         geon.Camera((0.0, 22.0, 1.3, 0.0), 55.0, 72, 72),
         geon.RenderConfig(step=0.05, max_steps=4000, samples_per_axis=2)),
    ]


def _wait(srv, tickets, timeout=120.0):
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


def test_serve_alone_vs_mix(ctx):
    import geon_amd as geon
    reqs = _requests()
    alone = []
    for m, cam, cfg in reqs:
        srv = geon.Server(ctx, hbm_budget=1 << 30, quantum_steps=128)
        t = srv.submit_render(m, cam, cfg, keep=True)
        res = _wait(srv, [t])[t]
        assert res["status"] == 0
        alone.append((srv.fetch_render(t, cam.width, cam.height), res))
        srv.close()
    srv = geon.Server(ctx, hbm_budget=1 << 30, quantum_steps=128)
    tickets = []
    for i, (m, cam, cfg) in enumerate(reqs):       # staggered arrivals
        tickets.append(srv.submit_render(m, cam, cfg, keep=True))
        time.sleep(0.003 * i)
    done = _wait(srv, tickets)
    for (m, cam, cfg), t, (out_a, res_a) in zip(reqs, tickets, alone):
        res = done[t]
        assert res["status"] == 0
        out = srv.fetch_render(t, cam.width, cam.height)
        for x, y in zip(out, out_a):
            assert np.array_equal(x, y)
        for k in ("captured", "escaped", "diverged", "ray_steps", "fp64_rays"):
            assert res["stats"][k] == res_a["stats"][k]
    # the served renders equal the batch path
    for (m, cam, cfg), t in zip(reqs, tickets):
        rgb, st, cls, d = geon.render(m, cam, cfg, ctx=ctx)
        out = srv.fetch_render(t, cam.width, cam.height)
        assert np.array_equal(out[1], cls) and np.array_equal(out[2], d)
    c = srv.counters()
    assert c["done"] == len(reqs) and c["quanta"] > 1
    srv.close()


def test_serve_bisection(ctx, O):
    import geon_amd as geon
    srv = geon.Server(ctx, hbm_budget=1 << 28, quantum_steps=512)
    r_obs = np.linspace(10.0, 50.0, 64)
    t = srv.submit_bisect(np.ones(64), r_obs, step=0.01, max_steps=200_000, iters=40)
    t2 = srv.submit_bisect([1.0], [20.0], step=0.01, max_steps=200_000, iters=60,
                           full_trace=True)
    res = _wait(srv, [t, t2], timeout=300.0)
    assert res[t]["status"] == 0 and res[t2]["status"] == 0
    a, b = srv.fetch_bisect(t, 64)
    # photon-sphere decisions converge on the analytic b_c = 3√3 M
    assert np.abs(b / (3 * math.sqrt(3)) - 1.0).max() < 1e-6
    # full traces reproduce validate_shadow (README.md:43-45) through the k-section rounds
    a2, b2 = srv.fetch_bisect(t2, 1)
    a_ref, b_ref = O.validate_shadow(1.0, 20.0, 0.01, 200_000, 60)
    assert abs(b2[0] - b_ref) < 1e-9
    assert f"{b2[0]:.6f}" == "5.196151"
    print(f"bisection: 64×40 latency {res[t]['latency_ms']:.1f} ms, ray-steps {res[t]['ray_steps']}; "
          f"full-trace 1×60 latency {res[t2]['latency_ms']:.1f} ms")
    srv.close()


def test_serve_admission_by_hbm(ctx):
    import geon_amd as geon
    m = geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=60.0)
    cfg = geon.RenderConfig(step=0.05, max_steps=2000, precision=geon.PREC_MIXED)
    cam = geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 128, 128)   # 655 KB of ray buffers
    srv = geon.Server(ctx, hbm_budget=1 << 21, quantum_steps=256)        # room for ~3 at once
    tickets = [srv.submit_render(m, cam, cfg) for _ in range(10)]
    res = _wait(srv, tickets)
    assert all(r["status"] == 0 for r in res.values())
    assert max(r["queue_ms"] for r in res.values()) > 0.0
    steps = {r["stats"]["ray_steps"] for r in res.values()}
    assert len(steps) == 1                                  # identical requests, identical work
    big = geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 1024, 1024)
    tb = srv.submit_render(m, big, cfg)
    rb = _wait(srv, [tb])[tb]
    assert rb["status"] == geon._abi.E_ADMISSION
    srv.close()


def test_serve_rejects_bad_requests(ctx):
    import geon_amd as geon
    srv = geon.Server(ctx, hbm_budget=1 << 24)
    with pytest.raises(geon.GeonError) as e:
        srv.submit_render(geon.Schwarzschild(), geon.Camera((0.0, 1.5, 1.0, 0.0), 60.0, 8, 8),
                          geon.RenderConfig())
    assert e.value.status == geon._abi.E_NOT_TIMELIKE
    with pytest.raises(geon.GeonError):
        srv.submit_bisect([1.0], [1.5])
    with pytest.raises(geon.GeonError) as e:
        srv.poll(987654)
    assert e.value.status == geon._abi.E_NOT_FOUND
    srv.close()
