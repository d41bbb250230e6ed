"""Seeded synthetic workloads of the benchmark configurations.

Each generator returns plain descriptions (metric, camera, render config or
bisection parameters) so that the same request stream can be sent to the
server, to the batch scheduler and to the plain FP64 kernel.
"""
from __future__ import annotations

import math

import numpy as np

from . import (PREC_FP64, PREC_MIXED, Alcubierre, Camera, MorrisThorne, RenderConfig,
               Schwarzschild)


def random_metric_camera(rng, n, r_cam=None, escape=80.0):
    """Metric uniform over {Schwarzschild M∈[0.5,2], Alcubierre v∈[0.1,0.9] (R=2,
    σ=8), Morris–Thorne b0∈[0.5,2]}; camera on the +x¹ axis looking inward."""
    k = int(rng.integers(3))
    if k == 0:
        m = Schwarzschild(mass=float(rng.uniform(0.5, 2.0)), horizon_pad=1.005,
                          escape_radius=escape)
        r = (r_cam if r_cam is not None else 30.0) * m.mass
        cam = Camera((0.0, r, math.pi / 2, 0.0), 60.0, n, n)
    elif k == 1:
        m = Alcubierre(bubble_radius=2.0, bubble_thickness=0.125,
                       ship_velocity=float(rng.uniform(0.1, 0.9)), escape_radius=escape)
        cam = Camera((0.0, r_cam if r_cam is not None else 10.0, 0.0, 0.0), 60.0, n, n)
    else:
        m = MorrisThorne(throat_radius=float(rng.uniform(0.5, 2.0)), escape_radius=escape)
        cam = Camera((0.0, r_cam if r_cam is not None else 15.0, math.pi / 2, 0.0), 60.0, n, n)
    return m, cam


def ragged_frames(seed=1234, n_frames=256, lo=64, hi=2048, max_steps=5000,
                  precision=PREC_MIXED):
    """256 frames, side log-uniform in [64, 2048], random metric; escape radius 80
    so that per-ray step counts span ~50–5000 at step 0.05."""
    rng = np.random.default_rng(seed)
    frames = []
    for _ in range(n_frames):
        n = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        m, cam = random_metric_camera(rng, n)
        frames.append((m, cam, RenderConfig(step=0.05, max_steps=max_steps, precision=precision)))
    return frames


def serving_stream(rate, n_requests, seed=7, precision=PREC_MIXED):
    """Poisson arrivals at `rate` req/s: 80% renders (side log-uniform 32..2048,
    random metric, camera r∈[10,50]), 20% shadow bisections (64 rays × 40
    halvings, M = 1, r_obs∈[10,50])."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n_requests):
        t += float(rng.exponential(1.0 / rate))
        if rng.uniform() < 0.8:
            n = int(round(math.exp(rng.uniform(math.log(32), math.log(2048)))))
            m, cam = random_metric_camera(rng, n, r_cam=float(rng.uniform(10.0, 50.0)),
                                          escape=100.0)
            out.append(dict(t=t, kind="render", metric=m, camera=cam,
                            cfg=RenderConfig(step=0.05, max_steps=5000, precision=precision)))
        else:
            out.append(dict(t=t, kind="bisect", mass=np.ones(64),
                            r_obs=rng.uniform(10.0, 50.0, 64), iters=40))
    return out
