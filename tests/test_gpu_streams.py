"""Device-pointer entry points on callers' streams.

Scratch the context keeps for "_dev" calls is held per stream, so calls
enqueued on different streams without any synchronisation in between must
each give the results they give alone (the host-buffer entry points, which
run on the context's stream, are the reference here, bit for bit).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import geon_amd as geon
    c = geon.Context(0)
    yield c
    c.close()


def _nec_dev(geon, ctx, metric, xs_dev, n, k, out_dev, stream):
    kk = (geon.C.c_double * 4)(*k)
    geon._abi.check(geon.lib().geon_nec_contraction_dev(
        ctx.handle, geon.C.byref(geon._cmetric(metric)), geon.C.c_void_p(xs_dev), n, kk,
        geon.C.c_void_p(out_dev), geon.C.c_void_p(stream)), "geon_nec_contraction_dev")


def test_nec_contraction_dev_on_two_streams(ctx):
    import torch

    import geon_amd as geon
    rng = np.random.default_rng(2)
    alc = geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5)
    mt = geon.MorrisThorne(throat_radius=1.0)
    xa = np.c_[np.zeros(1 << 16), rng.uniform(-4, 4, (1 << 16, 3))]
    xm = np.c_[np.zeros(4096), rng.uniform(-4, 4, 4096), rng.uniform(0.3, 2.8, 4096),
               np.zeros(4096)]
    ka, km = [1.0, 1.0, 0.0, 0.0], [1.0, -1.0, 0.0, 0.0]
    want_a = geon.nec_contraction(alc, xa, ka, ctx=ctx)
    want_m = geon.nec_contraction(mt, xm, km, ctx=ctx)
    da = torch.from_numpy(np.ascontiguousarray(xa.T)).cuda()
    dm = torch.from_numpy(np.ascontiguousarray(xm.T)).cuda()
    oa = torch.full((len(xa),), math.nan, dtype=torch.float64, device="cuda")
    om = torch.full((len(xm),), math.nan, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for _ in range(3):   # enqueue both, twice over, before waiting for either
        _nec_dev(geon, ctx, alc, da.data_ptr(), len(xa), ka, oa.data_ptr(), s1.cuda_stream)
        _nec_dev(geon, ctx, mt, dm.data_ptr(), len(xm), km, om.data_ptr(), s2.cuda_stream)
    s1.synchronize()
    s2.synchronize()
    assert np.array_equal(oa.cpu().numpy(), want_a)
    assert np.array_equal(om.cpu().numpy(), want_m)


def test_render_dev_on_two_streams(ctx):
    import torch

    import geon_amd as geon
    jobs = [
        (geon.Schwarzschild(mass=1.0, horizon_pad=1.005, escape_radius=100.0),
         geon.Camera((0.0, 30.0, math.pi / 2, 0.0), 60.0, 192, 160),
         geon.RenderConfig(step=0.05, max_steps=4000)),
        (geon.MorrisThorne(throat_radius=1.0, escape_radius=60.0),
         geon.Camera((0.0, 15.0, math.pi / 2, 0.0), 60.0, 96, 128),
         geon.RenderConfig(step=0.05, max_steps=3000)),
        (geon.Alcubierre(bubble_radius=2.0, bubble_thickness=0.125, ship_velocity=0.5,
                         escape_radius=60.0),
         geon.Camera((0.0, 10.0, 0.0, 0.0), 60.0, 128, 96),
         geon.RenderConfig(step=0.05, max_steps=3000, precision=geon.PREC_MIXED)),
    ]
    want = [geon.render(m, cam, cfg, ctx=ctx) for m, cam, cfg in jobs]
    streams = [torch.cuda.Stream() for _ in jobs]
    outs = []
    for (m, cam, cfg), st in zip(jobs, streams):
        n = cam.width * cam.height
        outs.append((torch.zeros(3 * n, dtype=torch.float32, device="cuda"),
                     torch.zeros(n, dtype=torch.int32, device="cuda"),
                     torch.zeros(3 * n, dtype=torch.float64, device="cuda")))
    torch.cuda.synchronize()
    # no statistics requested: each call enqueues and returns
    for (m, cam, cfg), st, o in zip(jobs, streams, outs):
        geon.render_dev(m, cam, cfg, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(),
                        stream_ptr=st.cuda_stream, want_stats=False, ctx=ctx)
    for st in streams:
        st.synchronize()
    for (rgb, stats, cls, d), o in zip(want, outs):
        assert np.array_equal(o[1].cpu().numpy(), cls)
        assert np.array_equal(o[2].cpu().numpy().reshape(3, -1).T, d)
        assert np.array_equal(o[0].cpu().numpy().reshape(-1, 3), rgb)
    # with statistics: each call waits for its own stream only
    for (m, cam, cfg), st, o, w in zip(jobs, streams, outs, want):
        s = geon.render_dev(m, cam, cfg, o[0].data_ptr(), o[1].data_ptr(), o[2].data_ptr(),
                            stream_ptr=st.cuda_stream, want_stats=True, ctx=ctx)
        for key in ("captured", "escaped", "diverged", "ray_steps", "max_null_residual"):
            assert s[key] == w[1][key], key
