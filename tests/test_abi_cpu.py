"""C ABI checks that need no GPU: the library loads, exports every symbol the
public header declares, and its host-side functions (defaults, Gram–Schmidt
tetrad, camera ray, tonemap) agree bit for bit with the oracle."""
import math
import os
import re

import numpy as np
import pytest

from tests.conftest import ROOT


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "geon_hip.h"), encoding="utf-8").read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(geon_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header(O):
    import geon_amd as geon
    L = geon.lib()
    syms = _header_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    declared = {name for name, _, _ in geon._abi.SIGNATURES}
    assert set(syms) == declared


def test_code_object_present():
    import geon_amd as geon
    assert os.path.exists(geon._abi.HSACO_PATH)


def test_metric_defaults_match_reference(O):
    import ctypes as C
    import geon_amd as geon
    for kind in range(4):
        a = geon._abi.Metric()
        assert geon.lib().geon_metric_default(kind, C.byref(a)) == 0
        b = O.metric(kind)
        assert bytes(a) == bytes(b)
    assert geon.lib().geon_metric_default(7, C.byref(a)) == geon._abi.E_INVALID
    # the dataclasses carry the same defaults
    for cls_, kind in [(geon.Minkowski, 0), (geon.Schwarzschild, 1), (geon.Alcubierre, 2),
                       (geon.MorrisThorne, 3)]:
        assert bytes(cls_().to_c()) == bytes(O.metric(kind))


@pytest.mark.parametrize("kind,kw,obs", [
    (1, dict(mass=1.0), (0.0, 30.0, math.pi / 2, 0.0)),
    (2, dict(bubble_thickness=0.125, ship_velocity=0.5), (0.0, 1.0, 0.3, 0.1)),
    (3, dict(throat_radius=1.0), (0.0, 15.0, math.pi / 2, 0.0)),
    (0, {}, (0.0, 20.0, 0.0, 0.0)),
])
def test_tetrad_and_rays_bitwise(O, kind, kw, obs):
    import geon_amd as geon
    om = O.metric(kind, **kw)
    st, ref = O.tetrad(om, obs)
    tet = geon.tetrad(om, obs)
    assert np.array_equal(tet, ref)
    import ctypes as C
    cam = O.camera(obs, 60.0, 7, 5)
    gcam = geon.Camera(obs, 60.0, 7, 5).to_c()
    for (i, j, du, dv) in [(0, 0, 0.5, 0.5), (6, 4, 0.25, 0.75), (3, 2, 0.5, 0.5)]:
        y = np.zeros(8)
        t = np.ascontiguousarray(tet.reshape(16))
        assert geon.lib().geon_ray_state(C.byref(gcam), t.ctypes.data_as(C.c_void_p), i, j, du, dv,
                                         y.ctypes.data_as(C.c_void_p)) == 0
        assert np.array_equal(y, O.ray_state_sub(cam, ref, i, j, du, dv))


def test_tetrad_errors():
    import geon_amd as geon
    with pytest.raises(geon.GeonError) as e:
        geon.tetrad(geon.Schwarzschild(), (0.0, 1.5, math.pi / 2, 0.0))
    assert e.value.status == geon._abi.E_NOT_TIMELIKE
    with pytest.raises(geon.GeonError) as e:
        geon.tetrad(geon.Schwarzschild(), (0.0, 10.0, 0.0, 0.0))
    assert e.value.status == geon._abi.E_GRAM_SCHMIDT


def test_tonemap_matches_oracle(O):
    import geon_amd as geon
    rng = np.random.default_rng(5)
    px = rng.uniform(-0.5, 1.5, 3000).astype(np.float32)
    assert np.array_equal(geon.tonemap_u8(px), O.tonemap_u8(px))


def test_write_png_pixels(O, tmp_path):
    """write_png: a valid 8-bit RGB PNG whose decoded rows are the oracle's
    tonemapped bytes, top row first; a mis-sized buffer is refused."""
    import struct
    import zlib
    import geon_amd as geon
    rng = np.random.default_rng(2)
    w, h = 7, 5
    px = np.concatenate([rng.uniform(-0.2, 1.2, 3 * w * h - 3), [0.0, 1.0, 0.5]]).astype(np.float32)
    path = tmp_path / "f.png"
    geon.write_png(path, px, w, h)
    data = path.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    chunks, pos = {}, 8
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag, body = data[pos + 4:pos + 8], data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(tag + body)
        chunks[tag] = chunks.get(tag, b"") + body
        pos += 12 + n
    assert struct.unpack(">IIBBBBB", chunks[b"IHDR"]) == (w, h, 8, 2, 0, 0, 0)
    raw = np.frombuffer(zlib.decompress(chunks[b"IDAT"]), np.uint8).reshape(h, 1 + 3 * w)
    assert (raw[:, 0] == 0).all()
    assert np.array_equal(raw[:, 1:].reshape(-1), O.tonemap_u8(px))
    with pytest.raises(ValueError):
        geon.write_png(tmp_path / "g.png", px[:-1], w, h)


def test_context_without_gpu_fails_cleanly():
    """On a machine with no visible device the context reports an error code."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import geon_amd as geon
    with pytest.raises(geon.GeonError):
        geon.Context(0)
