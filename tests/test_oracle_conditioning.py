"""The oracle's planted tanh perturbation (conditioning tests only) is inert
at 0 and moves the Alcubierre metric by the last bits when set. (A uniform
one-ulp shift of both tanh of the shape function cancels in their difference
when the two values share a binade, so not every setting moves g at a given
point; the per-evaluation random settings do.)"""
import numpy as np

from oracle import oracle as O


def test_tanh_shift_is_planted_and_removable():
    om = O.metric(O.ALCUBIERRE, bubble_radius=2.0, bubble_thickness=0.5, ship_velocity=2.0)
    x = np.array([0.0, 2.3, 0.4, -0.1])           # inside the wall: tanh unsaturated
    base = O.g(om, x)
    try:
        moved = []
        for k in (-1, 1, 2, 3):
            O.set_tanh_ulp_shift(k)
            g = O.g(om, x)
            moved.append(not np.array_equal(g, base))
            rel = np.abs(g - base).max() / np.abs(base).max()
            assert rel < 1e-14                       # a last-bit change, no more
        assert any(moved[2:])
    finally:
        O.set_tanh_ulp_shift(0)
    assert np.array_equal(O.g(om, x), base)
