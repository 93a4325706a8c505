"""The EGG kernel decides most ellipsoid tests with a division-free estimate and evaluates the exact expression
(the oracle's and the reference's, EGT_CUDA.cu:136-170) only within 1e-5 of the boundary:

    exact     v  = (e*e + (f*f) / aa) + (g*g) / aa        (float32, each operation rounded)
    estimate  vf = e*e + (f*f + g*g) * fl(1 / aa)

Both sum non-negative terms, so each is within a few units in the last place of the real value and of each
other. This pins that bound (|vf - v| <= 1e-6 v, a tenth of the band) on float32 arithmetic emulated in numpy,
over random magnitudes and the alphas the path uses (0.65, 1e-8, 1)."""
import numpy as np
import pytest


def _f(x):
    return np.float32(x)


@pytest.mark.parametrize("alpha", [0.65, 1e-8, 1.0, 0.3])
def test_estimate_within_band(alpha):
    rng = np.random.default_rng(7)
    n = 200_000
    aa = _f(alpha) * _f(alpha)
    inv = _f(1.0) / aa
    scale = np.float32(10.0) ** rng.uniform(-5, 0, (n, 1)).astype(np.float32)
    e, f, g = (rng.standard_normal((n, 3)).astype(np.float32) * scale).T
    e2, f2, g2 = e * e, f * f, g * g
    v = (e2 + f2 / aa) + g2 / aa
    vf = e2 + (f2 + g2) * inv
    assert v.dtype == np.float32 and vf.dtype == np.float32
    ok = v > 0
    rel = np.abs(vf[ok].astype(np.float64) - v[ok]) / v[ok]
    assert rel.max() <= 1e-6


def test_estimate_decides_like_exact_outside_band():
    """Where the estimate is outside dd (1 -+ 1e-5) it gives the exact expression's answer, including values
    placed right at the band edges."""
    rng = np.random.default_rng(8)
    aa = _f(0.65) * _f(0.65)
    inv = _f(1.0) / aa
    n = 200_000
    e, f, g = (rng.standard_normal((n, 3)) * 1e-3).astype(np.float32).T
    e2, f2, g2 = e * e, f * f, g * g
    v = (e2 + f2 / aa) + g2 / aa
    vf = e2 + (f2 + g2) * inv
    # thresholds just around each value, at and beyond the band
    for rel in (-2e-5, -1.0001e-5, 1.0001e-5, 2e-5, 1e-3):
        dd = (vf.astype(np.float64) / (1.0 + rel)).astype(np.float32)
        lo = dd * _f(1.0 - 1e-5)
        hi = dd * _f(1.0 + 1e-5)
        decided = (vf < lo) | (vf > hi)
        fast = vf < lo
        exact = v < dd
        assert np.array_equal(fast[decided], exact[decided])
