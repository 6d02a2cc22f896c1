"""The potential kernel's knot lookup (knot_interval in kernels.hip) against
the reference's 7-step binary search over the 128 spline knots
(eval_cubic_poly, src/cgpu_fcns.cuh:1496-1566), on the oracle's tables.

The kernel estimates the interval from the hardware log2 of R^2 (the knots
are log-spaced in R) and corrects it by one comparison on each side; the
index must equal the binary search's for every R^2, including each knot and
its neighbouring floats, and also when the log2 is off by several ulps (the
hardware log2 is not correctly rounded).  Restated here in numpy with the
kernel's float32 arithmetic; the GPU tests check the resulting potentials
bit for bit."""
import numpy as np
import pytest

from oracle import oracle


def _binary_search(tb, R2):
    i0 = np.zeros(R2.shape, int)
    ie = np.full(R2.shape, 127)
    for _ in range(7):
        im = (i0 + ie) >> 1
        lt = R2 < tb[im]
        ie = np.where(lt, im, ie)
        i0 = np.where(lt, i0, im)
    return i0


def _knot_interval(tb, R2, ulps):
    f32 = np.float32
    l0 = np.log2(f32(tb[0]), dtype=f32)
    c = f32(127) / f32(np.log2(f32(tb[127]), dtype=f32) - l0)
    with np.errstate(divide="ignore"):
        lg = np.log2(R2.astype(f32), dtype=f32)
    lg = np.where(np.isfinite(lg), lg + ulps * np.spacing(np.abs(lg)).astype(f32), lg)
    f = np.fmin(np.fmax((lg - l0) * c, f32(0)), f32(126))
    g = f.astype(int)
    dn = (g > 0) & (R2 < tb[g])
    up = ~dn & (g < 126) & (R2 >= tb[g + 1])
    return g - dn + up


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_knot_interval_equals_binary_search(precision):
    dt = np.float32 if precision == "f32" else np.float64
    rng = np.random.default_rng(7)
    n = 0
    for Z in (1, 6, 8, 14, 22, 29, 38, 47, 56, 74, 78, 79, 82, 92, 103):
        for dR in (0.01, 0.0383, 0.1):
            tb = oracle.atom_table(Z, dR, precision)["R2"].astype(dt)
            assert np.all(np.diff(tb) > 0)
            R2 = np.concatenate([
                np.array([0.0], dt), tb, np.nextafter(tb, dt(0)), np.nextafter(tb, dt(np.inf)),
                rng.uniform(0, tb[127] * 1.01, 2000).astype(dt),
                np.exp(rng.uniform(np.log(tb[0]) - 3, np.log(tb[127]), 2000)).astype(dt)])
            ref = _binary_search(tb, R2)
            for ulps in (-8, 0, 8):
                assert np.array_equal(_knot_interval(tb, R2, ulps), ref), (Z, dR, ulps)
            n += R2.size
    assert n > 100000


def test_knot_interval_catches_a_wrong_guess():
    """Without the correction, the guess alone disagrees with the binary
    search (so the test above is not passing trivially)."""
    tb = oracle.atom_table(79, 0.0383, "f32")["R2"].astype(np.float32)
    R2 = tb.copy()
    f32 = np.float32
    l0 = np.log2(tb[0], dtype=f32)
    c = f32(127) / f32(np.log2(tb[127], dtype=f32) - l0)
    guess = np.clip(((np.log2(R2 - np.spacing(R2), dtype=f32) - l0) * c).astype(int), 0, 126)
    assert not np.array_equal(guess, _binary_search(tb, R2 - np.spacing(R2)))
