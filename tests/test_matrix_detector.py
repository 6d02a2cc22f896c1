"""Matrix detectors (detector_type 3, detector.matrix(i).fR;
src/wave_function.cuh:125-129) in the oracle.

Pinned by the ring detector: a sensitivity map that is 1 on the pixels a
ring keeps (g2_in <= g^2 < g2_out) and 0 elsewhere gives the ring's image;
the map is linear (scaling it scales the image).  Maps are centred, as the
reference's example builds them (example_MULTEM_STEM_matrix_detector.m:115-130:
meshgrid((-nxh:nxh-1) dgx, (-nyh:nyh-1) dgy)), and the engine takes them to
the natural Fourier order with fft2_shift (src/wave_function.cuh:53-61): a
map in the natural order is a different detector.  The radial detector,
which the reference integrates to zero, is refused.
"""
import numpy as np
import pytest

from multem_amd import configs
from oracle import oracle
from tests.conftest import rel_l2


def ring_map(p, inner, outer):
    """The ring's pixel mask as the reference's example builds it: centred,
    gx = (-nxh:nxh-1) dgx, gy = (-nyh:nyh-1) dgy, MATLAB ny x nx; g^2
    evaluated as the FP64 oracle evaluates it."""
    lam = oracle.lib("f64").orc_lambda(float(p.E_0))
    lx, ly = float(p.spec_lx), float(p.spec_ly)
    dgx, dgy = 1.0 / lx, 1.0 / ly
    gx = (np.arange(p.nx) - p.nx // 2) * dgx
    gy = (np.arange(p.ny) - p.ny // 2) * dgy
    g2 = gx[:, None] * gx[:, None] + gy[None, :] * gy[None, :]
    g2i = (np.sin(inner * 1e-3) / lam) ** 2
    g2o = (np.sin(outer * 1e-3) / lam) ** 2
    return ((g2i <= g2) & (g2 < g2o)).astype(np.float64).T      # MATLAB ny x nx


@pytest.fixture(scope="module")
def stem():
    p = configs.au_stem_512(ns=3, nconf=1, n=128, ncells=2, nc=3)
    p.detector_cir = [(20.0, 60.0), (80.0, 200.0)]
    return p


def test_ring_map_is_the_ring(stem):
    ref = oracle.run(stem, "f64").image_tot
    q = stem.copy()
    q.detector_type = 3
    q.detector_fR = [ring_map(q, a, b) for a, b in stem.detector_cir]
    got = oracle.run(q, "f64").image_tot
    e = rel_l2(got, ref)
    print("ring maps vs rings: rel L2 %.2e" % e)
    assert e < 1e-13


def test_map_is_linear(stem):
    q = stem.copy()
    q.detector_type = 3
    rng = np.random.default_rng(3)
    f = rng.random((q.ny, q.nx))
    q.detector_fR = [f, 2.5 * f]
    img = oracle.run(q, "f64").image_tot
    assert rel_l2(img[:, 1], 2.5 * img[:, 0]) < 1e-14


def test_radial_detector_refused(stem):
    q = stem.copy()
    q.detector_type = 2
    with pytest.raises(RuntimeError):
        oracle.run(q, "f64")


def test_natural_order_map_is_another_detector(stem):
    """The quadrant swap matters: the ring's map in the natural Fourier
    order (ring at the grid corners after the engine's fft2_shift) is not
    the ring."""
    ref = oracle.run(stem, "f64").image_tot
    q = stem.copy()
    q.detector_type = 3
    q.detector_fR = [np.fft.ifftshift(ring_map(q, a, b)) for a, b in stem.detector_cir]
    got = oracle.run(q, "f64").image_tot
    assert rel_l2(got, ref) > 0.1
