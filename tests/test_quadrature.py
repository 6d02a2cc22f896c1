"""Integration rules of the lens spread functions (the oracle's restatement
of cond/obj_lens_temporal_spatial_quadratures, src/cpu_fcns.hpp:1503-1626).

Pinned by: the Gauss-Hermite and generalized Gauss-Laguerre rules of numpy
and scipy (nodes 1e-12 absolute, weights 1e-10 relative up to 100 points);
and by what the mathematics fixes for the spread functions: a Gaussian
defocus spread and an exponential source spread are integrated exactly by
their rules (unit mass, second moments sigma^2 and 6 beta^2), an
exponential defocus spread converges to unit mass and second moment
2 beta^2, and the objective lens's rule has unit mass and second moment 1/2.
"""
import numpy as np
import pytest
from scipy.special import roots_genlaguerre

from multem_amd import configs
from oracle import oracle


@pytest.mark.parametrize("n", [2, 7, 20, 64, 100])
def test_gauss_hermite_rule(n):
    x, w = oracle.gauss_rule(6, n, 0.0, 0.0, 1.0)
    X, W = np.polynomial.hermite.hermgauss(n)
    assert np.abs(x - X).max() < 1e-12
    assert np.abs(w / W - 1).max() < 1e-10
    b, a = 0.37, 1.5          # |x - a|^0 exp(-b (x - a)^2): nodes a + t / sqrt(b), weights W / sqrt(b)
    x, w = oracle.gauss_rule(6, n, 0.0, a, b)
    assert np.abs(x - (a + X / np.sqrt(b))).max() < 1e-11
    assert np.abs(w / (W / np.sqrt(b)) - 1).max() < 1e-10


@pytest.mark.parametrize("n", [2, 6, 30, 100])
@pytest.mark.parametrize("alpha", [0.0, 1.0])
def test_gauss_laguerre_rule(n, alpha):
    b = 2.5
    x, w = oracle.gauss_rule(5, n, alpha, 0.0, b)
    X, W = roots_genlaguerre(n, alpha)
    assert np.abs(x / (X / b) - 1).max() < 1e-12
    assert np.abs(w / (W / b ** (alpha + 1)) - 1).max() < 1e-10


def test_one_point_rule():
    for kind in (5, 6):
        x, w = oracle.gauss_rule(kind, 1, 1.0, 0.0, 3.0)
        assert x[0] == 0.0 and w[0] == 1.0


def _cbed():
    p = configs.pt_cbed_1024(nconf=1, n=256, na=4, nb=6, nc=12, thick=[15.0])
    p.illumination_model = 4
    p.cond_lens_si_sigma = 0.2
    return p


def test_defocus_spread_gaussian_exact():
    p = _cbed()
    p.cond_lens_ti_a, p.cond_lens_ti_sigma, p.cond_lens_ti_beta, p.cond_lens_ti_npts = 1.0, 30.0, 0.0, 8
    x, w = oracle.lens_quadrature(p, 0)
    assert len(x) == 8
    assert abs(w.sum() - 1) < 1e-13
    assert abs((w * x * x).sum() / 30.0 ** 2 - 1) < 1e-13
    assert abs((w * x).sum()) < 1e-10


def test_defocus_spread_exponential_converges():
    p = _cbed()
    errs = []
    for n in (8, 16, 32):
        p.cond_lens_ti_a, p.cond_lens_ti_sigma, p.cond_lens_ti_beta, p.cond_lens_ti_npts = 0.0, 30.0, 20.0, n
        x, w = oracle.lens_quadrature(p, 0)
        errs.append((abs(w.sum() - 1), abs((w * x * x).sum() / (2 * 20.0 ** 2) - 1)))
    print("exponential defocus spread, |mass - 1| and second-moment error:", errs)
    assert errs[0][0] > errs[1][0] > errs[2][0] and errs[2][0] < 0.03
    assert errs[2][1] < 0.01


def test_source_spread():
    p = _cbed()
    p.cond_lens_si_rad_npts, p.cond_lens_si_azm_npts = 5, 6
    p.cond_lens_si_a, p.cond_lens_si_sigma, p.cond_lens_si_beta = 0.0, 0.2, 0.15   # exponential: exact
    x, y, w = oracle.lens_quadrature(p, 1)
    assert len(w) == 30
    assert abs(w.sum() - 1) < 1e-13
    assert abs((w * (x * x + y * y)).sum() / (6 * 0.15 ** 2) - 1) < 1e-12
    assert abs((w * x).sum()) < 1e-14 and abs((w * y).sum()) < 1e-14
    p.cond_lens_si_a, p.cond_lens_si_beta = 1.0, 0.0                                # Gaussian: approximate
    p.cond_lens_si_rad_npts = 8
    x, y, w = oracle.lens_quadrature(p, 1)
    assert abs(w.sum() - 1) < 2e-3
    assert abs((w * (x * x + y * y)).sum() / (2 * 0.2 ** 2) - 1) < 0.02


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_objective_defocus_rule(precision):
    p = configs.si_ewrs_256(nc=4)
    p.simulation_type = 32
    p.illumination_model = 4
    p.temporal_spatial_incoh = 2
    p.obj_lens_ti_npts, p.obj_lens_ti_sigma = 9, 20.0
    x, w = oracle.lens_quadrature(p, 2, precision=precision)
    tol = 1e-14 if precision == "f64" else 1e-6
    assert len(x) == 9
    assert abs(w.sum() - 1) < tol and abs((w * x * x).sum() - 0.5) < tol


def test_hrtem_spatial_full_integration_refused():
    p = configs.si_ewrs_256(nc=4)
    p.simulation_type = 32
    p.illumination_model = 4
    p.cond_lens_si_sigma = 0.01            # spatial incoherence requested
    p.temporal_spatial_incoh = 1
    with pytest.raises(RuntimeError):
        oracle.lens_quadrature(p, 2)
