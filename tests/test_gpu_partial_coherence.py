"""Fully integrated partial coherence (illumination_model 4) against the
CPU oracle, through the C ABI: the condenser's defocus spread in STEM
(src/tem_simulation.cuh:138-178), its defocus and source spread in CBED
(:207-261), and the objective lens's defocus spread in HRTEM
(src/microscope_effects.cuh:150-202).

Tolerances (relative L2): 1e-4 for the FP32 path against the FP64 oracle,
as for the coherent runs; 1e-9 for the FP64 path.  The integration rules
themselves are pinned in tests/test_quadrature.py.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _stem(precision):
    p = configs.au_stem_512(ns=4, nconf=2, n=256, ncells=4, nc=4)
    p.system_conf.precision = precision
    p.illumination_model = 4
    p.cond_lens_c_10 = 40.0
    p.cond_lens_ti_a, p.cond_lens_ti_sigma, p.cond_lens_ti_beta, p.cond_lens_ti_npts = 0.8, 25.0, 10.0, 5
    return p


def _cbed(precision):
    p = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=10, thick=[12.0, 25.0])
    p.system_conf.precision = precision
    p.illumination_model = 4
    p.cond_lens_ti_a, p.cond_lens_ti_sigma, p.cond_lens_ti_beta, p.cond_lens_ti_npts = 1.0, 30.0, 0.0, 3
    p.cond_lens_si_a, p.cond_lens_si_sigma, p.cond_lens_si_beta = 0.8, 0.3, 0.1
    p.cond_lens_si_rad_npts, p.cond_lens_si_azm_npts = 2, 4
    return p


def _hrtem(precision):
    p = configs.si_ewrs_256(nc=4)
    p.system_conf.precision = precision
    p.simulation_type = 32
    p.illumination_model = 4
    p.temporal_spatial_incoh = 2
    p.obj_lens_c_10, p.obj_lens_c_30 = -55.0, 0.8
    p.obj_lens_outer_aper_ang = 25.0
    p.obj_lens_ti_sigma, p.obj_lens_ti_npts = 30.0, 7
    return p


TOL = {1: 1e-4, 2: 1e-9}


@pytest.mark.parametrize("precision", [1, 2])
def test_stem_defocus_spread(precision):
    p = _stem(precision)
    out = engine.run_multislice(p)
    assert "defocus_points=5" in engine.last_plan()
    ref = oracle.run(p, "f64")
    e = rel_l2(out.image_tot, ref.image_tot)
    coh = p.copy()
    coh.illumination_model = 1
    e_coh = rel_l2(oracle.run(coh, "f64").image_tot, ref.image_tot)
    print("STEM defocus spread (precision %d): rel L2 vs FP64 oracle %.3e; the coherent image differs by %.3e"
          % (precision, e, e_coh))
    assert e < TOL[precision]
    assert e_coh > 100 * TOL[1]     # the integration changes the image


def test_stem_one_defocus_point_is_the_coherent_run():
    p = _stem(1)
    p.cond_lens_ti_npts = 1
    q = p.copy()
    q.illumination_model = 1
    engine.set_pass_variants(0, 0, 0)
    try:
        a = engine.run_multislice(p).image_tot.copy()
        b = engine.run_multislice(q).image_tot.copy()
    finally:
        engine.set_pass_variants(-1, -1, -1)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("precision", [1, 2])
def test_cbed_source_and_defocus_spread(precision):
    p = _cbed(precision)
    out = engine.run_multislice(p)
    assert "waves_per_config=24" in engine.last_plan()
    ref = oracle.run(p, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        print("CBED partial coherence (precision %d) thickness %d: rel L2 vs FP64 oracle %.3e" % (precision, k, e))
        assert e < TOL[precision]


@pytest.mark.parametrize("precision", [1, 2])
def test_hrtem_defocus_spread(precision):
    p = _hrtem(precision)
    out = engine.run_multislice(p)
    assert "ctf_defocus_points=7" in engine.last_plan()
    ref = oracle.run(p, "f64")
    e = rel_l2(out.m2psi_tot, ref.m2psi_tot)
    print("HRTEM defocus spread (precision %d): rel L2 vs FP64 oracle %.3e" % (precision, e))
    assert e < TOL[precision]


def test_hrtem_spatial_full_integration_refused():
    p = _hrtem(1)
    p.temporal_spatial_incoh = 1
    p.cond_lens_si_sigma = 0.01
    with pytest.raises(engine.EngineError) as ei:
        engine.run_multislice(p)
    assert ei.value.status == 2
