"""ISTEM (12), CBEI (22), ED (31), PED (41) and HCTEM (42) on the GPU against
the CPU oracle, through the C ABI.

Tolerances (relative L2): 1e-4 for the FP32 path against the FP64 oracle,
as for the other runs; 1e-9 for the FP64 path.  The oracle's runs of these
types are pinned by identities with the pinned types in
tests/test_sim_types.py.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {1: 1e-4, 2: 1e-9}


def _case(st, precision, **kw):
    p = configs.si_ewrs_256(nc=3, n=128)
    p.system_conf.precision = precision
    p.simulation_type = st
    p.obj_lens_zero_defocus_type = 3
    p.cond_lens_outer_aper_ang = 15.0
    p.obj_lens_c_10, p.obj_lens_c_30, p.obj_lens_outer_aper_ang = -30.0, 0.5, 25.0
    p.illumination_model = 1
    p.scanning_type = 2
    p.scanning_periodic = 0
    p.scanning_ns = 3
    p.scanning_x0, p.scanning_y0 = 0.3 * p.spec_lx, 0.3 * p.spec_ly
    p.scanning_xe, p.scanning_ye = 0.6 * p.spec_lx, 0.6 * p.spec_ly
    p.pn_model = 3
    p.pn_nconf = 2
    p.pn_dim = 110
    p.pn_seed = 4321
    for k, v in kw.items():
        setattr(p, k, v)
    return p


CASES = {
    "istem": (12, dict(iw_type=2)),
    "istem_coherent": (12, dict(iw_type=2, pn_coh_contrib=1)),
    "istem_tilted": (12, dict(iw_type=2, theta=0.5, phi=30.0)),
    "istem_defocus_spread": (12, dict(iw_type=2, illumination_model=4, cond_lens_c_10=20.0,
                                      cond_lens_ti_sigma=30.0, cond_lens_ti_npts=3, cond_lens_si_sigma=0.0,
                                      temporal_spatial_incoh=2)),
    "cbei": (22, dict(iw_type=2, iw_x=np.array([9.0, 12.0]), iw_y=np.array([10.0, 8.0]))),
    "cbei_partial_coherent": (22, dict(iw_type=2, illumination_model=2)),
    "ed": (31, {}),
    "ed_coherent": (31, dict(pn_coh_contrib=1)),
    "ped": (41, dict(ped_theta=1.0, ped_nrot=6)),
    "hctem": (42, dict(hci_theta=0.8, hci_nrot=5)),
    "hctem_coherent": (42, dict(hci_theta=0.8, hci_nrot=4, pn_coh_contrib=1)),
}


def _outputs(o):
    return [a for a in (o.m2psi_tot, o.m2psi_coh) if a is not None]


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("name", sorted(CASES))
def test_sim_type_vs_oracle(name, precision):
    st, kw = CASES[name]
    p = _case(st, precision, **kw)
    if st == 22 and "iw_x" not in kw:
        p.iw_x = np.array([0.5 * p.spec_lx])
        p.iw_y = np.array([0.5 * p.spec_ly])
    out = engine.run_multislice(p)
    ref = oracle.run(p, "f64")
    assert out.output_type == ref.output_type
    assert np.allclose(out.x, ref.x) and np.allclose(out.y, ref.y)
    got, want = _outputs(out), _outputs(ref)
    assert len(got) == len(want) > 0
    for a, b in zip(got, want):
        for k in range(out.n_thk):
            e = rel_l2(a[k], b[k])
            print("%s (precision %d) thickness %d: rel L2 vs FP64 oracle %.3e" % (name, precision, k, e))
            assert e < TOL[precision]


@pytest.mark.parametrize("st,kw", [(12, dict(iw_type=2)), (41, dict(ped_theta=1.0, ped_nrot=6))])
def test_wave_chunks_do_not_change_the_result(st, kw):
    """Waves split over launches (a wave cap below the waves of one
    configuration: scan positions, azimuths) sum to the same output."""
    p = _case(st, 1, **kw)
    engine.set_pass_variants(0, 0, 0)
    try:
        a = engine.run_multislice(p).m2psi_tot.copy()
        engine.set_batch(4, True)
        b = engine.run_multislice(p).m2psi_tot.copy()
        plan = engine.last_plan()
    finally:
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)
    assert "waves_per_launch=4" in plan
    assert rel_l2(a, b) < 1e-6
