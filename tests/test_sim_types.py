"""Pins of the oracle's ISTEM (12), CBEI (22), ED (31), PED (41) and HCTEM (42)
runs.  Each new simulation type differs from one already pinned only in its
loop over waves (Multislice::STEM_ISTEM, CBED_CBEI, ED_HRTEM, PED_HCTEM,
src/tem_simulation.cuh:100-301) or in where its output goes
(set_m2psi_tot_psi_coh, src/wave_function.cuh:135-186), so each is pinned by
an identity with those runs: the same physics reached by another loop.
"""
import numpy as np
import pytest

from multem_amd import configs
from oracle import oracle
from tests.conftest import rel_l2


def _case(st, **kw):
    p = configs.si_ewrs_256(nc=3, n=128)
    p.simulation_type = st
    p.obj_lens_zero_defocus_type = 3
    p.cond_lens_outer_aper_ang = 15.0
    p.obj_lens_c_10, p.obj_lens_c_30, p.obj_lens_outer_aper_ang = -30.0, 0.5, 25.0
    p.illumination_model = 1
    p.scanning_type = 2
    p.scanning_periodic = 0
    p.scanning_ns = 2
    p.scanning_x0, p.scanning_y0 = 0.3 * p.spec_lx, 0.3 * p.spec_ly
    p.scanning_xe, p.scanning_ye = 0.6 * p.spec_lx, 0.6 * p.spec_ly
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def _frozen(p, nconf=2):
    p.pn_model = 3
    p.pn_nconf = nconf
    p.pn_dim = 110
    p.pn_seed = 1234
    return p


def test_ed_is_the_ewfs_diffraction_pattern():
    """ED runs the EWFS loop (ED_HRTEM -> EWFS_EWRS) and accumulates
    |psi(g)|^2: with phonons it is EWFS's m2psi_tot."""
    ed = oracle.run(_frozen(_case(31)), "f64")
    ew = oracle.run(_frozen(_case(51)), "f64")
    assert ed.output_type == 4 and ew.output_type == 5
    assert rel_l2(ed.m2psi_tot, ew.m2psi_tot) < 1e-13
    assert abs(ed.m2psi_tot[0].sum() - 1.0) < 1e-9      # flux of a plane wave, no aperture


def test_cbei_with_an_identity_lens_is_the_convergent_exit_wave():
    """With no aberrations and an aperture beyond the grid's largest g the
    objective lens is the identity, so the CBEI image is |psi|^2 of the
    convergent-beam exit wave (EWRS with iw_type 2)."""
    kw = dict(iw_type=2, obj_lens_c_10=0.0, obj_lens_c_30=0.0, obj_lens_outer_aper_ang=400.0)
    cbei = oracle.run(_frozen(_case(22, **kw)), "f64")
    ewrs = oracle.run(_frozen(_case(52, **kw)), "f64")
    assert cbei.output_type == 4
    assert rel_l2(cbei.m2psi_tot, ewrs.m2psi_tot) < 1e-12


def test_cbei_equals_hrtem_of_the_same_incident_wave():
    """CBEI and HRTEM share the objective lens; given the convergent probe
    as a user-defined incident wave, HRTEM gives the CBEI image."""
    p = _case(22, iw_type=2)
    p.iw_x = np.array([0.5 * p.spec_lx])
    p.iw_y = np.array([0.5 * p.spec_ly])
    cbei = oracle.run(p, "f64").m2psi_tot
    psi0 = oracle.incident_wave(p, "f64")
    h = _case(32, iw_type=3)
    h.iw_psi = psi0.T                   # MATLAB's ny x nx view
    h.iw_x, h.iw_y = p.iw_x.copy(), p.iw_y.copy()
    hr = oracle.run(h, "f64").m2psi_tot
    assert rel_l2(cbei, hr) < 1e-9


def test_istem_is_the_sum_of_cbei_images_over_the_scan():
    """ISTEM accumulates the image of every scan position with the phonon
    weight alone (STEM_ISTEM, src/tem_simulation.cuh:179-198): the sum of
    CBEI images of a probe at each position."""
    p = _frozen(_case(12, iw_type=2))
    ist = oracle.run(p, "f64")
    assert ist.output_type == 4
    n = p.scanning_ns
    xs = [p.scanning_x0 + i * (p.scanning_xe - p.scanning_x0) / (n - 1) for i in range(n)]
    ys = [p.scanning_y0 + i * (p.scanning_ye - p.scanning_y0) / (n - 1) for i in range(n)]
    acc = np.zeros_like(ist.m2psi_tot)
    for x in xs:
        for y in ys:
            q = _frozen(_case(22, iw_type=2))
            q.iw_x, q.iw_y = np.array([x]), np.array([y])
            acc += oracle.run(q, "f64").m2psi_tot
    assert rel_l2(ist.m2psi_tot, acc) < 1e-9


def test_istem_defocus_integration_is_the_weighted_sum_over_defocus():
    """Fully integrated ISTEM integrates the condenser defocus spread only:
    sum_t w_t ISTEM(c_10 + x_t), with the nodes and weights of the
    condenser rule (no objective focal spread, so the objective lens's own
    integration is the coherent lens)."""
    kw = dict(iw_type=2, cond_lens_c_10=20.0, cond_lens_ti_sigma=30.0, cond_lens_ti_npts=4,
              cond_lens_si_sigma=0.0, temporal_spatial_incoh=2, obj_lens_ti_sigma=0.0)
    p = _case(12, illumination_model=4, **kw)
    full = oracle.run(p, "f64").m2psi_tot
    x, w = oracle.lens_quadrature(p, 0)
    assert len(x) == 4
    acc = np.zeros_like(full)
    for xt, wt in zip(x, w):
        q = _case(12, illumination_model=1, **dict(kw, cond_lens_c_10=20.0 + xt))
        acc += wt * oracle.run(q, "f64").m2psi_tot
    assert rel_l2(full, acc) < 1e-9


def _ewfs_tilted(theta, phi):
    p = _case(51, theta=theta, phi=phi)
    return np.abs(oracle.run(p, "f64").psi_coh[0]) ** 2      # one configuration: psi_coh only


def test_ped_single_azimuth_is_the_unshifted_tilted_pattern():
    """PED leaves out the output ramp exp(-2 pi i g0.r) (dp_Shift,
    src/wave_function.cuh:77-90): with one azimuth it is the tilted EWFS
    pattern moved back by g0, a whole number of pixels here (theta is
    quantised to multiples of lambda dg, src/input_multislice.cuh:891-895),
    so the direct beam stays at the centre for every azimuth."""
    theta = 1.0
    ped = oracle.run(_case(41, ped_theta=theta, ped_nrot=1), "f64").m2psi_tot[0]
    ew = _ewfs_tilted(theta, 0.0)
    p = _case(51)
    lam = 0.0196875
    n = int(round(np.sin(np.deg2rad(theta)) / (lam / p.spec_lx)))
    assert n > 0
    assert rel_l2(ped, np.roll(ew, n, axis=0)) < 1e-10
    c = p.nx // 2
    assert ped[c, c] == ped.max()
    assert ew[c - n, c] == ew.max()


def test_ped_averages_the_azimuths():
    """Two azimuths (phi = 0, pi): the mean of the two tilted patterns,
    each moved back by its g0 (the pi azimuth has a residual sin(pi) of
    1e-7 in float, hence the tolerance)."""
    theta = 1.0
    ped = oracle.run(_case(41, ped_theta=theta, ped_nrot=2), "f64").m2psi_tot[0]
    p = _case(51)
    lam = 0.0196875
    n = int(round(np.sin(np.deg2rad(theta)) / (lam / p.spec_lx)))
    want = 0.5 * (np.roll(_ewfs_tilted(theta, 0.0), n, axis=0) + np.roll(_ewfs_tilted(theta, 180.0), -n, axis=0))
    assert rel_l2(ped, want) < 1e-4


@pytest.mark.parametrize("nrot", [1, 3])
def test_hctem_is_the_mean_of_tilted_hrtem_images(nrot):
    """HCTEM = HRTEM with the tilt azimuth stepped over the cone and the
    images averaged (PED_HCTEM, weight 1/(nconf nrot))."""
    theta = 0.8
    hc = oracle.run(_case(42, hci_theta=theta, hci_nrot=nrot), "f64").m2psi_tot
    acc = np.zeros_like(hc)
    for r in range(nrot):
        acc += oracle.run(_case(32, theta=theta, phi=360.0 * r / nrot), "f64").m2psi_tot
    acc /= nrot
    assert rel_l2(hc, acc) < (1e-12 if nrot == 1 else 1e-6)


def test_output_axes_follow_the_grid_space():
    """Real-space axes for the images (ISTEM, CBEI, HCTEM), reciprocal for
    the patterns (ED, PED) (Output_Multislice::set_output_grid)."""
    for st, rs in [(12, True), (22, True), (42, True), (31, False), (41, False)]:
        kw = dict(iw_type=2) if st in (12, 22) else {}
        p = _case(st, **kw)
        d = oracle.run(p, "f64")
        assert np.isclose(d.dx, p.spec_lx / p.nx if rs else 1.0 / p.spec_lx), st
