"""The FP64 wave path (system_conf.precision = 2, mt_hip_run_multislice_f64)
against the CPU oracle built in double, through the C ABI.

Both sides evaluate every step in double: the potential splines, exp(i sigma
V), the propagator, the probe and the FFTs.  They differ only in the FFT
factorisation and in the libm of sincos/exp, so the tolerances (relative L2)
are set at rounding level: 1e-12 for the projected potential (same sums in
the same order), 1e-11 for single pieces, 1e-9 for whole simulations (tens
of slices and FFT round trips).
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL_PIECE = 1e-11
TOL_RUN = 1e-9


def f64(p):
    q = p.copy()
    q.system_conf.precision = 2
    return q


def small_stem(n=256, ns=4, nconf=2, nc=5, ncells=4):
    return f64(configs.au_stem_512(ns=ns, nconf=nconf, n=n, ncells=ncells, nc=nc))


def test_projected_potential_and_transmission():
    p = f64(configs.si_ewrs_256())
    for islice in (0, 7, 40):
        V = engine.projected_potential(p, islice)
        V64 = oracle.projected_potential(p, islice, precision="f64")
        e = rel_l2(V, V64)
        print("V slice %d: rel L2 vs FP64 %.2e, identical %s" % (islice, e, np.array_equal(V, V64)))
        assert e < 1e-12
        T = engine.transmission_function(p, islice)
        T64 = oracle.transmission_function(p, islice, precision="f64")
        e = rel_l2(T, T64)
        print("T slice %d: rel L2 vs FP64 %.2e" % (islice, e))
        assert e < TOL_PIECE


def test_pieces_probe_and_propagate():
    p = small_stem()
    p.simulation_type = 21
    p.iw_type = 2
    p.cond_lens_c_10, p.cond_lens_c_30 = 35.0, 0.002
    p.cond_lens_c_12, p.cond_lens_phi_12 = 15.0, 30.0
    p.iw_x = np.array([0.3 * p.spec_lx])
    p.iw_y = np.array([0.6 * p.spec_ly])
    e = rel_l2(engine.incident_wave(p), oracle.incident_wave(p, precision="f64"))
    print("probe rel L2 vs FP64 %.2e" % e)
    assert e < TOL_PIECE
    rng = np.random.default_rng(5)
    psi = rng.standard_normal((p.nx, p.ny)) + 1j * rng.standard_normal((p.nx, p.ny))
    for space in (1, 2):
        e = rel_l2(engine.propagate(p, psi, 7.5, space), oracle.propagate(p, psi, 7.5, space, precision="f64"))
        print("propagate space %d rel L2 vs FP64 %.2e" % (space, e))
        assert e < TOL_PIECE


def test_stem_fp64():
    p = small_stem()
    out = engine.run_multislice(p)
    assert " fp64 " in " " + engine.last_plan() + " "
    r64 = oracle.run(p, "f64")
    e = rel_l2(out.image_tot, r64.image_tot)
    print("STEM FP64 rel L2 vs FP64 oracle %.3e" % e)
    assert e < TOL_RUN


def test_forced_entries_follow_their_precision():
    p = small_stem(ns=3, nconf=1)
    q = p.copy()
    q.system_conf.precision = 1
    a = engine.run_multislice(q, precision=2).image_tot.copy()     # _f64 entry, float system_conf
    b = engine.run_multislice(p).image_tot.copy()                  # precision 2 through the dispatcher
    c = engine.run_multislice(p, precision=1).image_tot.copy()     # _f32 entry, double system_conf
    d = engine.run_multislice(q).image_tot.copy()
    assert np.array_equal(a, b)
    assert np.array_equal(c, d)
    r64 = oracle.run(p, "f64").image_tot
    e64, e32 = rel_l2(a, r64), rel_l2(c, r64)
    print("forced entries: FP64 %.2e, FP32 %.2e vs FP64 oracle" % (e64, e32))
    assert e64 < TOL_RUN and 1e-9 < e32 < 1e-4


def test_exit_wave_and_cbed_fp64():
    p = f64(configs.si_ewrs_256())
    e = rel_l2(engine.run_multislice(p).psi_coh, oracle.run(p, "f64").psi_coh)
    print("exit wave FP64 rel L2 vs FP64 oracle %.3e" % e)
    assert e < TOL_RUN
    q = f64(configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=12, thick=[15.0, 30.0]))
    out, ref = engine.run_multislice(q), oracle.run(q, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        print("CBED FP64 thickness %d rel L2 %.3e" % (k, e))
        assert e < TOL_RUN


def test_tilted_and_coherent_stem_fp64():
    p = small_stem(ns=3, nconf=2)
    p.theta, p.phi = 0.5, 20.0
    p.pn_coh_contrib = 1
    out, ref = engine.run_multislice(p), oracle.run(p, "f64")
    e_tot, e_coh = rel_l2(out.image_tot, ref.image_tot), rel_l2(out.image_coh, ref.image_coh)
    print("tilted coherent STEM FP64: image_tot %.3e, image_coh %.3e" % (e_tot, e_coh))
    assert e_tot < TOL_RUN and e_coh < TOL_RUN


@pytest.mark.parametrize("illum", [1, 2])
def test_hrtem_fp64(illum):
    p = f64(configs.si_ewrs_256(nc=4))
    p.simulation_type = 32
    p.illumination_model = illum
    p.obj_lens_c_10, p.obj_lens_c_30 = -55.0, 0.8
    p.obj_lens_outer_aper_ang = 25.0
    p.obj_lens_ti_sigma, p.cond_lens_si_sigma = 30.0, 0.005
    e = rel_l2(engine.run_multislice(p).m2psi_tot, oracle.run(p, "f64").m2psi_tot)
    print("HRTEM FP64 illumination %d rel L2 %.3e" % (illum, e))
    assert e < TOL_RUN


@pytest.mark.parametrize("n", [512, 1024, 2048])
def test_stem_fp64_large_grids(n):
    """The narrowed FP64 column tiles at 1024^2 and 2048^2 (4 and 8 columns
    per workgroup) and the detector partials they lay out."""
    p = small_stem(n=n, ns=2, nconf=1, nc=2, ncells=4 * n // 256)
    out = engine.run_multislice(p)
    e = rel_l2(out.image_tot, oracle.run(p, "f64").image_tot)
    print("STEM FP64 %d^2 rel L2 vs FP64 oracle %.3e" % (n, e))
    assert e < TOL_RUN


def test_group_fp64_matches_solo():
    ps = [small_stem(n=256, ns=3, nconf=2), small_stem(n=512, ns=2, nconf=1, ncells=8)]
    solo = [engine.run_multislice(p).image_tot.copy() for p in ps]
    grp = engine.run_multislice_group(ps)
    assert engine.last_plan().count("stem fp64") == 2
    for a, o in zip(solo, grp):
        assert np.array_equal(a, o.image_tot)


def _run_forced(p, row, col, col_det):
    engine.set_pass_variants(row, col, col_det)
    try:
        return engine.run_multislice(p).image_tot.copy()
    finally:
        engine.set_pass_variants(-1, -1, -1)


# FP64 launch variants: the FP32 set without the three that differ only in
# where twiddles live (1, 4, 6: the FP64 passes always read the table);
# those keeping the default radix schedule are bit-identical to variant 0.
FP64_TWIDDLE_TWINS = (1, 4, 6)
FP64_SAME_SCHEDULE = (2, 5, 8, 10)


@pytest.mark.parametrize("n", [256, 512, 1024])
def test_fp64_pass_variant_sweep(n):
    """Each launch variant of the FP64 row and column slice passes, one pass
    at a time: same-schedule variants bit-identical to variant 0, the
    alternative schedule within 1e-12 of it, and (256^2, 512^2) within the
    FP64 run tolerance of the oracle.  Variants whose double LDS tile would
    not fit are refused."""
    p = small_stem(n=n, ns=3, nconf=1)
    base = _run_forced(p, 0, 0, 0)
    if n <= 512:
        e = rel_l2(base, oracle.run(p, "f64").image_tot)
        print("n=%d variant 0: rel L2 vs FP64 oracle %.2e" % (n, e))
        assert e < TOL_RUN
    ran = 0
    for v in range(1, engine.PASS_VARIANTS):
        for which, (rv, cv) in (("row", (v, 0)), ("col", (0, v))):
            try:
                img = _run_forced(p, rv, cv, 0)
            except engine.EngineError as err:
                assert err.status == 1, err
                print("n=%d %s variant %d: refused" % (n, which, v))
                continue
            assert v not in FP64_TWIDDLE_TWINS, (which, v)
            ran += 1
            e = rel_l2(img, base)
            print("n=%d %s variant %d: rel L2 vs variant 0 %.2e" % (n, which, v, e))
            if v in FP64_SAME_SCHEDULE:
                assert np.array_equal(img, base), (which, v)
            else:
                assert e < 1e-12, (which, v)
    assert ran >= 8


def test_fp64_detector_pass_has_one_variant():
    p = small_stem(n=256, ns=2, nconf=1)
    with pytest.raises(engine.EngineError) as ei:
        _run_forced(p, 0, 0, 2)
    assert ei.value.status == 1
