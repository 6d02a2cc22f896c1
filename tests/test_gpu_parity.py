"""Parity of the HIP engine against the CPU oracle, through the C ABI.

Tolerances (relative L2):
  projected potential: bit-identical to the FP32 restatement (same float
      operations in the same order), 1e-5 vs FP64;
  transmission: 5e-7 vs the FP32 restatement (sincosf differs by ulps);
  single pieces (probe, propagate): 1e-5 vs FP64;
  exit wave (Si [001], 256^2, 41 slices): 1e-5 vs the FP32 restatement of
      the reference path, 1e-4 vs FP64;
  STEM / CBED images: 1e-4 vs FP64.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def small_stem(n=256, ns=4, nconf=2, nc=5, ncells=4):
    return configs.au_stem_512(ns=ns, nconf=nconf, n=n, ncells=ncells, nc=nc)


def test_projected_potential_all_slices_si():
    p = configs.si_ewrs_256()
    ns = len(oracle.spec_slicing(p))
    for islice in range(ns):
        V = engine.projected_potential(p, islice)
        V32 = oracle.projected_potential(p, islice, precision="f32")
        V64 = oracle.projected_potential(p, islice, precision="f64")
        # same float operations in the same order as the FP32 restatement
        assert np.array_equal(V, V32), islice
        assert rel_l2(V, V64) < 1e-5


def test_projected_potential_phonon_au():
    p = small_stem()
    for iconf in (1, 2):
        for islice in (0, 3, 10):
            V = engine.projected_potential(p, islice, iconf)
            V32 = oracle.projected_potential(p, islice, iconf, precision="f32")
            assert np.array_equal(V, V32), (iconf, islice)


def test_transmission_function():
    p = configs.si_ewrs_256()
    for islice in (0, 1, 20, 40):
        T = engine.transmission_function(p, islice)
        T32 = oracle.transmission_function(p, islice, precision="f32")
        T64 = oracle.transmission_function(p, islice, precision="f64")
        assert rel_l2(T, T32) < 5e-7     # V identical; sincosf differs by ulps
        assert rel_l2(T, T64) < 1e-5


def test_transmission_function_bwl():
    p = configs.si_ewrs_256()
    p.bwl = 1
    T = engine.transmission_function(p, 3)
    T64 = oracle.transmission_function(p, 3, precision="f64")
    assert rel_l2(T, T64) < 1e-5


def test_incident_wave_probe():
    p = small_stem()
    p.simulation_type = 21
    p.iw_type = 2
    p.cond_lens_c_10 = 35.0
    p.cond_lens_c_30 = 0.002
    p.iw_x = np.array([0.3 * p.spec_lx])
    p.iw_y = np.array([0.6 * p.spec_ly])
    psi = engine.incident_wave(p)
    ref = oracle.incident_wave(p, precision="f64")
    assert rel_l2(psi, ref) < 1e-5
    assert abs(np.sum(np.abs(psi) ** 2) / psi.size - 1.0) < 1e-5


@pytest.mark.parametrize("space", [1, 2])
def test_propagate(space):
    p = configs.si_ewrs_256()
    rng = np.random.default_rng(1)
    psi = rng.standard_normal((p.nx, p.ny)) + 1j * rng.standard_normal((p.nx, p.ny))
    out = engine.propagate(p, psi, 7.5, space)
    ref = oracle.propagate(p, psi, 7.5, space, precision="f64")
    assert rel_l2(out, ref) < 1e-5


def test_exit_wave_si():
    p = configs.si_ewrs_256()
    out = engine.run_multislice(p)
    r32 = oracle.run(p, "f32")
    r64 = oracle.run(p, "f64")
    e32, e64 = rel_l2(out.psi_coh, r32.psi_coh), rel_l2(out.psi_coh, r64.psi_coh)
    print("exit wave rel L2 vs FP32 restatement %.3e, vs FP64 %.3e" % (e32, e64))
    assert e32 < 1e-5
    assert e64 < 1e-4


def test_stem_small():
    p = small_stem()
    out = engine.run_multislice(p)
    r64 = oracle.run(p, "f64")
    e = rel_l2(out.image_tot, r64.image_tot)
    print("STEM image rel L2 vs FP64 %.3e" % e)
    assert e < 1e-4


def test_stem_batch_and_graph_invariance():
    # fixed launch variants: the batch width and the graph replay alone must not
    # change a single bit
    p = small_stem(ns=5, nconf=1)
    engine.set_pass_variants(0, 0, 0)
    try:
        engine.set_batch(0, True)
        a = engine.run_multislice(p).image_tot.copy()
        engine.set_batch(3, True)
        b = engine.run_multislice(p).image_tot.copy()
        engine.set_batch(7, False)
        c = engine.run_multislice(p).image_tot.copy()
    finally:
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)
    assert np.array_equal(a, b)
    assert np.array_equal(a, c)


# launch variants with the default radix schedule (Sched<N, 0>): same
# arithmetic in the same order as variant 0
SAME_SCHEDULE = (0, 1, 2, 5, 6, 8, 10)


def _run_forced(p, row, col, col_det):
    engine.set_pass_variants(row, col, col_det)
    try:
        return engine.run_multislice(p)
    finally:
        engine.set_pass_variants(-1, -1, -1)


@pytest.mark.parametrize("n", [256, 512])
def test_stem_pass_variant_sweep(n):
    """Every launch variant of the row, column and detector column passes
    against the FP64 oracle; variants that keep the radix schedule are
    bit-identical to variant 0."""
    p = small_stem(n=n, ns=3, nconf=1)
    r64 = oracle.run(p, "f64")
    base = _run_forced(p, 0, 0, 0).image_tot.copy()
    assert rel_l2(base, r64.image_tot) < 1e-4
    for v in range(1, engine.PASS_VARIANTS):
        try:
            img = _run_forced(p, v, v, v).image_tot.copy()
            vdet = v
        except engine.EngineError as e:
            # detector epilogues keep the default column width
            assert e.status == 1, e
            img = _run_forced(p, v, v, 0).image_tot.copy()
            vdet = 0
        e64 = rel_l2(img, r64.image_tot)
        print("n=%d variant %d (det %d): rel L2 vs FP64 %.3e" % (n, v, vdet, e64))
        assert e64 < 1e-4, v
        if v in SAME_SCHEDULE:
            assert np.array_equal(img, base), v


def test_exit_wave_pass_variants():
    p = configs.si_ewrs_256()
    r32 = oracle.run(p, "f32")
    base = _run_forced(p, 0, 0, -1).psi_coh.copy()
    for v in range(1, engine.PASS_VARIANTS):
        psi = _run_forced(p, v, v, -1).psi_coh.copy()
        e32 = rel_l2(psi, r32.psi_coh)
        print("exit wave variant %d: rel L2 vs FP32 restatement %.3e" % (v, e32))
        assert e32 < 1e-5, v
        if v in SAME_SCHEDULE:
            assert np.array_equal(psi, base), v


def test_autotune_log_and_cache():
    p = small_stem(ns=4, nconf=1)
    engine.set_autotune(2)
    try:
        engine.run_multislice(p)
        log = engine.tune_log()
        plan = engine.last_plan()
    finally:
        engine.set_autotune(1)
    print(log)
    for name in ("row", "col", "col_det"):
        assert (" %s=v" % name) in " " + log
    chosen = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in plan.split() if kv.endswith(tuple("0123456789"))
              and "_variant=" in kv}
    assert set(chosen) == {"row_variant", "col_variant", "col_det_variant"}
    assert all(0 <= v < engine.PASS_VARIANTS for v in chosen.values())
    engine.run_multislice(p)           # mode 1: the same shape is served from the cache
    engine.run_multislice(p)
    assert engine.tune_log().startswith("cached: ")


def test_forced_variant_out_of_range():
    p = small_stem(ns=2, nconf=1)
    with pytest.raises(engine.EngineError) as ei:
        _run_forced(p, engine.PASS_VARIANTS, -1, -1)
    assert ei.value.status == 1


def test_cbed_small():
    p = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=12, thick=[15.0, 30.0])
    out = engine.run_multislice(p)
    r64 = oracle.run(p, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], r64.m2psi_tot[k])
        print("CBED thickness", k, "rel L2 vs FP64 %.3e" % e)
        assert e < 1e-4


def _group_inputs():
    return [configs.sto_stem(n=256, ncells=6, nc=3, ns=6, nconf=2),
            configs.sto_stem(n=512, ncells=12, nc=2, ns=5, nconf=1),
            small_stem(n=256, ns=4, nconf=3)]


@pytest.mark.parametrize("batch", [0, 3])
def test_group_uneven_sizes_matches_solo_runs(batch):
    """Specimens of different grid sizes in one group run give, bit for bit,
    what each gives alone (fixed launch variants; batch 3 makes the
    specimens run out of work in different rounds)."""
    ps = _group_inputs()
    engine.set_pass_variants(0, 0, 0)
    engine.set_batch(batch, True)
    try:
        solo = [engine.run_multislice(p).image_tot.copy() for p in ps]
        grp = engine.run_multislice_group(ps)
        plan = engine.last_plan()
    finally:
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)
    assert plan.startswith("group ") and plan.count("stem fp32 nx=") == 3
    for a, o in zip(solo, grp):
        assert np.array_equal(a, o.image_tot)


def test_group_vs_oracle():
    ps = _group_inputs()[:2]
    outs = engine.run_multislice_group(ps)
    for p, o in zip(ps, outs):
        e = rel_l2(o.image_tot, oracle.run(p, "f64").image_tot)
        print("group member %d^2: rel L2 vs FP64 %.3e" % (p.nx, e))
        assert e < 1e-4


def test_group_refuses_non_stem():
    with pytest.raises(engine.EngineError) as ei:
        engine.run_multislice_group([small_stem(), configs.si_ewrs_256()])
    assert ei.value.status == 2


def test_tilted_exit_wave():
    p = configs.si_ewrs_256()
    p.theta, p.phi = 0.6, 30.0                    # degrees
    out = engine.run_multislice(p)
    e32 = rel_l2(out.psi_coh, oracle.run(p, "f32").psi_coh)
    e64 = rel_l2(out.psi_coh, oracle.run(p, "f64").psi_coh)
    print("tilted exit wave: rel L2 vs FP32 restatement %.3e, vs FP64 %.3e" % (e32, e64))
    assert e32 < 1e-5 and e64 < 1e-4


def test_tilted_stem_and_cbed():
    p = small_stem(ns=3, nconf=2)
    p.theta, p.phi = 0.5, 0.0
    e = rel_l2(engine.run_multislice(p).image_tot, oracle.run(p, "f64").image_tot)
    print("tilted STEM rel L2 vs FP64 %.3e" % e)
    assert e < 1e-4
    q = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=12, thick=[15.0, 30.0])
    q.theta, q.phi = 0.3, 45.0
    out, ref = engine.run_multislice(q), oracle.run(q, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        print("tilted CBED thickness %d rel L2 vs FP64 %.3e" % (k, e))
        assert e < 1e-4


def test_tilted_pieces():
    p = small_stem()
    p.simulation_type = 21
    p.iw_type = 2
    p.cond_lens_c_10 = 20.0
    p.iw_x = np.array([0.4 * p.spec_lx])
    p.iw_y = np.array([0.5 * p.spec_ly])
    p.theta, p.phi = 0.4, 60.0
    assert rel_l2(engine.incident_wave(p), oracle.incident_wave(p, precision="f64")) < 1e-5
    rng = np.random.default_rng(3)
    psi = rng.standard_normal((p.nx, p.ny)) + 1j * rng.standard_normal((p.nx, p.ny))
    for space in (1, 2):
        a = engine.propagate(p, psi, 37.0, space)
        b = oracle.propagate(p, psi, 37.0, space, precision="f64")
        assert rel_l2(a, b) < 1e-5, space


@pytest.mark.parametrize("theta", [0.0, 0.5])
def test_stem_coherent_image(theta):
    p = small_stem(ns=3, nconf=3)
    p.pn_coh_contrib = 1
    p.theta = theta
    out, ref = engine.run_multislice(p), oracle.run(p, "f64")
    e_tot = rel_l2(out.image_tot, ref.image_tot)
    e_coh = rel_l2(out.image_coh, ref.image_coh)
    print("coherent STEM theta=%.1f: image_tot %.3e, image_coh %.3e" % (theta, e_tot, e_coh))
    assert e_tot < 1e-4 and e_coh < 1e-4


def test_stem_coherent_image_in_scan_blocks():
    """A memory budget too small for every scan point's coherent sums runs
    the scan in blocks; with the launch variants and batch fixed the images
    are bit-identical to the unblocked run."""
    p = small_stem(ns=4, nconf=2)
    p.pn_coh_contrib = 1
    d = engine.output_dims(p)
    nscan = d.nx * d.ny
    nxy = p.nx * p.ny
    ns = len(oracle.spec_slicing(p))
    per_scan = d.n_thk * nxy * 16
    budget = ns * nxy * 8 + nscan * per_scan // 2 + 4 * nxy * 8 * 2
    engine.set_pass_variants(0, 0, 0)
    engine.set_batch(4, True)
    try:
        a = engine.run_multislice(p)
        engine.set_memory_budget(budget)
        b = engine.run_multislice(p)
        plan = engine.last_plan()
    finally:
        engine.set_memory_budget(0)
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)
    assert "scan_blocks=" in plan and "scan_blocks=1 " not in plan, plan
    assert np.array_equal(a.image_tot, b.image_tot)
    assert np.array_equal(a.image_coh, b.image_coh)


def test_projected_potential_sum():
    p = configs.si_ewrs_256(nc=3)
    ns = len(oracle.spec_slicing(p))
    want = np.zeros((p.nx, p.ny))
    for i in range(ns):
        want += oracle.projected_potential(p, i, precision="f32")
    assert np.array_equal(engine.projected_potential_sum(p), want)


def test_azimuthal_aberrations():
    # two-fold and three-fold astigmatism, axial coma, a vortex term
    p = small_stem(ns=3, nconf=1)
    p.cond_lens_c_12, p.cond_lens_phi_12 = 15.0, 30.0
    p.cond_lens_c_23, p.cond_lens_phi_23 = 300.0, 70.0
    p.cond_lens_c_21, p.cond_lens_phi_21 = 200.0, 10.0
    p.cond_lens_m = 1
    e = rel_l2(engine.run_multislice(p).image_tot, oracle.run(p, "f64").image_tot)
    print("STEM with azimuthal aberrations: rel L2 vs FP64 %.3e" % e)
    assert e < 1e-4
    q = p.copy()
    q.simulation_type = 21
    q.iw_type = 2
    q.iw_x = np.array([0.45 * q.spec_lx])
    q.iw_y = np.array([0.55 * q.spec_ly])
    assert rel_l2(engine.incident_wave(q), oracle.incident_wave(q, precision="f64")) < 1e-5


@pytest.mark.parametrize("illum", [1, 2])
def test_hrtem(illum):
    # objective lens with defocus, Cs and (partially coherent) focal/angular spread
    p = configs.si_ewrs_256(nc=4)
    p.simulation_type = 32
    p.illumination_model = illum
    p.obj_lens_c_10, p.obj_lens_c_30 = -55.0, 0.8
    p.obj_lens_outer_aper_ang = 25.0
    p.obj_lens_ti_sigma, p.cond_lens_si_sigma = 30.0, 0.005
    out = engine.run_multislice(p)
    e64 = rel_l2(out.m2psi_tot, oracle.run(p, "f64").m2psi_tot)
    e32 = rel_l2(out.m2psi_tot, oracle.run(p, "f32").m2psi_tot)
    print("HRTEM illumination %d: rel L2 vs FP64 %.3e, vs FP32 restatement %.3e" % (illum, e64, e32))
    assert e64 < 1e-4 and e32 < 1e-4


@pytest.mark.parametrize("illum,theta", [(1, 0.0), (2, 0.0), (4, 0.0), (1, 0.5)])
def test_hrtem_coherent_image(illum, theta):
    """HRTEM with pn_coh_contrib: m2psi_coh = |IFFT(lens psi_coh)|^2 of the
    configuration-averaged wave (set_m2psi_coh, src/wave_function.cuh:206-215)."""
    p = configs.si_ewrs_256(nc=4)
    p.simulation_type = 32
    p.pn_model, p.pn_nconf, p.pn_coh_contrib = 3, 2, 1
    p.illumination_model = illum
    p.temporal_spatial_incoh = 2
    p.theta, p.phi = theta, 20.0
    p.obj_lens_c_10, p.obj_lens_c_30, p.obj_lens_outer_aper_ang = -55.0, 0.8, 25.0
    p.obj_lens_ti_sigma, p.obj_lens_ti_npts = 30.0, 5
    out, ref = engine.run_multislice(p), oracle.run(p, "f64")
    e_tot, e_coh = rel_l2(out.m2psi_tot, ref.m2psi_tot), rel_l2(out.m2psi_coh, ref.m2psi_coh)
    print("HRTEM illum %d theta %.1f: m2psi_tot %.2e, m2psi_coh %.2e vs FP64" % (illum, theta, e_tot, e_coh))
    assert e_tot < 1e-4 and e_coh < 1e-4


def _user_probe(p):
    """A convergent probe at the box centre, as a user-defined incident wave (ny x nx)."""
    q = p.copy()
    q.simulation_type, q.iw_type = 21, 2
    q.iw_x, q.iw_y = np.array([0.5 * q.spec_lx]), np.array([0.5 * q.spec_ly])
    q.cond_lens_c_10 = 25.0
    return oracle.incident_wave(q, precision="f64").T


def test_user_defined_incident_wave():
    """iw_type 3: the user's wave moved to iw_x[0], iw_y[0] (EW, CBED) or to
    each scan position (STEM), against the FP64 oracle."""
    p = small_stem(ns=3, nconf=1)
    p.iw_type = 3
    p.iw_psi = _user_probe(p)
    e = rel_l2(engine.run_multislice(p).image_tot, oracle.run(p, "f64").image_tot)
    print("STEM with a user incident wave: rel L2 vs FP64 %.2e" % e)
    assert e < 1e-4
    q = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=8, thick=[10.0, 20.0])
    q.iw_type = 3
    q.iw_psi = _user_probe(q)
    out, ref = engine.run_multislice(q), oracle.run(q, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        print("CBED with a user incident wave, thickness %d: rel L2 vs FP64 %.2e" % (k, e))
        assert e < 1e-4
    assert rel_l2(engine.incident_wave(q), oracle.incident_wave(q, precision="f64")) < 1e-5


def test_rotated_specimen():
    """spec_rot_theta about a tilted axis: the engine slices the same rotated
    specimen as the oracle (pinned against scipy in tests/test_rotation.py)."""
    p = configs.si_ewrs_256(nc=4)
    p.spec_rot_theta, p.spec_rot_u0 = 12.0, np.array([1.0, 0.3, 0.0])
    p.potential_slicing = 2          # a rotated crystal has no atomic planes along z
    assert len(oracle.spec_slicing(p)) > 10
    out = engine.run_multislice(p)
    e32 = rel_l2(out.psi_coh, oracle.run(p, "f32").psi_coh)
    e64 = rel_l2(out.psi_coh, oracle.run(p, "f64").psi_coh)
    print("rotated Si exit wave: rel L2 vs FP32 restatement %.2e, vs FP64 %.2e" % (e32, e64))
    assert e32 < 1e-5 and e64 < 1e-4


@pytest.mark.parametrize("precision,theta", [(1, 0.0), (2, 0.0), (1, 0.5), (2, 0.5)])
def test_matrix_detectors(precision, theta):
    """detector_type 3: sensitivity maps weight |psi(g)|^2 in the detector
    pass and in the coherent image, against the oracle; with tilt the
    output ramp comes first."""
    p = small_stem(ns=3, nconf=2)
    p.system_conf.precision = precision
    p.pn_coh_contrib = 1
    p.theta = theta
    p.detector_type = 3
    rng = np.random.default_rng(11)
    p.detector_fR = [rng.random((p.ny, p.nx)), (rng.random((p.ny, p.nx)) > 0.7).astype(float)]
    out, ref = engine.run_multislice(p), oracle.run(p, "f64")
    e_tot, e_coh = rel_l2(out.image_tot, ref.image_tot), rel_l2(out.image_coh, ref.image_coh)
    print("matrix detectors (precision %d, theta %.1f): image_tot %.2e, image_coh %.2e vs FP64"
          % (precision, theta, e_tot, e_coh))
    tol = 1e-4 if precision == 1 else 1e-9
    assert e_tot < tol and e_coh < tol


def test_matrix_detector_ring_map_equals_ring():
    """The reference example's centred ring map gives the ring detector's
    image on the GPU (example_MULTEM_STEM_matrix_detector.m:115-130)."""
    from tests.test_matrix_detector import ring_map
    p = small_stem(ns=3, nconf=1)
    p.detector_cir = [(20.0, 60.0), (80.0, 200.0)]
    ring = oracle.run(p, "f64").image_tot          # the rings, FP64 g^2 as the map's
    q = p.copy()
    q.detector_type = 3
    q.detector_fR = [ring_map(q, a, b) for a, b in p.detector_cir]
    e = rel_l2(engine.run_multislice(q).image_tot, ring)
    print("centred ring maps (GPU) vs rings (FP64 oracle): rel L2 %.2e" % e)
    assert e < 1e-4


def test_tuner_cache_separates_generic_and_specialised_passes():
    """A bandwidth-limited STEM run (generic column kernels, one launch
    variant) and then an unlimited one at the same grid and batch in the
    same process: the second must tune its own specialised kernels, not
    reuse the first's choice, and both match the FP64 oracle."""
    p = small_stem(ns=3, nconf=1)
    engine.set_batch(9, True)
    try:
        for bwl in (1, 0, 1):
            q = p.copy()
            q.bwl = bwl
            out = engine.run_multislice(q)
            e = rel_l2(out.image_tot, oracle.run(q, "f64").image_tot)
            print("bwl=%d: %s | %s; rel L2 vs FP64 %.2e" % (bwl, engine.last_plan(), engine.tune_log(), e))
            assert e < 1e-4
            if bwl:
                assert "col_det=v0(generic)" in engine.tune_log() or "cached" in engine.tune_log()
    finally:
        engine.set_batch(0, True)


@pytest.mark.parametrize("n", [256, 512])
def test_wave_pitch_does_not_change_the_image(n):
    """STEM wave batches stored at row pitch ny (the default), ny + 16 and
    ny + 40: bit-identical images (launch variants are fixed so that all
    three runs use the same kernels)."""
    p = small_stem(n=n, ns=3, nconf=2)
    imgs = []
    engine.set_pass_variants(0, 3, 0)
    try:
        for pad in (0, 16, 40):
            engine.set_wave_pad(pad)
            imgs.append(engine.run_multislice(p).image_tot.copy())
    finally:
        engine.set_wave_pad(-1)
        engine.set_pass_variants(-1, -1, -1)
    assert np.array_equal(imgs[0], imgs[1]) and np.array_equal(imgs[0], imgs[2])
    assert rel_l2(imgs[1], oracle.run(p, "f64").image_tot) < 1e-4


@pytest.mark.parametrize("kw", [dict(), dict(theta=0.4, phi=20.0, pn_coh_contrib=1)])
def test_batch_chains_do_not_change_the_image(kw):
    """A batch split into 1, 2 or 3 concurrent chains (each on its own
    stream in the batch's graph): bit-identical images, incoherent and, with
    tilted illumination, coherent (the output-ramp scratch and the coherent
    sums are indexed per chain)."""
    p = small_stem(n=256, ns=6, nconf=2)
    for k, v in kw.items():
        setattr(p, k, v)
    outs, plans = [], []
    engine.set_pass_variants(0, 3, 0)
    engine.set_batch(18, True)
    try:
        for chains in (1, 2, 3):
            engine.set_chains(chains)
            o = engine.run_multislice(p)
            outs.append([a.copy() for a in (o.image_tot, o.image_coh) if a is not None])
            plans.append(engine.last_plan())
    finally:
        engine.set_chains(0)
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)
    assert "chains=2" in plans[1] and "chains=3" in plans[2], plans
    for o in outs[1:]:
        for a, b in zip(o, outs[0]):
            assert np.array_equal(a, b)
    ref = oracle.run(p, "f64")
    assert rel_l2(outs[0][0], ref.image_tot) < 1e-4


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_projected_potential_many_atom_types(precision):
    """Ten elements in one specimen: the spline tables of every type share
    the potential kernel's LDS (51 KB of FP64 tables); V is bit-identical to
    the FP32 restatement and within 1e-12 of the FP64 oracle."""
    p = configs.si_ewrs_256(nc=3)
    Zs = np.array([6, 8, 14, 22, 26, 29, 38, 47, 56, 79], dtype=float)
    p.spec_atoms = p.spec_atoms.copy()
    p.spec_atoms[:, 0] = Zs[np.arange(len(p.spec_atoms)) % len(Zs)]
    if precision == "f64":
        p.system_conf.precision = 2
    for islice in (0, 3, 7):
        V = engine.projected_potential(p, islice)
        ref = oracle.projected_potential(p, islice, precision=precision)
        if precision == "f32":
            assert np.array_equal(V, ref), islice
        else:
            assert rel_l2(V, ref) < 1e-12, islice
