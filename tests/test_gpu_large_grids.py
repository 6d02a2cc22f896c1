"""The FP32 slice passes at 1024^2 and 2048^2, and the benchmarked STEM plan
itself, against the FP64 oracle through the C ABI.

Every grid of the reference goes through the same FFT<float> plan
(src/fft.cuh:422-474); here each grid size has its own radix schedules and
launch variants, so each size is swept on its own.  Tolerance: 1e-4 relative
L2 of the image or diffraction pattern against the FP64 oracle; variants
with the default radix schedule are bit-identical to variant 0.

Workloads:
  SrTiO3 [001] 1024^2, 24x24 cells, 50 cells thick (101 slices): the 1024^2
      member of the uneven-size group, alone and inside a group.  Its atomic
      planes at z = 78.1 A and 156.2 A (40 and 80 times a/2) sit on edges of
      the plane finder's 0.1 A histogram, and the reference's FP32 path,
      which bins z in float (Identify_Planes, src/types.cuh:1166-1372, with
      Input_Multislice<float>), places the slice boundaries around them
      differently from its FP64 path: slices 40-42 and 80-82 change
      thickness, and the FP32 image differs from the FP64 one by 6e-3.  The
      engine follows the FP32 path, so this member is held to the FP32
      oracle (same slicing); the slicing difference is asserted by the CPU
      test test_boundary.py::test_fp32_slicing_quirk_srtio3;
  Au [001] 2048^2, 16x16 cells, 2 cells thick (5 slices);
  Au [001] 512^2, 8x8x20 cells (41 slices), 11x11 scan, 2 configurations,
      batch 118 and launch variants 8/3/3 (row, column, detector) -- the
      specimen, grid, batch and variants of the 64x64-scan,
      16-configuration benchmark;
  Pt [110] CBED 1024^2, 200 kV, 10 mrad, -100 A defocus, 20 configurations,
      full depth (thicknesses every 50 A to 450 A, 325 slices): FP32 wave
      against the FP64 wave path, per thickness; the FP64 path is itself
      checked against the FP64 oracle at 1024^2 on a 100 A Pt [110] crystal.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SAME_SCHEDULE = (0, 1, 2, 5, 6, 8, 10)


def _forced(p, row, col, col_det, batch=0):
    engine.set_pass_variants(row, col, col_det)
    engine.set_batch(batch, True)
    try:
        out = engine.run_multislice(p)
        return out, engine.last_plan()
    finally:
        engine.set_pass_variants(-1, -1, -1)
        engine.set_batch(0, True)


def _sweep(p, ref, label):
    """Every launch variant (detector pass: the same variant, or the default
    width where the variant does not apply to it) against the oracle."""
    base, _ = _forced(p, 0, 0, 0)
    base = base.image_tot.copy()
    e0 = rel_l2(base, ref)
    print("%s variant 0: rel L2 vs oracle %.3e" % (label, e0))
    assert e0 < 1e-4
    ran = [0]
    for v in range(1, engine.PASS_VARIANTS):
        try:
            out, plan = _forced(p, v, v, v)
        except engine.EngineError as e:
            assert e.status == 1, e
            out, plan = _forced(p, v, v, 0)
        img = out.image_tot.copy()
        e64 = rel_l2(img, ref)
        print("%s %s: rel L2 vs oracle %.3e" % (label, plan.split(" batch=")[1], e64))
        assert e64 < 1e-4, v
        if v in SAME_SCHEDULE:
            assert np.array_equal(img, base), v
        ran.append(v)
    assert ran == list(range(engine.PASS_VARIANTS))


def _sto1024(ns=2):
    return configs.sto_stem(n=1024, ncells=24, nc=50, ns=ns, nconf=1)


@pytest.fixture(scope="module")
def sto1024_ref():
    """The FP32 oracle's image (the reference's FP32 slicing) of the SrTiO3
    1024^2 member."""
    p = _sto1024()
    return p, oracle.run(p, "f32").image_tot


def test_sto_1024_tuned_and_variant_sweep(sto1024_ref):
    p, ref = sto1024_ref
    out = engine.run_multislice(p)               # the tuner's choice
    e = rel_l2(out.image_tot, ref)
    print("SrTiO3 1024^2 (101 slices), tuned %s: rel L2 vs FP32 oracle %.3e" % (engine.last_plan(), e))
    assert e < 1e-4
    _sweep(p, ref, "SrTiO3 1024^2")


def test_sto_1024_inside_group(sto1024_ref):
    """The 1024^2 member run inside an uneven-size group with a 256^2 and a
    512^2 SrTiO3 specimen: bit-identical to its solo run with the same
    variants, and within 1e-4 of the FP64 oracle."""
    p, ref = sto1024_ref
    ps = [configs.sto_stem(n=256, ncells=6, nc=10, ns=3, nconf=2),
          configs.sto_stem(n=512, ncells=12, nc=4, ns=2, nconf=1), p]
    for v in (0, 3):
        engine.set_pass_variants(v, v, 0)
        try:
            solo = engine.run_multislice(p).image_tot.copy()
            grp = engine.run_multislice_group(ps)
            plan = engine.last_plan()
        finally:
            engine.set_pass_variants(-1, -1, -1)
        assert plan.count("stem fp32 nx=") == 3 and "nx=1024" in plan
        assert np.array_equal(grp[2].image_tot, solo)
        e = rel_l2(grp[2].image_tot, ref)
        print("group (row/col variant %d) 1024^2 member: rel L2 vs FP32 oracle %.3e" % (v, e))
        assert e < 1e-4
    e256 = rel_l2(grp[0].image_tot, oracle.run(ps[0], "f64").image_tot)
    print("group 256^2 member: rel L2 vs FP64 %.3e" % e256)
    assert e256 < 1e-4


def test_stem_2048_variant_sweep():
    p = configs.au_stem_512(ns=2, nconf=1, n=2048, ncells=16, nc=2)
    ref = oracle.run(p, "f64").image_tot
    out = engine.run_multislice(p)
    e = rel_l2(out.image_tot, ref)
    print("Au 2048^2 tuned %s: rel L2 vs FP64 %.3e" % (engine.last_plan(), e))
    assert e < 1e-4
    _sweep(p, ref, "Au 2048^2")


def test_benchmarked_stem_plan():
    """The benchmark's specimen and plan (Au [001] 8x8x20 cells, 41 slices,
    512^2, batch 118, variants 8/3/3: persistent row pass, alternative
    schedule column and detector passes) on an 11x11 scan over one cell
    with 2 phonon configurations: 121 probes = one full batch and a tail of
    three."""
    p = configs.au_stem_512(ns=11, nconf=2)
    out, plan = _forced(p, 8, 3, 3, batch=118)
    assert "batch=118 " in plan and "row_variant=8 col_variant=3 col_det_variant=3" in plan, plan
    ref = oracle.run(p, "f64").image_tot
    e = rel_l2(out.image_tot, ref)
    print("benchmarked plan (%s) on 11x11 probes x 2 configs: rel L2 vs FP64 %.3e" % (plan, e))
    assert e < 1e-4


def _pt1024(**kw):
    return configs.pt_cbed_1024(n=1024, **kw)


def test_cbed_pt110_1024_fp64_path_vs_oracle():
    """The FP64 wave path at 1024^2 on a 100 A Pt [110] crystal (36 cells,
    73 slices, 2 configurations) against the FP64 oracle: that path is the
    reference of the full-depth precision study below.  The FP32 path is
    held to 1e-4 on the same run."""
    p = _pt1024(nconf=2, nc=36, thick=[50.0, 100.0])
    p.system_conf.precision = 2
    out, ref = engine.run_multislice(p), oracle.run(p, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        print("Pt[110] 1024^2 FP64 path, thickness %.0f A: rel L2 vs FP64 oracle %.3e" % (p.thick[k], e))
        assert e < 1e-9
    q = p.copy()
    q.system_conf.precision = 1
    out32 = engine.run_multislice(q)
    for k in range(out32.n_thk):
        e = rel_l2(out32.m2psi_tot[k], ref.m2psi_tot[k])
        print("Pt[110] 1024^2 FP32 path, thickness %.0f A: rel L2 vs FP64 oracle %.3e" % (p.thick[k], e))
        assert e < 1e-4


def test_cbed_pt110_1024_full_depth_precision_study():
    """FP32 wave + FP64 phases and accumulation against the full FP64 wave
    path: Pt [110] 1024^2, 20 configurations, thicknesses 50..450 A."""
    p = _pt1024(nconf=20)
    p64 = p.copy()
    p64.system_conf.precision = 2
    out32, out64 = engine.run_multislice(p), engine.run_multislice(p64)
    assert len(p.thick) == 9
    errs = []
    for k in range(out32.n_thk):
        e = rel_l2(out32.m2psi_tot[k], out64.m2psi_tot[k])
        errs.append(e)
        print("Pt[110] CBED 1024^2, 20 configs, thickness %5.1f A: FP32 vs FP64 rel L2 %.3e" % (p.thick[k], e))
    assert max(errs) < 1e-4
