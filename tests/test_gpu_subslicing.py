"""Sub-slicing (potential_slicing = 3) on the GPU against the oracle through
the C ABI: the per-slice potentials of configurations with z displacements,
the whole-specimen STEM image and exit wave, FP32 and FP64, the Memory_Slice
fallback, and one identity independent of the oracle: the sub-sliced
potential summed over all slices is the plane-sliced one (both are the
specimen's projected potential).

Tolerances: potentials 1e-5 (FP32; the splines are built in FP64 and
cast, so only the evaluation rounds in float) / 1e-11 (FP64) relative L2
against the oracle; images 1e-4 (FP32) / 1e-9 (FP64) against the FP64
oracle; summed potential 1e-5 against the oracle's sum.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _sub(nconf=2, pn_dim=111, dz=1.0, precision=1, ns=3):
    p = configs.au_stem_512(ns=ns, nconf=nconf, n=256, ncells=4, nc=3)
    p.potential_slicing = 3
    p.spec_dz = dz
    p.pn_dim = pn_dim
    p.system_conf.precision = precision
    return p


@pytest.mark.parametrize("precision", [1, 2])
def test_subsliced_potential_per_slice(precision):
    p = _sub(precision=precision)
    ns = len(engine.spec_slicing(p))
    o = "f32" if precision == 1 else "f64"
    tol = 1e-5 if precision == 1 else 1e-11
    for iconf in (1, 2):
        for islice in (0, 3, ns // 2, ns - 1):
            V = engine.projected_potential(p, islice, iconf)
            ref = oracle.projected_potential(p, islice, iconf, precision=o)
            e = rel_l2(V, ref)
            print("precision %d conf %d slice %d/%d: rel L2 vs oracle %.2e" % (precision, iconf, islice, ns, e))
            assert np.abs(ref).max() > 0
            assert e < tol, (iconf, islice)
    # z displacements: configurations 1 and 2 differ
    assert rel_l2(engine.projected_potential(p, ns // 2, 1), engine.projected_potential(p, ns // 2, 2)) > 1e-3


def test_subsliced_sum_is_the_projected_potential():
    """All slices' sub-sliced potentials add up to the plane-sliced projected
    potential, except for the tails the atom ranges leave out by design (an
    atom farther than R_int_max from a slice's centre is not in it, though
    its potential still reaches the slice's near edge): 4e-5 in the oracle.
    The engine's sum equals the oracle's."""
    p = _sub(nconf=1, pn_dim=110)
    p.pn_model = 1
    q = p.copy()
    q.potential_slicing = 1
    Vs = engine.projected_potential_sum(p)
    Vp = engine.projected_potential_sum(q)
    ns = len(engine.spec_slicing(p))
    Vo = sum(oracle.projected_potential(p, i, 1, precision="f32") for i in range(ns))
    e, eo = rel_l2(Vs, Vp), rel_l2(Vs, Vo)
    print("sub-sliced (%d slices) vs plane-sliced (%d slices) summed potential: rel L2 %.2e; vs the oracle's sum %.2e"
          % (ns, len(engine.spec_slicing(q)), e, eo))
    assert ns > len(engine.spec_slicing(q))
    assert e < 1e-4 and eo < 1e-5


@pytest.mark.parametrize("precision", [1, 2])
def test_stem_subslicing_vs_oracle(precision):
    p = _sub(precision=precision)
    out = engine.run_multislice(p)
    ref = oracle.run(p, "f64").image_tot
    e = rel_l2(out.image_tot, ref)
    print("STEM sub-sliced, precision %d (%s): rel L2 vs FP64 oracle %.2e" % (precision, engine.last_plan(), e))
    assert e < (1e-4 if precision == 1 else 1e-9)


def test_exit_wave_subslicing_vs_oracle():
    p = configs.si_ewrs_256(nc=4)
    p.potential_slicing = 3
    p.spec_dz = 0.7
    out = engine.run_multislice(p)
    e32 = rel_l2(out.psi_coh, oracle.run(p, "f32").psi_coh)
    e64 = rel_l2(out.psi_coh, oracle.run(p, "f64").psi_coh)
    print("EWRS sub-sliced (%d slices): rel L2 vs FP32 oracle %.2e, vs FP64 %.2e" % (len(engine.spec_slicing(p)), e32, e64))
    assert e32 < 1e-5 and e64 < 1e-4


def test_subslicing_rebuilt_slices_bit_identical():
    """The Memory_Slice fallback with sub-slicing: each slice's transmission
    rebuilt from the configuration's pair coefficients equals the resident
    stack's, bit for bit."""
    p = _sub()
    engine.set_pass_variants(0, 0, 0)
    engine.set_batch(2, True)
    try:
        ref = engine.run_multislice(p).image_tot.copy()
        nxy = p.nx * p.ny
        engine.set_memory_budget(nxy * 8 + 3 * nxy * 4 + 2 * nxy * 8 + nxy * 2)
        out = engine.run_multislice(p).image_tot.copy()
        plan = engine.last_plan()
    finally:
        engine.set_memory_budget(0)
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)
    print(plan)
    assert "slices=rebuilt stored_potentials=3/" in plan
    assert np.array_equal(out, ref)
