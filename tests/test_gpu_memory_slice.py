"""Memory_Slice fallback (Transmission_Function, src/transmission_function.cuh:
146-215, used at :84-103 and :118-134) behind mt_hip_set_memory_budget.

With a budget below one configuration's transmission stack the engine
stores the potentials of as many slices as fit and rebuilds each slice's
transmission -- from its stored potential, or from the atoms beyond those --
just before the row pass that uses it.  The images must be bit-identical to
the fully resident run (same kernels, same launch variants; the batch width
does not change a bit), for all, some and none of the slices stored, with
and without the bandwidth limit, in STEM, CBED and the exit wave, FP32 and
FP64.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

OOM = 3


def _run(p, budget=0, batch=0):
    engine.set_pass_variants(0, 0, 0 if p.simulation_type == 11 else -1)
    engine.set_batch(batch, True)
    engine.set_memory_budget(budget)
    try:
        out = engine.run_multislice(p)
        return out, engine.last_plan()
    finally:
        engine.set_memory_budget(0)
        engine.set_batch(0, True)
        engine.set_pass_variants(-1, -1, -1)


def _sizes(p):
    nxy = p.nx * p.ny
    real = 8 if p.system_conf.precision == 2 else 4
    ns = len(oracle.spec_slicing(p))
    return nxy, real, ns


def _stem_budget(p, n_pot):
    """STEM: slot + n_pot stored potentials + two waves (+ slack below one more potential)."""
    nxy, real, ns = _sizes(p)
    cplx = 2 * real
    return nxy * cplx + n_pot * nxy * real + 2 * nxy * cplx + nxy * real // 2


def _ew_budget(p, n_pot):
    """Exit wave / CBED: output buffers + slot + n_pot potentials + one wave's buffers."""
    nxy, real, ns = _sizes(p)
    cplx = 2 * real
    d = engine.output_dims(p)
    fixed = d.n_thk * (d.nx * d.ny * 3 + 2 * nxy) * 8
    return fixed + nxy * cplx + n_pot * nxy * real + 2 * nxy * cplx + nxy * real // 2


def _stem(precision=1, bwl=0):
    p = configs.au_stem_512(ns=3, nconf=2, n=256, ncells=4, nc=5)
    p.system_conf.precision = precision
    p.bwl = bwl
    return p


@pytest.mark.parametrize("precision,bwl", [(1, 0), (1, 1), (2, 0)])
def test_stem_rebuilt_slices_bit_identical(precision, bwl):
    p = _stem(precision, bwl)
    _, _, ns = _sizes(p)
    ref, plan0 = _run(p, batch=2)
    assert "slices=rebuilt" not in plan0
    for n_pot in (ns, ns // 2, 0):
        out, plan = _run(p, _stem_budget(p, n_pot))
        print("STEM precision %d bwl %d, %s" % (precision, bwl, plan))
        assert "slices=rebuilt stored_potentials=%d/%d" % (n_pot, ns) in plan, plan
        assert "batch=2 " in plan, plan
        assert np.array_equal(out.image_tot, ref.image_tot), n_pot
    r64 = oracle.run(p, "f64").image_tot
    assert rel_l2(ref.image_tot, r64) < (1e-4 if precision == 1 else 1e-9)


def test_cbed_rebuilt_slices_bit_identical():
    p = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=12, thick=[15.0, 30.0])
    _, _, ns = _sizes(p)
    ref, _ = _run(p)
    for n_pot in (ns, 5, 0):
        out, plan = _run(p, _ew_budget(p, n_pot))
        print("CBED %s" % plan)
        assert "slices=rebuilt stored_potentials=%d/%d" % (n_pot, ns) in plan, plan
        assert np.array_equal(out.m2psi_tot, ref.m2psi_tot), n_pot


@pytest.mark.parametrize("bwl", [0, 1])
def test_exit_wave_rebuilt_slices_bit_identical(bwl):
    p = configs.si_ewrs_256(nc=4)
    p.bwl = bwl
    _, _, ns = _sizes(p)
    ref, _ = _run(p)
    for n_pot in (ns, 3, 0):
        out, plan = _run(p, _ew_budget(p, n_pot))
        print("EWRS bwl %d %s" % (bwl, plan))
        assert "slices=rebuilt stored_potentials=%d/%d" % (n_pot, ns) in plan, plan
        assert np.array_equal(out.psi_coh, ref.psi_coh), n_pot
    assert rel_l2(ref.psi_coh, oracle.run(p, "f32").psi_coh) < 1e-5


def test_budget_below_one_slice_and_one_wave():
    p = _stem()
    nxy = p.nx * p.ny
    with pytest.raises(engine.EngineError) as ei:
        _run(p, 2 * nxy * 8)
    assert ei.value.status == OOM
    q = configs.si_ewrs_256(nc=4)
    with pytest.raises(engine.EngineError) as ei:
        _run(q, nxy * 8)
    assert ei.value.status == OOM
    # the engine is usable afterwards, at the default budget
    out, plan = _run(p)
    assert "slices=rebuilt" not in plan and np.isfinite(out.image_tot).all()
