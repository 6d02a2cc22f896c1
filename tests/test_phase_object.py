"""The (weak) phase-object interaction models in the oracle
(interaction_model 2 and 3; src/slicing.hpp:326-355, src/spec.hpp:65-77,
src/cgpu_fcns.cuh:3071-3086).

Pinned by what the model means: with a plane incident wave, no
back-propagation (zero-defocus plane at the exit face) and no bandwidth
limit, the exit wave of the phase object is exp(i sigma sum_slices V_s) and
that of the weak-phase object 1 + i sigma sum_slices V_s, where V_s are the
multislice slices' projected potentials (every atom counted once, no
propagation inside the specimen).  A specimen of a single atomic plane is a
phase object whatever interaction model is asked for.
"""
import numpy as np
import pytest

from multem_amd import configs
from oracle import oracle
from tests.conftest import rel_l2


def _sigma_vsum(p, precision):
    q = p.copy()
    q.interaction_model = 1
    ns = len(oracle.spec_slicing(q, precision=precision))
    V = sum(oracle.projected_potential(q, i, precision=precision) for i in range(ns))
    L = oracle.lib(precision)
    w = L.orc_vr_factor(float(p.E_0), 0.0)
    return w * V


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_phase_object_exit_wave(precision):
    p = configs.si_ewrs_256(nc=3)
    p.interaction_model = 2
    psi = oracle.run(p, precision).psi_coh[0]
    want = np.exp(1j * _sigma_vsum(p, precision))
    e = rel_l2(psi, want)
    print("phase object (%s): rel L2 vs exp(i sigma sum V) %.2e" % (precision, e))
    assert e < (1e-12 if precision == "f64" else 2e-6)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_weak_phase_object_exit_wave(precision):
    p = configs.si_ewrs_256(nc=3)
    p.interaction_model = 3
    psi = oracle.run(p, precision).psi_coh[0]
    want = 1 + 1j * _sigma_vsum(p, precision)
    e = rel_l2(psi, want)
    print("weak phase object (%s): rel L2 vs 1 + i sigma sum V %.2e" % (precision, e))
    assert e < (1e-12 if precision == "f64" else 2e-6)


def _one_plane(p):
    a = np.asarray(p.spec_atoms)
    z0 = a[:, 3].min()
    p.spec_atoms = a[np.abs(a[:, 3] - z0) < 1e-3]
    return p


def test_single_plane_is_a_phase_object():
    p = _one_plane(configs.si_ewrs_256(nc=1))
    assert oracle.run(p, "f64").psi_coh is not None
    q = p.copy()
    q.interaction_model = 2
    a, b = oracle.run(p, "f64").psi_coh, oracle.run(q, "f64").psi_coh
    assert np.array_equal(a, b)
    # one slice, and the multislice request (model 1) ran as that phase object
    assert len(oracle.spec_slicing(p)) == 1
