"""Specimen rotation (spec_rot_theta, spec_rot_u0, spec_rot_center_*;
src/input_multislice.cuh:707-722) in the oracle, pinned by scipy's
rotation: the atoms the oracle slices are the input atoms turned by theta
about the unit axis u0 through the chosen centre (right-handed)."""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from multem_amd import configs
from oracle import oracle


def _cluster():
    p = configs.si_ewrs_256(nc=2)
    p.spec_lx = p.spec_ly = 40.0
    p.spec_lz = 20.0
    rng = np.random.default_rng(7)
    xyz = 20.0 + rng.uniform(-4, 4, size=(12, 3))
    xyz[:, 2] -= 10.0
    p.spec_atoms = np.column_stack([np.full(12, 14.0), xyz, np.full(12, 0.085), np.ones(12), np.zeros(12),
                                    np.zeros(12)])
    p.thick_type = 2
    p.thick = np.array([5.0])
    return p


def _sorted(a):
    a = np.asarray(a)[:, :3]
    return a[np.lexsort((a[:, 1], a[:, 0], a[:, 2]))]


@pytest.mark.parametrize("theta,u0,center", [(30.0, [0, 0, 1], 1), (75.0, [1, 1, 0], 1),
                                              (20.0, [0.3, -0.5, 0.8], 2)])
def test_rotated_atoms(theta, u0, center):
    p = _cluster()
    p.spec_rot_theta, p.spec_rot_u0, p.spec_rot_center_type = theta, np.array(u0, float), center
    p.spec_rot_center_p = np.array([19.0, 21.0, 9.0])
    xyz = np.asarray(p.spec_atoms)[:, 1:4]
    c = xyz.mean(axis=0) if center == 1 else p.spec_rot_center_p
    u = np.asarray(u0, float) / np.linalg.norm(u0)
    want = Rotation.from_rotvec(np.deg2rad(theta) * u).apply(xyz - c) + c
    got = oracle.displaced_atoms(p, 1)
    assert len(got) == len(want)
    assert np.abs(_sorted(got) - _sorted(want)).max() < 1e-9
    # a rotated specimen is simulated whole (thick_type 1)
    assert len(oracle.run(p, "f64").thick) == 1


def test_full_turn_is_the_identity():
    p = _cluster()
    p.thick_type = 1
    q = p.copy()
    q.spec_rot_theta, q.spec_rot_u0 = 360.0, np.array([0.2, 0.4, 1.0])
    assert np.array_equal(oracle.run(p, "f64").psi_coh, oracle.run(q, "f64").psi_coh)


@pytest.mark.parametrize("pot_sli", [1, 2])
def test_engine_slicing_of_a_rotated_specimen(pot_sli):
    """The C ABI's host model slices the rotated specimen as the oracle does
    (slice bounds, atom ranges, thickness slice); no device work."""
    from multem_amd import engine
    p = configs.si_ewrs_256(nc=4)
    p.spec_rot_theta, p.spec_rot_u0 = 12.0, np.array([1.0, 0.3, 0.0])
    p.potential_slicing = pot_sli
    a, b = engine.spec_slicing(p), oracle.spec_slicing(p, precision="f32")
    assert len(a) == len(b)
    for x, y in zip(a, b):
        assert np.allclose(x[:4], y[:4], rtol=0, atol=1e-4) and tuple(x[4:]) == tuple(y[4:]), (x, y)
