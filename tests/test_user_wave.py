"""User-defined incident wave (iw_type 3) in the oracle
(Incident_Wave, src/incident_wave.cuh:50-55, 92-110).

Pinned by the translation theorem: the user wave is moved to (iw_x[0],
iw_y[0]) by a phase ramp on its spectrum, so a convergent probe given as
the user wave at the box centre and moved to P equals the probe computed
at P; and a uniform user wave is the plane wave (only g = 0 is present).
"""
import numpy as np
import pytest

from multem_amd import configs
from oracle import oracle
from tests.conftest import rel_l2


def _probe_case():
    p = configs.au_stem_512(ns=4, nconf=1, n=256, ncells=4, nc=3)
    p.simulation_type = 21
    p.iw_type = 2
    p.cond_lens_c_10, p.cond_lens_c_30 = 30.0, 0.002
    return p


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_user_wave_is_translated(precision):
    p = _probe_case()
    p.iw_x, p.iw_y = np.array([0.5 * p.spec_lx]), np.array([0.5 * p.spec_ly])
    centred = oracle.incident_wave(p, precision="f64")          # natural layout [ix, iy]
    P = (0.3 * p.spec_lx, 0.65 * p.spec_ly)
    p.iw_x, p.iw_y = np.array([P[0]]), np.array([P[1]])
    want = oracle.incident_wave(p, precision="f64")
    q = p.copy()
    q.iw_type = 3
    q.iw_psi = centred.T                                        # MATLAB's ny x nx view
    got = oracle.incident_wave(q, precision=precision)
    e = rel_l2(got, want)
    print("translated user wave (%s): rel L2 %.2e" % (precision, e))
    assert e < (1e-12 if precision == "f64" else 1e-5)


def test_uniform_user_wave_is_the_plane_wave():
    p = configs.si_ewrs_256(nc=3)
    q = p.copy()
    q.iw_type = 3
    q.iw_psi = np.ones((p.ny, p.nx), dtype=np.complex128)
    q.iw_x, q.iw_y = np.array([0.37 * p.spec_lx]), np.array([0.81 * p.spec_ly])
    a, b = oracle.run(p, "f64").psi_coh, oracle.run(q, "f64").psi_coh
    e = rel_l2(b, a)
    print("uniform user wave vs plane wave: rel L2 %.2e" % e)
    assert e < 1e-12
