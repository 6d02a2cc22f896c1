"""Phase object and weak-phase object (interaction_model 2, 3) on the GPU
against the oracle, through the C ABI.  Tolerances (relative L2) as for
multislice runs: 1e-5 against the FP32 restatement and 1e-4 against FP64
for the FP32 path; 1e-9 for the FP64 path.  The oracle's models are pinned
in tests/test_phase_object.py."""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", [2, 3])
@pytest.mark.parametrize("precision", [1, 2])
def test_exit_wave(model, precision):
    p = configs.si_ewrs_256(nc=3)
    p.interaction_model = model
    p.system_conf.precision = precision
    out = engine.run_multislice(p)
    e64 = rel_l2(out.psi_coh, oracle.run(p, "f64").psi_coh)
    e32 = rel_l2(out.psi_coh, oracle.run(p, "f32").psi_coh)
    print("model %d precision %d: rel L2 vs FP64 %.2e, vs FP32 restatement %.2e" % (model, precision, e64, e32))
    if precision == 1:
        assert e32 < 1e-5 and e64 < 1e-4
    else:
        assert e64 < 1e-9


def test_stem_and_cbed_phase_object():
    p = configs.au_stem_512(ns=4, nconf=2, n=256, ncells=4, nc=3)
    p.interaction_model = 2
    e = rel_l2(engine.run_multislice(p).image_tot, oracle.run(p, "f64").image_tot)
    print("STEM phase object: rel L2 vs FP64 %.2e" % e)
    assert e < 1e-4
    q = configs.pt_cbed_1024(nconf=2, n=256, na=4, nb=6, nc=8, thick=[10.0, 20.0])
    q.interaction_model = 3
    q.bwl = 1
    out, ref = engine.run_multislice(q), oracle.run(q, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        print("CBED weak phase object, bwl, thickness %d: rel L2 vs FP64 %.2e" % (k, e))
        assert e < 1e-4


def test_single_plane_runs_as_phase_object():
    p = configs.si_ewrs_256(nc=1)
    a = np.asarray(p.spec_atoms)
    p.spec_atoms = a[np.abs(a[:, 3] - a[:, 3].min()) < 1e-3]
    out = engine.run_multislice(p)
    e = rel_l2(out.psi_coh, oracle.run(p, "f64").psi_coh)
    print("single atomic plane (requested as multislice): rel L2 vs FP64 %.2e" % e)
    assert e < 1e-4
