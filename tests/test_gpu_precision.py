"""Precision study on thick CBED: Pt [110], 200 kV, 10 mrad, -100 A defocus,
20 frozen-phonon configurations, thicknesses every 50 A down to 450 A
(325 slices).  The engine (FP32 wave and FFTs, FP64 phase arguments and
accumulation) against the full-FP64 oracle, relative L2 of the diffraction
pattern at every thickness.

At 512^2 (the lateral box is the same 10 x 14 projected cells) the engine is
compared with the oracle at full depth and phonon count.  The full 1024^2
study, FP32 path against the engine's FP64 path (which the oracle pins at
1024^2 separately; the oracle itself would need minutes at this size), is
tests/test_gpu_large_grids.py::test_cbed_pt110_1024_full_depth_precision_study.
"""
import os

import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _threads():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return max(1, min(n, 16))


def _study(n):
    p = configs.pt_cbed_1024(nconf=20, n=n)
    out = engine.run_multislice(p)
    ref = oracle.run(p, "f64", nthreads=_threads())
    errs = []
    for k in range(out.n_thk):
        e = rel_l2(out.m2psi_tot[k], ref.m2psi_tot[k])
        errs.append(e)
        print("Pt[110] CBED %d^2, thickness %5.1f A: rel L2 vs FP64 %.3e" % (n, p.thick[k], e))
    return np.array(errs)


def test_cbed_pt110_full_depth_512():
    errs = _study(512)
    assert len(errs) == 9
    assert errs.max() < 1e-4
