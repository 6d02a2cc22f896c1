"""Grids of 3 x 2^k and 5 x 2^k points (the reference plans any size through
FFTW / cuFFT, src/fft.cuh:119-135,422-429): the mixed-radix slice passes
(radix-3 and radix-5 butterflies as the last Stockham stage) against the CPU
oracle, whose FFT splits an m x 2^k transform into m radix-2 transforms and
one direct radix-m combination (pinned against numpy in
tests/test_oracle_pins.py).

Tolerances (relative L2), as for the power-of-two grids: 1e-5 for single
pieces and 1e-4 for whole runs (FP32 path vs the FP64 oracle); 1e-11 and
1e-9 for the FP64 path.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

LENGTHS_3 = [96, 192, 384, 768, 1536]
LENGTHS_5 = [80, 160, 320, 640, 1280]
TOL_PIECE = {1: 1e-5, 2: 1e-11}
TOL_RUN = {1: 1e-4, 2: 1e-9}


def _ew(nx, ny, precision, nc=3):
    p = configs.si_ewrs_256(nc=nc, n=128)
    p.nx, p.ny = nx, ny
    p.system_conf.precision = precision
    return p


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("nx,ny", [(96, 80), (160, 192), (384, 320), (640, 768), (1536, 1280), (1280, 64), (128, 1536)])
def test_propagate_mixed(nx, ny, precision):
    """One propagation in real and in reciprocal space: the column and row
    passes of every new length, alone and beside a power of two."""
    p = _ew(nx, ny, precision)
    rng = np.random.default_rng(nx * 7 + ny)
    psi = (rng.standard_normal((nx, ny)) + 1j * rng.standard_normal((nx, ny))) / np.sqrt(2 * nx * ny)
    for space in (1, 2):
        got = engine.propagate(p, psi, 7.5, space)
        ref = oracle.propagate(p, psi, 7.5, space, precision="f64")
        e = rel_l2(got, ref)
        print("%dx%d precision %d space %d: rel L2 vs FP64 oracle %.2e" % (nx, ny, precision, space, e))
        assert e < TOL_PIECE[precision]


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("n", LENGTHS_3 + LENGTHS_5)
def test_probe_mixed(n, precision):
    p = configs.au_stem_512(ns=2, nconf=1, n=n, ncells=4, nc=3)
    p.system_conf.precision = precision
    p.simulation_type = 21
    p.cond_lens_c_10, p.cond_lens_c_30 = 25.0, 0.001
    p.cond_lens_c_12, p.cond_lens_phi_12 = 10.0, 20.0
    p.iw_x = np.array([0.37 * p.spec_lx])
    p.iw_y = np.array([0.58 * p.spec_ly])
    e = rel_l2(engine.incident_wave(p), oracle.incident_wave(p, precision="f64"))
    print("probe %d precision %d: rel L2 vs FP64 oracle %.2e" % (n, precision, e))
    assert e < TOL_PIECE[precision]


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("nx,ny", [(96, 96), (160, 384), (768, 640), (1536, 1280)])
def test_exit_wave_mixed(nx, ny, precision):
    """Si exit wave (EWRS) through every slice of the run."""
    p = _ew(nx, ny, precision)
    out = engine.run_multislice(p)
    ref = oracle.run(p, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.psi_coh[k], ref.psi_coh[k])
        print("EWRS %dx%d precision %d thickness %d: rel L2 vs FP64 oracle %.2e" % (nx, ny, precision, k, e))
        assert e < TOL_RUN[precision]


@pytest.mark.parametrize("precision", [1, 2])
@pytest.mark.parametrize("nx,ny", [(192, 192), (320, 384), (640, 640)])
def test_stem_mixed(nx, ny, precision):
    """STEM with phonons: probe pass, slice loop, detector epilogue."""
    p = configs.au_stem_512(ns=3, nconf=2, n=nx, ncells=4, nc=4)
    p.nx, p.ny = nx, ny
    p.system_conf.precision = precision
    p.detector_cir = [(0.0, 15.0), (20.0, 60.0)]
    out = engine.run_multislice(p)
    ref = oracle.run(p, "f64")
    for k in range(out.n_thk):
        e = rel_l2(out.image_tot[k], ref.image_tot[k])
        print("STEM %dx%d precision %d thickness %d: rel L2 vs FP64 oracle %.2e" % (nx, ny, precision, k, e))
        assert e < TOL_RUN[precision]
