"""mt_hip_propagate_noalloc against mt_hip_propagate through the C ABI.

The allocation-free entry runs the same kernels on buffers carved from one
static per-GPU device array, so its output must equal the allocating entry's
bit for bit: FP32 and FP64, real and reciprocal space, tilted illumination
with the bandwidth limit, and two calls in a row on the shared array with
the second on a smaller grid (the array keeps the first call's contents).
A grid above the array's limit is refused with MT_ERR_INVALID_ARGUMENT.
"""
import numpy as np
import pytest

from multem_amd import configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _wave(p, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((p.nx, p.ny)) + 1j * rng.standard_normal((p.nx, p.ny))


@pytest.mark.parametrize("precision", [1, 2])
def test_noalloc_equals_propagate_two_calls(precision):
    big = configs.si_ewrs_256(n=512)
    small = configs.si_ewrs_256(n=256)
    for p in (big, small):
        p.system_conf.precision = precision
    for p, seed, z in ((big, 11, 7.5), (small, 12, -23.0)):
        psi = _wave(p, seed)
        for space in (1, 2):
            a = engine.propagate_noalloc(p, psi, z, space)
            b = engine.propagate(p, psi, z, space)
            assert np.array_equal(a, b), (p.nx, space)
            ref = oracle.propagate(p, psi, z, space, precision="f64")
            assert rel_l2(a, ref) < (1e-5 if precision == 1 else 1e-11), (p.nx, space)


def test_noalloc_tilt_and_bandwidth_limit():
    p = configs.si_ewrs_256()
    p.theta, p.phi, p.bwl = 0.4, 60.0, 1
    psi = _wave(p, 5)
    for z in (37.0, 0.0):
        for space in (1, 2):
            assert np.array_equal(engine.propagate_noalloc(p, psi, z, space), engine.propagate(p, psi, z, space))


@pytest.mark.parametrize("nx,ny", [(768, 640), (96, 1280)])
def test_noalloc_mixed_lengths(nx, ny):
    """Grids of 3 x 2^k and 5 x 2^k points through the allocation-free entry."""
    p = configs.si_ewrs_256()
    p.nx, p.ny = nx, ny
    psi = _wave(p, 7)
    for space in (1, 2):
        a = engine.propagate_noalloc(p, psi, 12.5, space)
        assert np.array_equal(a, engine.propagate(p, psi, 12.5, space)), space
        assert rel_l2(a, oracle.propagate(p, psi, 12.5, space, precision="f64")) < 1e-5


def test_noalloc_refuses_grid_above_limit():
    p = configs.si_ewrs_256(n=2048)
    assert p.nx * p.ny > engine.NOALLOC_MAX_NXY
    with pytest.raises(engine.EngineError) as ei:
        engine.propagate_noalloc(p, np.zeros((p.nx, p.ny), complex), 5.0, 1)
    assert ei.value.status == 1
