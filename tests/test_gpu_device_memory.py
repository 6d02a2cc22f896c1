"""Device blocks kept between runs (mt_hip_kept_device_bytes,
mt_hip_release_device_memory; no counterpart in the reference, which
allocates and frees its buffers in every call): a run leaves its blocks
kept, the next run of the same simulation reuses them and gives the same
image bit for bit, the release returns them to the runtime, and a run after
the release is again bit-identical (buffers are cleared where they
accumulate, not assumed zero from a fresh allocation).  A run of another
size in between (other block sizes) changes nothing either."""
import numpy as np
import pytest

from multem_amd import configs, engine

pytestmark = pytest.mark.gpu


def test_kept_blocks_reused_and_released():
    p = configs.au_stem_512(ns=4, nconf=2, n=256, ncells=4, nc=5)
    engine.release_device_memory(0)
    assert engine.kept_device_bytes(0) == 0
    a = engine.run_multislice(p).image_tot.copy()
    kept = engine.kept_device_bytes(0)
    print("kept after one run: %.1f MB" % (kept / 1e6))
    assert kept > 0
    b = engine.run_multislice(p).image_tot.copy()
    assert np.array_equal(a, b)
    assert engine.kept_device_bytes(0) == kept
    q = configs.au_stem_512(ns=4, nconf=1, n=512, ncells=4, nc=3)
    engine.run_multislice(q)
    c = engine.run_multislice(p).image_tot.copy()
    assert np.array_equal(a, c)
    engine.release_device_memory(0)
    assert engine.kept_device_bytes(0) == 0
    d = engine.run_multislice(p).image_tot.copy()
    assert np.array_equal(a, d)


def test_kept_blocks_count_as_free_for_the_planner():
    """The plan of a run does not change because an earlier run's blocks
    are kept."""
    p = configs.au_stem_512(ns=6, nconf=2, n=256, ncells=4, nc=5)
    engine.release_device_memory(0)
    engine.run_multislice(p)
    plan0 = engine.last_plan()
    engine.run_multislice(p)
    assert engine.last_plan() == plan0


def test_release_rejects_a_bad_device():
    with pytest.raises(engine.EngineError):
        engine.release_device_memory(-1)
