"""Run control through the C ABI (Multislice::ext_stop_sim, ext_niter and
ext_iter, src/tem_simulation.cuh:47-49,103-133,512-519).

mt_hip_request_stop, called from a second host thread while a run is in
flight, ends that run with MT_ERR_STOPPED (the reference breaks out of its
loops and keeps the partial sums; here the call reports the stop), and the
engine runs cleanly afterwards.  mt_hip_progress rises monotonically from
(0, 0) to (niter, niter), niter being scan points x configurations for STEM.
"""
import threading
import time

import numpy as np
import pytest

from multem_amd import capi, configs, engine
from oracle import oracle
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

STOPPED = 5
assert capi.STATUS_NAMES[STOPPED] == "MT_ERR_STOPPED"


def _stem(ns=8, nconf=4):
    return configs.au_stem_512(ns=ns, nconf=nconf, n=256, ncells=4, nc=5)


def _start(p):
    res = {}

    def work():
        try:
            res["out"] = engine.run_multislice(p)
        except engine.EngineError as e:
            res["err"] = e

    th = threading.Thread(target=work)
    th.start()
    return th, res


def test_progress_rises_to_niter():
    p = _stem()
    engine.set_batch(8, True)                  # 8 batches x 4 configurations = 32 work items
    try:
        th, res = _start(p)
        seen = []
        while th.is_alive():
            seen.append(engine.progress())
            time.sleep(0.0002)
        th.join()
    finally:
        engine.set_batch(0, True)
    assert "out" in res, res.get("err")
    it, nit = engine.progress()
    assert nit == 8 * 8 * 4 and it == nit
    its = [a for a, _ in seen] + [it]
    assert all(b >= a for a, b in zip(its, its[1:])), "progress went backwards"
    assert all(0 <= a <= nit for a in its)
    assert all(b in (0, nit) for _, b in seen)
    print("progress samples: %d, distinct values %d" % (len(seen), len(set(its))))


def test_request_stop_from_second_thread():
    p = _stem(ns=16, nconf=8)
    niter = 16 * 16 * 8
    engine.set_batch(4, True)                  # 64 batches x 8 configurations: many stop points
    try:
        th, res = _start(p)
        t0 = time.time()

        def started():
            # this run's counters (an earlier run's stay until this one resets them)
            it, nit = engine.progress()
            return nit == niter and it > 0

        while th.is_alive() and not started() and time.time() - t0 < 60:
            time.sleep(0.0005)
        engine.request_stop()
        th.join()
        stopped_at = engine.progress()
    finally:
        engine.set_batch(0, True)
    assert "err" in res, "the run finished before the stop request"
    assert res["err"].status == STOPPED, res["err"]
    assert 0 < stopped_at[0] < stopped_at[1], stopped_at
    print("stopped at %d of %d" % stopped_at)
    # the stop flag is cleared by the next call, which runs to completion
    q = _stem(ns=4, nconf=2)
    out = engine.run_multislice(q)
    assert engine.progress() == (16 * 2, 16 * 2)
    e = rel_l2(out.image_tot, oracle.run(q, "f64").image_tot)
    assert e < 1e-4, e


def test_stop_before_a_run_is_cleared():
    """A stop requested between runs does not cancel the next run (the flag
    belongs to the run in flight, as ext_stop_sim is reset by the caller)."""
    engine.request_stop()
    q = _stem(ns=3, nconf=1)
    out = engine.run_multislice(q)
    assert np.isfinite(out.image_tot).all()
    assert engine.progress() == (9, 9)
