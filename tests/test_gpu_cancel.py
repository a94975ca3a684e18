"""Progress and cancellation of the distance calls on the GPU.

A cancellation (ppsk_cancel) asked for while a long pass runs makes its
kernels take no further work items; every workgroup finishes normally (no
trap) and the call returns PPSK_ERR_CANCELLED.  Checked for the persistent
kernel, the per-(tile, k) kernel and a batch: the call stops early (progress
below its total), and after ppsk_cancel_reset the same calls give results
equal to the oracle, so no counter or queue state is left behind.  A
request made before a call starts cancels it before any device work.
"""
import threading
import time

import numpy as np
import pytest

import oracle as O
import ppsk_hip as H
from helpers import KLIST8, synth

pytestmark = pytest.mark.gpu

N_BIG = 24576  # self pass of ~302M pairs, ~0.34 s on the persistent kernel
S64 = 156


@pytest.fixture(scope="module")
def big():
    raw, le, _ = H.synth(N_BIG, KLIST8, S64, seed=31)
    s = H.SketchSet(raw, KLIST8, S64, lengths=le)
    del raw
    yield s
    s.free()


def _run_and_cancel(call):
    """Runs call() on a thread, cancels once progress shows work done, and
    returns (exception or None, progress at the cancel, final progress)."""
    H.cancel_reset(0)
    box = {}

    def target():
        try:
            call()
        except H.PpskError as e:
            box["err"] = e

    th = threading.Thread(target=target)
    th.start()
    seen = None
    t0 = time.perf_counter()
    while th.is_alive() and time.perf_counter() - t0 < 60:
        st = H.progress(0)
        if st["running"] and st["done"] > 0:
            H.cancel(0)
            seen = st
            break
        time.sleep(0.0005)
    th.join()
    final = H.progress(0)
    H.cancel_reset(0)
    return box.get("err"), seen, final


def _check_small_against_oracle():
    raw, le, _ = synth(300)
    with H.SketchSet(raw, KLIST8, S64, lengths=le) as s:
        d, nf, _ = H.dense(s, random_mode=H.RANDOM_EQUAL_BASE)
        c, _, _ = H.dense(s, out="counts")
    want, counts, rnf, _ = O.query_db(raw, KLIST8, S64, 14, le, random=O.Random(1, True),
                                      want_counts=True)
    assert nf == rnf == 0
    assert np.array_equal(c, counts)
    assert np.abs(d - want).max() <= 1e-6


def test_cancel_before_call_does_no_work():
    raw, le, _ = synth(300)
    with H.SketchSet(raw, KLIST8, S64, lengths=le) as s:
        H.cancel(0)
        try:
            with pytest.raises(H.PpskError) as e:
                H.dense(s)
            assert e.value.code == H.ERR_CANCELLED
        finally:
            H.cancel_reset(0)
        d, nf, _ = H.dense(s)
    assert nf == 0 and d.shape == (300 * 299 // 2, 2)


def test_progress_of_a_finished_call(big):
    H.cancel_reset(0)
    dptr, nf = H.dense_device(big, random_mode=H.RANDOM_EQUAL_BASE)
    try:
        st = H.progress(0)
        assert nf == 0
        assert not st["running"] and st["total"] > 0
        assert st["done"] == st["total"]  # every (tile, k) item reported
    finally:
        H.device_free(dptr)


@pytest.mark.parametrize("persistent", ["1", "0"])
def test_cancel_mid_pass(big, monkeypatch, persistent):
    # "1": one persistent launch (workgroups stop taking tiles); "0": chunks of
    # the per-(tile, k) kernel (workgroups skip their matching but keep the
    # completion counters consistent)
    monkeypatch.setenv("PPSK_PERSISTENT", persistent)
    rows = N_BIG * (N_BIG - 1) // 2
    dptr = H.device_alloc(rows * 8)
    try:
        err, seen, final = _run_and_cancel(
            lambda: H.dense_device(big, random_mode=H.RANDOM_EQUAL_BASE, dptr=dptr))
    finally:
        H.device_free(dptr)
    assert seen is not None, "the pass ended before any progress was seen"
    assert err is not None and err.code == H.ERR_CANCELLED
    assert final["done"] < final["total"]
    assert not final["running"]
    # the device state is intact: same calls, correct results
    _check_small_against_oracle()


def test_cancel_batch_then_rerun():
    raw, le, _ = synth(2000)
    with H.SketchSet(raw[:1500], KLIST8, S64) as r, H.SketchSet(raw[1500:], KLIST8, S64) as q:
        with H.Batch([dict(ref=r, query=q), dict(ref=r.prefix(700), query=q.prefix(3))]) as b:
            # cancelled before the launch: no items taken
            H.cancel(0)
            try:
                with pytest.raises(H.PpskError) as e:
                    b.run(2)
                assert e.value.code == H.ERR_CANCELLED
            finally:
                H.cancel_reset(0)
            nf, used, _ = b.run(2)
            assert used == 2 and (nf == 0).all()
            want, _, _ = H.dense(r, q)
            assert np.array_equal(b.result(0).view(np.uint8), want.view(np.uint8))
