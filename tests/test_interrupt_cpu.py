"""Progress, cancellation and Ctrl-C handling, on the host side (no GPU).

The C ABI keeps each device's progress / cancellation state in host memory,
so ppsk_progress, ppsk_cancel and ppsk_cancel_reset work without a device.
The pp_sketchlib module runs its distance calls on a worker thread while the
calling thread polls PyErr_CheckSignals, as the reference's distance loops do
(src/api.cpp:217-226, 285-287): _interruptProbe stands in for a long call
(it returns once the device's calls are cancelled), so a SIGINT during it
must raise KeyboardInterrupt promptly and leave no cancellation standing.
"""
import ctypes
import os
import signal
import threading
import time

import pytest

import ppsk_hip as H


def test_cancel_state_without_device():
    H.cancel_reset(3)
    st = H.progress(3)
    assert st["pending"] is False and st["running"] is False
    H.cancel(3)
    assert H.progress(3)["pending"] is True
    assert H.progress(2)["pending"] is False  # per device
    H.cancel_reset(3)
    assert H.progress(3)["pending"] is False


def test_cancel_rejects_bad_device():
    lib = H.lib()
    for dev in (-1, 64):
        assert lib.ppsk_cancel(dev) == H.ERR_ARG
        assert lib.ppsk_cancel_reset(dev) == H.ERR_ARG
        assert lib.ppsk_progress(dev, None, None, None, None) == H.ERR_ARG
    assert lib.ppsk_progress(0, None, None, None, None) == H.OK


def test_probe_returns_without_signal():
    import pp_sketchlib as pp
    H.cancel_reset(0)
    t = time.perf_counter()
    assert pp._interruptProbe(0.2, 0) is False
    assert time.perf_counter() - t >= 0.2


def test_probe_sees_cancel_from_another_thread():
    import pp_sketchlib as pp
    H.cancel_reset(1)
    threading.Timer(0.1, H.cancel, args=(1,)).start()
    try:
        assert pp._interruptProbe(5.0, 1) is True
    finally:
        H.cancel_reset(1)


def test_sigint_raises_keyboard_interrupt_promptly():
    import pp_sketchlib as pp
    H.cancel_reset(0)
    if threading.current_thread() is not threading.main_thread():
        pytest.skip("signals are delivered to the main thread only")
    timer = threading.Timer(0.3, os.kill, args=(os.getpid(), signal.SIGINT))
    t = time.perf_counter()
    timer.start()
    with pytest.raises(KeyboardInterrupt):
        pp._interruptProbe(20.0, 0)
    elapsed = time.perf_counter() - t
    timer.join()
    # the worker saw the cancellation and returned: far less than 20 s
    assert elapsed < 5.0
    # the request was withdrawn after the worker finished
    assert H.progress(0)["pending"] is False
