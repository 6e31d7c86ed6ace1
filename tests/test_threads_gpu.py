"""Concurrent calls from several host threads.

The reference's engine is synchronous and reentrant per call (fd.h); so is
this one: every call allocates its own fields and streams, the last-error
string is per thread, and the tuner and profiler tables are locked.  ctypes
releases the GIL around foreign calls, so the threads below are inside the
library at the same time.  Every result must still equal the reference's
golden vectors bit for bit.
"""
import threading

import numpy as np
import pytest

import test_gpu_parity as parity
from conftest import load_fixture

pytestmark = pytest.mark.gpu


def test_concurrent_calls_match_reference(engine):
    cases = [load_fixture(n) for n in parity.SMALL]
    errors = []

    def worker(t):
        try:
            for rep in range(3):
                c = cases[(t + rep) % len(cases)]
                a = parity._args(c)
                assert np.array_equal(engine.modeling(*a), c["rec"]), ("modeling", t, rep)
                assert np.array_equal(engine.forward(*a, c["dv"]), c["born"]), ("born", t, rep)
                assert np.array_equal(engine.adjoint(*a, c["rec_in"]), c["img"]), ("rtm", t, rep)
        except BaseException as e:  # reported by the main thread
            errors.append(e)

    print("case 4 threads x 3 rounds of modeling/born/rtm", flush=True)
    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=100)
    assert not any(th.is_alive() for th in threads), "a worker thread did not finish"
    assert not errors, errors


def test_argument_errors_are_per_thread(engine):
    """An argument error in one thread leaves another thread's last-error
    string alone."""
    from seis_amd._lib import SeisError
    msgs = {}
    gz = np.zeros(4, np.int32)
    bad = threading.Event()

    def failing():
        try:
            engine.modeling(0.001, 10, 10, 10, 0, 4, 4, 2, 9, 0, gz, np.zeros(10, np.float32),
                            np.ones((4, 4), np.float32))
        except SeisError:
            msgs["failing"] = engine.lib().seis_last_error().decode()
        bad.set()

    def clean():
        c = load_fixture(parity.SMALL[0])
        bad.wait(timeout=60)
        engine.modeling(*parity._args(c))
        msgs["clean"] = engine.lib().seis_last_error().decode()

    ts = [threading.Thread(target=failing), threading.Thread(target=clean)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=100)
    assert "source" in msgs["failing"]
    assert "source" not in msgs["clean"]
