"""The library's error state is per host thread (pp_last_error: "the last
failing call on this thread"), which the thread safety across distinct
handles rests on. Argument checks fail before any device call, so this runs
without a GPU."""

import sys
import threading

from pawpyseed_amd.core._lib import PP_OK, lib


def test_last_error_is_per_thread():
    L = lib()
    cases = {
        "ctx": (lambda: L.pp_ctx_create(None, 0), "pp_ctx_create"),
        "rebuild": (lambda: L.pp_sphere_rebuild(None, None), "pp_sphere_rebuild"),
    }
    n = 2000
    start = threading.Barrier(len(cases))
    bad = {}

    def work(name):
        call, token = cases[name]
        start.wait(timeout=30)
        wrong = 0
        for _ in range(n):
            status = call()
            msg = L.pp_last_error().decode()
            # another thread's failure in between must not change this thread's message
            if status == PP_OK or token not in msg or L.pp_last_status() != status:
                wrong += 1
        bad[name] = wrong

    # switch threads as often as the interpreter allows, so that the other
    # thread's failing call often lands between a call and its message read
    old_interval = sys.getswitchinterval()
    sys.setswitchinterval(1e-6)
    try:
        threads = [threading.Thread(target=work, args=(k,)) for k in cases]
        for t in threads:
            t.start()
        for t in threads:
            t.join(timeout=60)
    finally:
        sys.setswitchinterval(old_interval)
    assert not any(t.is_alive() for t in threads)
    assert bad == {k: 0 for k in cases}, bad
