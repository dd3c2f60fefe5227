"""Context teardown: spk_ctx_destroy (the Session destructor) returns every
device buffer the context allocated, the column-test and loop buffers
included."""
import gc

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def P():
    import psacking_amd
    return psacking_amd


def free_bytes():
    """hipMemGetInfo's free device memory on the calling thread's device"""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")   # the runtime libspk.so already loaded
    free, total = C.c_size_t(), C.c_size_t()
    assert hip.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
    return free.value


def test_session_teardown_releases_device_memory(P):
    """After two warm-up cycles, six more create/pack/destroy cycles on a
    240x123x100 tray leave free device memory where it was.  A live context
    holds far more than the bound (its spectra alone take hundreds of MiB),
    and the column-test buffers of one such context take several MiB, so a
    context that kept them would exceed the bound within the six cycles."""
    from psacking_amd import workloads as W
    from psacking_amd.rotations import orientation_maps
    w = W.readme(seed=5, n=6)
    perms, flips = orientation_maps(24)
    order = np.arange(len(w.items), dtype=np.int32)

    def cycle():
        s = P.Session(0)
        s.set_tray(np.zeros(w.tray_size, np.int32), "edt", "nonzero", 0)
        s.pack(list(w.items), order, perms, flips, "cube", False, 0)
        held = free_bytes()
        del s
        gc.collect()
        return held

    say("teardown: 2 warm-up cycles, README tray, 6 items")
    cycle()
    cycle()
    free0 = free_bytes()
    say("teardown: 6 measured cycles")
    held = [cycle() for _ in range(6)]
    free1 = free_bytes()
    say("teardown: free MiB before", free0 >> 20, "after", free1 >> 20,
        "least while a context lived", min(held) >> 20)
    assert free0 - min(held) > 64 << 20
    assert free0 - free1 < 8 << 20
