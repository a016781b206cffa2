"""The host-array entry point called from several host threads at once.

winds_bs_eval keeps one context per device (streams and grow-only buffers)
and serialises the calls that share it; each thread's error text is its
own.  Four threads evaluate different wakes (and one a bad call) at the
same time, 8 times each; every result must equal the same call made alone,
bit for bit (FP64 MEX arithmetic), and the failing call's message must not
leak into the others."""
import math
import threading

import numpy as np
import pytest

import winds
from conftest import flush_print

pytestmark = pytest.mark.gpu


def wake(seed, n_seg, n_fil, n_t):
    rng = np.random.default_rng(seed)
    F1 = rng.normal(0, 2, (n_seg, 3, n_fil, 1))
    F2 = F1 + rng.normal(0, 0.2, F1.shape)
    P = rng.normal(0, 2, (n_t, 3, 1, 1))
    return F1, F2, P, rng.normal(0, 1, n_seg * n_fil), np.full(n_seg * n_fil, 1e-3)


def call(w, type_="fuln"):
    F1, F2, P, g, rc = w
    return winds.BiotSavartMex(F1, F2, P, g, rc, 1e-4, "none", math.inf, type_)[0]


def test_concurrent_host_calls_equal_serial_calls(need_gpu):
    flush_print("[threads] 4 threads x 8 calls")
    wakes = [wake(s, 50 + 30 * s, 4 + s, 300 + 100 * s) for s in range(4)]
    ref = [call(w) for w in wakes]
    out = [[] for _ in wakes]
    errs = []

    def run(i):
        try:
            for _ in range(8):
                out[i].append(call(wakes[i]))
                if i == 3:
                    with pytest.raises(winds.WindsError, match="unknown type"):
                        call(wakes[i], "xxxx")
        except Exception as e:  # reported below
            errs.append((i, repr(e)))

    ts = [threading.Thread(target=run, args=(i,)) for i in range(4)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errs, errs
    for i in range(4):
        assert len(out[i]) == 8
        for u in out[i]:
            assert np.array_equal(u, ref[i]), i
