"""GPU test of the shared default context under threads: the drop-in calls
of several threads, each on its own stream, share one library context (and
its workspace); the library serialises the calls and orders every stream
switch after the work queued before it, so each thread's result equals the
sequential one (triangle sums and counts bit for bit; values that contain the
P(k) shell powers, whose float64 atomic bin sums are order-dependent, to
round-off)."""
import threading

import numpy as np
import pytest
import torch

from pyspectrum_amd import pyspectrum as P

pytestmark = pytest.mark.gpu


def _box(seed, n=20000, lbox=500.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, lbox, size=(3, n)), lbox


def _bk(xyz, lbox):
    return P.Bk_periodic(xyz, Lbox=lbox, Ngrid=48, step=3, Ncut=3, Nmax=8, precision="fp32")


def test_threads_share_the_default_context():
    print("[case] threads sharing the default context", flush=True)
    boxes = [_box(s) for s in (41, 42, 43)]
    ref = [_bk(x, l) for x, l in boxes]
    out = [None] * len(boxes)
    errors = []

    def work(i):
        try:
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                for _ in range(2):
                    out[i] = _bk(*boxes[i])
            s.synchronize()
        except Exception as e:  # surfaced below
            errors.append(e)

    ths = [threading.Thread(target=work, args=(i,)) for i in range(len(boxes))]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors, errors
    for a, b in zip(out, ref):
        assert np.array_equal(a["counts"], b["counts"])
        assert np.array_equal(a["i_k1"], b["i_k1"])
        for k in ("b123", "q123", "p0k1"):
            den = np.maximum(np.abs(b[k]), 1e-6 * np.max(np.abs(b[k])))
            assert np.max(np.abs(a[k] - b[k]) / den) < 1e-10, k
