"""GPU tests of the continuous-batching request loop (pyspectrum_amd/serve.py):
requests overlap on the device, and each result equals the result of the
same request run alone bit for bit (every stage, the P_l binning included,
sums in a fixed order)."""
import numpy as np
import pytest
import torch

from pyspectrum_amd import serve

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _stream():
    return serve.make_stream(6, rate=200.0, seed=7, grids=(64, 96, 128), boxes=(500., 1000.),
                             nrange=(2e4, 2e5), mesh=128)


def _shrink(reqs, nmax=12):
    for r in reqs:
        r.Nmax = nmax
    return reqs


def test_results_identical_alone_and_in_mix():
    _say("serve mix vs alone")
    mix, wall = serve.Server(nslots=3).serve(_shrink(_stream()))
    assert all(r.result is not None for r in mix)
    # concurrency actually happened: some request started before another finished
    starts = np.array([r.t_start for r in mix]); ends = np.array([r.t_done for r in mix])
    assert np.any(starts[1:] < ends[:-1])
    alone_server = serve.Server(nslots=1)
    for r in mix:
        ref = serve.Request(r.rid, r.xyz, r.Lbox, r.Ngrid, Nmax=r.Nmax, precision=r.precision)
        out, _ = alone_server.serve([ref])
        a, b = out[0].result, r.result
        for key in ("b_counts", "b123", "q123", "p0k1", "p0k2", "p0k3", "b123_sn", "counts",
                    "p0k", "p2k", "p4k", "k"):
            assert np.array_equal(a[key], b[key], equal_nan=True), (r.rid, key)


def test_admission_respects_budget():
    _say("serve admission budget")
    reqs = _shrink(serve.make_stream(4, rate=0.0, seed=8, grids=(64,), boxes=(500.,),
                                     nrange=(2e4, 5e4), mesh=64))
    need = serve.footprint_bytes(64, 5e4, 12)
    # a budget below two requests forces one at a time
    server = serve.Server(nslots=4, hbm_budget=int(1.5 * need))
    out, _ = server.serve(reqs)
    iv = sorted((r.t_start, r.t_done) for r in out)
    for (s0, e0), (s1, e1) in zip(iv, iv[1:]):
        assert s1 >= e0 - 1e-3


def test_failed_request_is_isolated():
    # an unsupported grid (62 = 2 x 31) fails with the library's error; the
    # other requests complete and equal their results run alone
    _say("serve failure isolation")
    reqs = _shrink(serve.make_stream(4, rate=0.0, seed=10, grids=(48,), boxes=(500.,),
                                     nrange=(2e4, 5e4), mesh=64))
    reqs[1].Ngrid = 62
    out, _ = serve.Server(nslots=2).serve(reqs)
    bad = [r for r in out if r.rid == reqs[1].rid][0]
    assert bad.result is None and bad.error is not None and "Ngrid must be" in bad.error
    good = [r for r in out if r.error is None]
    assert len(good) == 3 and all(r.result is not None for r in good)
    alone = serve.Server(nslots=1)
    for r in good:
        ref = serve.Request(r.rid, r.xyz, r.Lbox, r.Ngrid, Nmax=r.Nmax, precision=r.precision)
        o, _ = alone.serve([ref])
        assert np.array_equal(o[0].result["b_counts"], r.result["b_counts"])
    st = serve.latency_stats(out)
    assert np.isfinite(st["p50_ms"])
