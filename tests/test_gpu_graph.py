"""GPU tests of the captured B(k) graph (pyspectrum_amd/graphs.py).

Every replay is checked against the eager psx_bk_raw result of the same half
field: the triangle sums bit for bit (shell fields, contraction and its
fixed-order reduction are deterministic), the shell powers p0k to float64
round-off (their k-space bins are float64 atomic sums).  Replays alternate
between two inputs, with eager library work and host synchronisation between
them, so a graph that kept stale state or mis-ordered a node would fail."""
import numpy as np
import pytest
import torch

from pyspectrum_amd import _lib as L
from pyspectrum_amd import graphs, mocks, serve

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _half(ctx, xyz, lbox, N, prec, out):
    n = xyz.shape[1]
    L.call("psx_fft_periodic", ctx.handle, L.ptr(xyz), None, n, N, float(lbox), float(n), float(n),
           L.ptr(out), L.PREC[prec])


def _eager(ctx, half, N, step, ncut, nmax, prec):
    ntri = len(L.triangles(step, ncut, nmax))
    braw = torch.empty(ntri, dtype=torch.float64, device=half.device)
    p0k = torch.empty(nmax, dtype=torch.float64, device=half.device)
    L.call("psx_bk_raw", ctx.handle, L.ptr(half), N, step, ncut, nmax, L.ptr(braw), L.ptr(p0k),
           L.PREC[prec])
    return braw.cpu().numpy(), p0k.cpu().numpy()


@pytest.mark.parametrize("prec", ["bf16x3", "fp32", "fp64"])
def test_replays_match_eager(prec):
    N, step, ncut, nmax, lbox = 64, 3, 3, 12, 500.0
    _say("graph replays", prec, "N", N, "Nmax", nmax)
    g = graphs.BkGraph(N, step, ncut, nmax, prec)
    ctx = L.Context(0)
    refs = []
    for seed in (11, 12):
        xyz, _ = mocks.lognormal_box(20000, lbox, seed=seed, mesh=64)
        h = torch.empty_like(g.half)
        _half(ctx, xyz, lbox, N, prec, h)
        refs.append((h, _eager(ctx, h, N, step, ncut, nmax, prec)))
    assert not np.array_equal(refs[0][1][0], refs[1][1][0])
    for rep in range(3):
        for k, (h, (b_ref, p_ref)) in enumerate(refs):
            g.half.copy_(h)
            g.replay()
            b, p = g.braw.cpu().numpy(), g.p0k.cpu().numpy()
            assert np.array_equal(b, b_ref), (prec, rep, k)
            assert np.allclose(p, p_ref, rtol=1e-12, atol=0), (prec, rep, k)
            # eager library work on the device between replays
            _eager(ctx, refs[1 - k][0], N, step, ncut, nmax, prec)
            torch.cuda.synchronize()
    g.close()
    ctx.close()


def test_server_with_graphs_matches_eager():
    _say("serve with graphs vs eager")
    reqs = serve.make_stream(4, rate=0.0, seed=9, grids=(64, 96), boxes=(500., 1000.),
                             nrange=(2e4, 1e5), mesh=64)
    for r in reqs:
        r.Nmax = 12
    eager = serve.Server(nslots=2)
    out_e, _ = eager.serve(reqs)
    res_e = {r.rid: r.result for r in out_e}
    eager.close()
    for r in reqs:
        r.result = None
    server = serve.Server(nslots=2, use_graphs=True)
    out_g, _ = server.serve(reqs)
    assert server.graph_build_s > 0 and all(s.graphs for s in server.slots)
    for r in out_g:
        a, b = res_e[r.rid], r.result
        assert np.array_equal(a["b_counts"], b["b_counts"]), r.rid
        for key in ("b123", "q123", "p0k1", "b123_sn", "p0k", "p2k"):
            den = np.maximum(np.abs(a[key]), np.max(np.abs(a[key])) * 1e-6)
            assert np.max(np.abs(a[key] - b[key]) / den) < 1e-10, (r.rid, key)
    server.close()
