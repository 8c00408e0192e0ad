"""GPU tests of the batched B(k) path (psx_bk_periodic_batch, one launch per
stage for up to 16 same-shape boxes) and of the request loop's continuous
batching: every box's outputs bit-identical to the single-box path."""
import numpy as np
import pytest
import torch

from pyspectrum_amd import _lib as L
from pyspectrum_amd import batch as B
from pyspectrum_amd import mocks
from pyspectrum_amd import pyspectrum as P
from pyspectrum_amd import serve

pytestmark = pytest.mark.gpu
KEYS = ("i_k1", "i_k2", "i_k3", "p0k1", "p0k2", "p0k3", "b123", "q123", "counts", "b123_sn")


def _say(*a):
    print("[case]", *a, flush=True)


def _same(a, b):
    for k in KEYS:
        assert np.array_equal(a[k], b[k], equal_nan=True), k
    assert a["p0k_sn"] == b["p0k_sn"]


@pytest.mark.parametrize("prec", ["bf16x3", "fp32", "fp64"])
def test_batch_equals_single_boxes(prec):
    """Five boxes of uneven size (7 to 1.2e5 points), box length and weights
    at Ngrid 64, Nmax 12: the split-BF16 batch runs the fused contraction for
    all boxes at once, FP32/FP64 their staged contraction box by box."""
    _say("batch vs single", prec)
    rng = np.random.default_rng(21)
    sizes, boxes = (2000, 50000, 300, 120000, 7), (500.0, 800.0, 500.0, 1000.0, 650.0)
    xyzs = [rng.uniform(0, L_, (3, n)).astype(np.float32) for n, L_ in zip(sizes, boxes)]
    ws = [None, rng.uniform(0.5, 1.5, 50000), None, None, None]
    got = B.Bk_periodic_batch(xyzs, ws=[w if w is not None else np.ones(x.shape[1])
                                        for w, x in zip(ws, xyzs)],
                              Lbox=list(boxes), Ngrid=64, step=3, Ncut=3, Nmax=12, precision=prec)
    for x, w, L_, g in zip(xyzs, ws, boxes, got):
        ref = P.Bk_periodic(x, w=w if w is not None else np.ones(x.shape[1]), Lbox=L_, Ngrid=64,
                            step=3, Ncut=3, Nmax=12, precision=prec)
        _same(g, ref)


def test_batch_clustered_boxes_full_size():
    """Three clustered boxes at the single-box workload's shape (Ngrid 360,
    Nmax 40, split-BF16), 2e5 to 1e6 points: one batched call equals three
    single calls bit for bit."""
    _say("batch 3 boxes at 360/40")
    xyzs = [mocks.lognormal_box(n, 2500.0, seed=s)[0] for n, s in ((1_000_000, 2), (200_000, 3),
                                                                   (500_000, 4))]
    got = B.Bk_periodic_batch(xyzs, Lbox=2500.0, Ngrid=360, step=3, Ncut=3, Nmax=40,
                              precision="bf16x3")
    for x, g in zip(xyzs, got):
        _same(g, P.Bk_periodic(x, Lbox=2500.0, Ngrid=360, step=3, Ncut=3, Nmax=40,
                               precision="bf16x3"))
    del xyzs
    torch.cuda.empty_cache()


def test_batch_limits():
    _say("batch limits")
    with pytest.raises(ValueError):
        B.run_batch(L.context(0), [], None, [], 64, 3, 3, 12, "fp32", 0)
    with pytest.raises(L.PsxError):
        L.context(0).call("psx_bk_periodic_batch", 17, (L.PsxBox * 17)(), 64, 3, 3, 12, None,
                          None, 1.0, 1.0, None, None, None, 1)


def test_server_batches_equal_alone():
    """The request loop with batches of up to 8 same-shape requests (16
    requests over two grids, submitted at once): P_l and B of every request
    bit-identical to the same request run alone."""
    _say("server batching")
    reqs = serve.make_stream(16, rate=0.0, seed=31, grids=(128, 192), nrange=(2e5, 1e6),
                             mesh=128)
    server = serve.Server(nslots=2, max_batch=8)
    mix, _ = server.serve(reqs)
    assert all(r.error is None for r in mix), [r.error for r in mix if r.error]
    server.close()
    alone = serve.Server(nslots=1)
    keys = ("k", "p0k", "p2k", "p4k", "counts", "b123", "q123", "p0k1", "p0k2", "p0k3",
            "b_counts", "b123_sn")
    for r in mix:
        ref = serve.Request(r.rid, r.xyz, r.Lbox, r.Ngrid, step=r.step, Ncut=r.Ncut, Nmax=r.Nmax,
                            Nmubin=r.Nmubin, precision=r.precision)
        out, _ = alone.serve([ref])
        assert out[0].error is None
        for k in keys:
            assert np.array_equal(out[0].result[k], r.result[k], equal_nan=True), (r.rid, k)
    alone.close()
