"""GPU test of psx_bin_power_noalloc, the entry point that allocates no
device memory (its block partials live in a fixed device array): two calls
in a row -- a 96^3 field, then a smaller 48^3 one -- each equal bit for bit
to psx_bin_power on the same input, no workspace in the calling context, and
a clear error beyond its limit of 512 bins."""
import numpy as np
import pytest
import torch

from pyspectrum_amd import _lib as L

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda", 0)


def _say(*a):
    print("[case]", *a, flush=True)


def _half(ctx, N, seed, prec):
    """Half field of a random real field's transform (complex [N][N][N/2+1])."""
    rng = np.random.default_rng(seed)
    f = torch.from_numpy(np.fft.fftn(rng.standard_normal((N, N, N)))).to(DEV)
    cd = torch.complex128 if prec == "fp64" else torch.complex64
    half = torch.empty((N, N, N // 2 + 1), dtype=cd, device=DEV)
    sx, sy, sz = f.stride()
    ctx.call("psx_hermitian_half", L.ptr(f), L.PREC["fp64"], sx, sy, sz, L.ptr(half), N,
             L.PREC[prec])
    return half


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_noalloc_bin_power_matches_bin_power(prec):
    _say("bin_power_noalloc", prec)
    cases = ((96, 1, 48, 0.01), (48, 2, 12, 0.03))  # (Ngrid, step, nbins, kf): second smaller
    ref_ctx = L.Context(0)
    halves, refs = [], []
    for k, (N, step, nbins, kf) in enumerate(cases):
        halves.append(_half(ref_ctx, N, 50 + k, prec))
        ref = torch.empty(3 * nbins, dtype=torch.float64, device=DEV)
        ref_ctx.call("psx_bin_power", L.ptr(halves[-1]), N, step, nbins, kf, L.ptr(ref), L.PREC[prec])
        refs.append(ref)
    # a context of its own that only ever runs the no-allocation entry point
    ctx = L.Context(0)
    outs = []
    for (N, step, nbins, kf), half in zip(cases, halves):
        out = torch.full((3 * nbins,), np.nan, dtype=torch.float64, device=DEV)
        ctx.call("psx_bin_power_noalloc", L.ptr(half), N, step, nbins, kf, L.ptr(out), L.PREC[prec])
        outs.append(out)
    torch.cuda.synchronize()
    for (N, _, nbins, _), out, ref in zip(cases, outs, refs):
        assert np.array_equal(out.cpu().numpy(), ref.cpu().numpy()), (N, nbins)
    assert ctx.workspace_bytes() == 0
    with pytest.raises(L.PsxError, match="512"):
        big = torch.empty(3 * 600, dtype=torch.float64, device=DEV)
        ctx.call("psx_bin_power_noalloc", L.ptr(halves[0]), 96, 1, 600, 0.01, L.ptr(big),
                 L.PREC[prec])
    ctx.close()
    ref_ctx.close()
