"""GPU tests of the input-side kernels (psx_mock.hip) against their numpy
restatement (tests/mockref.py) and of psx_apply_rsd against util.applyRSD
(the reference's util.py:54-75 arithmetic), bit for bit where the kernels
and numpy run the same float64 operations."""
import numpy as np
import pytest
import torch

import mockref as M
from pyspectrum_amd import _lib as L
from pyspectrum_amd import mocks
from pyspectrum_amd import util

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("mesh,seed", [(32, 11), (48, 2 ** 40 + 3)])
def test_gaussian_field_matches_numpy(mesh, seed):
    _say("mock gaussian field", mesh, seed)
    Lbox = 500.0
    g, sig = mocks.gaussian_field(Lbox, seed, mesh=mesh)
    ref, sig_ref = M.gaussian_field(mesh, Lbox, seed)
    got = g.cpu().numpy().astype(np.float64)
    # float32 FFTs on the device against float64 numpy: unit-variance field
    assert np.max(np.abs(got - ref)) < 2e-5
    assert abs(sig / sig_ref - 1) < 1e-5
    assert abs(got.mean()) < 1e-6 and abs(got.std(ddof=1) - 1) < 1e-6


def test_sampler_matches_numpy_and_is_deterministic():
    _say("mock sampler")
    mesh, Lbox, seed, n = 32, 500.0, 5, 20000
    g, _ = mocks.gaussian_field(Lbox, seed, mesh=mesh)
    a = mocks.sample(g, n, Lbox, seed)
    b = mocks.sample(g, n, Lbox, seed)
    want = M.sample(g.cpu().numpy(), n, Lbox, seed)
    assert np.array_equal(a.cpu().numpy(), want)
    assert torch.equal(a, b)
    assert float(a.min()) >= 0.0 and float(a.max()) < Lbox
    assert not torch.equal(a, mocks.sample(g, n, Lbox, seed + 1))


def test_fused_rsd_equals_velocities_then_apply_rsd():
    """The shift fused into the sampling pass equals sampling, drawing the
    velocities and applying psx_apply_rsd, which equals util.applyRSD."""
    _say("mock fused rsd")
    mesh, Lbox, seed, n, z, sv = 32, 1000.0, 9, 30000, 0.5, 300.0
    g, _ = mocks.gaussian_field(Lbox, seed, mesh=mesh)
    real = mocks.sample(g, n, Lbox, seed)
    v = mocks.gaussian_velocities(n, sv, seed)
    vref = M.velocities(n, seed, sv)
    ulp = np.spacing(np.abs(vref).astype(np.float32))
    assert np.all(np.abs(v.cpu().numpy() - vref) <= ulp)
    for los in range(3):
        fused = mocks.sample(g, n, Lbox, seed, sigma_v=sv, redshift=z, los=los)
        sep = mocks.apply_rsd(real, v, z, Lbox, los=los)
        assert torch.equal(fused, sep), los
        host = util.applyRSD(real.cpu().numpy(), v.cpu().numpy(), z, h=0.7, omega0_m=0.3,
                             LOS="xyz"[los], Lbox=Lbox)
        assert np.array_equal(sep.cpu().numpy(), host), los


@pytest.mark.parametrize("los", ["x", "y", "z"])
def test_apply_rsd_matches_util_bitwise(los):
    _say("apply_rsd", los)
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1000, (3, 50000)).astype(np.float32)
    x[:, :4] = np.array([0.0, 999.99994, 500.0, 1000 * (1 - 1e-6)], dtype=np.float32)
    v = rng.normal(0, 300, (3, 50000)).astype(np.float32)
    v[:, 0] = -3e5   # a shift longer than the box: negative before the wrap
    v[:, 1] = 3e5
    want = util.applyRSD(x, v, 0.5, h=0.7, omega0_m=0.3, LOS=los, Lbox=1000.0)
    got = mocks.apply_rsd(torch.from_numpy(x).to(DEV), torch.from_numpy(v).to(DEV), 0.5, 1000.0,
                          los="xyz".index(los))
    assert np.array_equal(got.cpu().numpy(), want)


def test_lognormal_box_full_size():
    """The clustered-box input (1e6 points, 512^3 mesh, Lbox 2500, seed 2):
    every point inside the box, the same bits on a second call, and the
    counts in 64^3 cells consistent with the field's density: with
    lambda_C = n W_C / W (W_C the summed weights exp(g - 1/2) of the 8^3 mesh
    cells in C), sum (N_C - lambda_C)^2 / lambda_C is within 7 standard
    deviations of its expectation, the number of cells."""
    _say("lognormal box 1e6 on 512^3")
    n, Lbox = 1_000_000, 2500.0
    a, A = mocks.lognormal_box(n, Lbox, seed=2)
    b, _ = mocks.lognormal_box(n, Lbox, seed=2)
    assert a.shape == (3, n) and a.dtype == torch.float32
    assert torch.equal(a, b) and A > 0
    assert float(a.min()) >= 0.0 and float(a.max()) < Lbox
    g, _ = mocks.gaussian_field(Lbox, 2, mesh=512)
    w = torch.exp(g.double() - 0.5).reshape(64, 8, 64, 8, 64, 8).sum(dim=(1, 3, 5))
    lam = (n * w / w.sum()).reshape(-1)
    del g, w
    cell = torch.clamp((a.double() / (Lbox / 64)).long(), 0, 63)
    counts = torch.bincount((cell[0] * 64 + cell[1]) * 64 + cell[2], minlength=64 ** 3).double()
    chi2 = float(((counts - lam) ** 2 / lam).sum())
    ncell = 64 ** 3
    _say("  chi2 / ncell = %.4f, var(lambda)/mean = %.3f" % (chi2 / ncell, float(lam.var() / lam.mean())))
    assert abs(chi2 - ncell) < 7 * np.sqrt(2 * ncell)
    assert float(lam.var()) > 0.01 * float(lam.mean()) ** 2   # the density is clustered
