"""Seeded synthetic inputs generated on the GPU by the library's kernels:
lognormal mocks, Gaussian velocities and the redshift-space shift.

Lognormal recipe (the clustered-box workload): a Gaussian field on an M^3
mesh with P(k) proportional to k^slope exp(-(k/kcut)^2) (k in h/Mpc),
normalised to unit cell variance; exactly N points drawn with probability
proportional to exp(delta_G - 1/2) = 1 + delta_LN per cell (inverse-CDF
sampling of the cell) and placed uniformly inside it.  PyTorch only holds the
buffers; the field, the prefix sums, the draws and the shift run in
psx_mock.hip (Philox4x32-10 random numbers keyed by the seed).

``apply_rsd`` is applyRSD (pyspectrum/util.py:54-75) for float32 device
arrays: x_los <- (x_los + (f v_los + Lbox)) % Lbox with f = (1+z)/(100 E(z)),
the sum rounded to float32 before the wrap, as numpy does in the reference.
"""
import numpy as np
import torch

from . import _lib as L
from . import util


def _ctx(device):
    return L.context(device)


def gaussian_field(Lbox, seed, mesh=512, slope=-1.5, kcut=0.5, device=0):
    """float32 (mesh, mesh, mesh) unit-variance Gaussian field on ``cuda:device``
    and the standard deviation of the unnormalised draw."""
    g = torch.empty((mesh, mesh, mesh), dtype=torch.float32, device=torch.device("cuda", device))
    sig = L._d()
    _ctx(device).call("psx_mock_gaussian", mesh, float(Lbox), int(seed) & (2 ** 64 - 1),
                      float(slope), float(kcut), L.ptr(g), L.ctypes.byref(sig))
    return g, sig.value


def sample(g, n, Lbox, seed, sigma_v=0.0, redshift=0.0, omega0_m=0.3, los=2, device=0):
    """float32 (3, n) points drawn from the lognormal density of field ``g``;
    with sigma_v > 0, moved to redshift space along ``los`` in the same pass."""
    mesh = g.shape[0]
    xyz = torch.empty((3, n), dtype=torch.float32, device=g.device)
    fac = float(util.rsd_shift(redshift, omega0_m=omega0_m)) if sigma_v > 0 else 0.0
    _ctx(device).call("psx_mock_sample", L.ptr(g), mesh, float(Lbox), int(n),
                      int(seed) & (2 ** 64 - 1), float(sigma_v), fac, int(los), L.ptr(xyz))
    return xyz


def lognormal_box(n, Lbox, seed, mesh=512, slope=-1.5, kcut=0.5, device=0):
    """float32 (3, n) positions on ``cuda:device`` and the amplitude A of the
    power law (in units of a unit-variance white-noise draw normalised like an
    inverse real FFT) that gives the cell sigma_G = 1."""
    g, sig = gaussian_field(Lbox, seed, mesh, slope, kcut, device)
    xyz = sample(g, n, Lbox, seed, device=device)
    del g
    A = (float(mesh) ** 3 / sig) ** 2
    return xyz, A


def lognormal_rsd_box(n, Lbox, seed, sigma_v=300.0, redshift=0.5, omega0_m=0.3, los=2, mesh=512,
                      slope=-1.5, kcut=0.5, device=0):
    """Redshift-space lognormal box: the points of lognormal_box(seed) moved
    along ``los`` by Gaussian velocities (gaussian_velocities(seed)), the shift
    fused into the sampling pass."""
    g, _ = gaussian_field(Lbox, seed, mesh, slope, kcut, device)
    return sample(g, n, Lbox, seed, sigma_v, redshift, omega0_m, los, device)


def gaussian_velocities(n, sigma, seed, device=0):
    """float32 (3, n) velocities sigma * N(0, 1)."""
    v = torch.empty((3, n), dtype=torch.float32, device=torch.device("cuda", device))
    _ctx(device).call("psx_mock_velocities", int(n), int(seed) & (2 ** 64 - 1), float(sigma),
                      L.ptr(v))
    return v


def apply_rsd(xyz, vxyz, redshift, Lbox, omega0_m=0.3, los=2, h=0.7):
    """Redshift-space positions along axis ``los`` (util.py:54-75) for float32
    (3, N) device arrays."""
    assert xyz.dtype == torch.float32 and vxyz.dtype == torch.float32
    xyz, vxyz = xyz.contiguous(), vxyz.contiguous()
    out = torch.empty_like(xyz)
    fac = float(util.rsd_shift(redshift, h, omega0_m))
    _ctx(xyz.device.index or 0).call("psx_apply_rsd", L.ptr(xyz), L.ptr(vxyz), xyz.shape[1], fac,
                                     float(Lbox), int(los), L.ptr(out))
    return out


def uniform_box(n, Lbox, seed):
    """Uniform randoms as float32 (3, n) numpy (the CPU-runnable workload)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0, Lbox, size=(3, n)).astype(np.float32)
