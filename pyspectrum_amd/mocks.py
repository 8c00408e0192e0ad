"""Seeded synthetic inputs generated on the GPU: lognormal mocks and the RSD shift.

Lognormal recipe (the clustered-box workload): a Gaussian field on an M^3
mesh with P(k) = A k^-1.5 exp(-(k/0.5)^2) (k in h/Mpc), A chosen so the cell
sigma_G = 1; delta_LN = exp(delta_G - sigma^2/2) - 1; exactly N points drawn
with probability proportional to 1 + delta_LN per cell (inverse-CDF sampling
of the cell, i.e. a multinomial draw) and placed uniformly inside the cell.

``apply_rsd`` restates applyRSD (pyspectrum/util.py:54-75) for a flat LCDM
E(z) = sqrt(Om (1+z)^3 + 1 - Om): x_los <- (x_los + v_los (1+z)/(100 E(z)) + L) mod L.
"""
import math

import numpy as np
import torch


def lognormal_box(n, Lbox, seed, mesh=512, slope=-1.5, kcut=0.5, device=0):
    """float32 (3, n) positions on ``cuda:device`` (and the amplitude A used)."""
    dev = torch.device("cuda", device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    M = mesh
    white = torch.randn((M, M, M), generator=g, device=dev, dtype=torch.float32)
    wk = torch.fft.rfftn(white)
    del white
    kf = 2 * math.pi / Lbox
    fx = torch.fft.fftfreq(M, d=1.0 / M, device=dev) * kf
    fz = torch.fft.rfftfreq(M, d=1.0 / M, device=dev) * kf
    k2 = fx[:, None, None] ** 2 + fx[None, :, None] ** 2 + fz[None, None, :] ** 2
    k = torch.sqrt(k2)
    amp = torch.where(k > 0, k.clamp_min(1e-30) ** (slope / 2) * torch.exp(-0.5 * (k / kcut) ** 2),
                      torch.zeros_like(k))
    wk *= amp
    del k, k2, amp
    dg = torch.fft.irfftn(wk, s=(M, M, M))
    del wk
    sig = float(dg.double().std().item())
    dg /= sig
    A = 1.0 / sig ** 2  # amplitude (in the units of the unnormalised draw) giving sigma_G = 1
    w = torch.exp(dg.double() - 0.5).reshape(-1)
    del dg
    cdf = torch.cumsum(w, 0)
    del w
    u = torch.rand(n, generator=g, device=dev, dtype=torch.float64) * cdf[-1]
    cell = torch.searchsorted(cdf, u).clamp_max(M ** 3 - 1)
    del cdf, u
    ix = cell // (M * M)
    iy = (cell // M) % M
    iz = cell % M
    off = torch.rand((3, n), generator=g, device=dev, dtype=torch.float64)
    h = Lbox / M
    xyz = torch.stack([(ix + off[0]) * h, (iy + off[1]) * h, (iz + off[2]) * h])
    xyz = torch.remainder(xyz, Lbox).to(torch.float32).contiguous()
    return xyz, A


def gaussian_velocities(n, sigma, seed, device=0):
    dev = torch.device("cuda", device)
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed) + 7919)
    return torch.randn((3, n), generator=g, device=dev, dtype=torch.float32) * sigma


def apply_rsd(xyz, vxyz, redshift, Lbox, omega0_m=0.3, los=2):
    """Redshift-space positions along axis ``los`` (util.py:54-75)."""
    ez = math.sqrt(omega0_m * (1 + redshift) ** 3 + 1 - omega0_m)
    fac = (1 + redshift) / (100.0 * ez)
    out = xyz.clone()
    out[los] = torch.remainder(xyz[los].double() + fac * vxyz[los].double() + Lbox, Lbox).to(xyz.dtype)
    return out


def uniform_box(n, Lbox, seed):
    """Uniform randoms as float32 (3, n) numpy (the CPU-runnable workload)."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0, Lbox, size=(3, n)).astype(np.float32)
