"""GPU parity of the quadrupole machinery (pyspectrum_amd/quad.py through
the C ABI) against the oracle (oracle/spectrum.py quadrupole section, pinned
bit for bit to the reference Fortran): FiveDelta2g_1/2 and build_quad bit for
bit on float32 half fields, the Q_ij / Q_ijkl particle weights bit for bit,
and a Q_ij density field through FFT_periodic."""
import numpy as np
import pytest
import torch

import pyspectrum_amd._lib as L
from oracle import spectrum as S
from pyspectrum_amd import pyspectrum as P
from pyspectrum_amd import quad as Q

pytestmark = pytest.mark.gpu


def _say(*a):
    print("[case]", *a, flush=True)


def _field(N, seed, dtype=np.complex64):
    rng = np.random.default_rng(seed)
    s = (N // 2 + 1, N, N)
    return (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(dtype)


@pytest.mark.parametrize("N", [24, 96])
@pytest.mark.parametrize("unit", [False, True])
def test_five_delta2g_match_oracle(N, unit):
    _say("five_delta2g", N, unit)
    xx, yy, zz, xy, yz, zx, g = (_field(N, 60 + k) for k in range(7))
    exp1 = S.five_delta2g_1(xx, yy, zz, N, unit=unit)
    got1 = xx.copy()
    Q.FiveDelta2g_1(got1, yy, zz, N, unit_vectors=unit)
    assert np.array_equal(got1, exp1)
    exp2 = S.five_delta2g_2(g, xx, xy, yz, zx, N, unit=unit)
    got2 = xx.copy()
    Q.FiveDelta2g_2(g, got2, xy, yz, zx, N, unit_vectors=unit)
    assert np.array_equal(got2, exp2)


@pytest.mark.parametrize("irsd", [1, 2, 3])
def test_build_quad_matches_oracle(irsd):
    _say("build_quad", irsd)
    N = 48
    d1, d2 = _field(N, 70), _field(N, 71)
    exp = S.build_quad(d1, d2, irsd, N)
    got = d2.copy()
    Q.build_quad(d1, got, irsd, N)
    assert np.array_equal(got, exp)


def test_quad_fp64_and_device_tensors():
    # complex128 runs the float64 path (within float32 rounding of the oracle);
    # CUDA tensors in the [iz, iy, ix] layout are updated in place
    _say("quad fp64 / device")
    N = 32
    xx, yy, zz = (_field(N, 80 + k, np.complex128) for k in range(3))
    exp = S.five_delta2g_1(xx, yy, zz, N, unit=True)
    got = xx.copy()
    Q.FiveDelta2g_1(got, yy, zz, N, unit_vectors=True)
    assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) < 1e-6
    d1, d2 = _field(N, 84), _field(N, 85)
    t1 = torch.from_numpy(np.ascontiguousarray(d1.transpose(2, 1, 0))).cuda()
    t2 = torch.from_numpy(np.ascontiguousarray(d2.transpose(2, 1, 0))).cuda()
    Q.build_quad(t1, t2, 3, N)
    assert np.array_equal(t2.cpu().numpy().transpose(2, 1, 0), S.build_quad(d1, d2, 3, N))


@pytest.mark.parametrize("iabcd", [(0, 0, 0, 0), (1, 1, 0, 0), (2, 3, 0, 0), (1, 2, 3, 1)])
def test_quad_weights_match_oracle(iabcd):
    _say("quad_weights", iabcd)
    rng = np.random.default_rng(90)
    xyz = rng.uniform(0, 500, size=(3, 30001)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=30001).astype(np.float32)
    assert np.array_equal(Q.quad_weights(xyz, w, *iabcd), S.quad_weights(xyz, w, *iabcd))
    assert np.array_equal(Q.quad_weights(xyz, None, *iabcd), S.quad_weights(xyz, None, *iabcd))


def test_qij_field_through_fft_periodic():
    # a Q_xy density field: the library's weights into FFT_periodic equal the
    # oracle's FFT_periodic of the oracle's weights
    _say("Q_xy field")
    rng = np.random.default_rng(91)
    N, Lbox = 48, 600.0
    xyz = rng.uniform(0, Lbox, size=(3, 20000)).astype(np.float32)
    wq = Q.quad_weights(xyz, None, 1, 2, 0, 0)
    got = P.FFT_periodic(xyz, w=wq, Lbox=Lbox, Ngrid=N)
    exp = S.FFT_periodic(xyz.astype(np.float64), w=S.quad_weights(xyz, None, 1, 2, 0, 0).astype(np.float64),
                         Lbox=Lbox, Ngrid=N)
    assert np.max(np.abs(got - exp)) / np.max(np.abs(exp)) < 1e-5


def test_quad_argument_errors():
    _say("quad errors")
    N = 16
    d = _field(N, 92)
    with pytest.raises(ValueError):
        Q.build_quad(d, d.copy(), 4, N)
    with pytest.raises(L.PsxError):
        Q.quad_weights(np.zeros((3, 10), np.float32), None, 4, 1, 0, 0)
    with pytest.raises(ValueError):
        Q.FiveDelta2g_1(d, d.astype(np.complex128), d, N)
