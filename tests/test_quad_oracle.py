"""CPU tests that pin the quadrupole oracle (oracle/spectrum.py, quadrupole
section): against the reference's own Fortran FiveDelta2g_1/2, build_quad and
the Q_ij / Q_ijkl weight branch of assign_quad (oracle/_ref), and against
closed forms that hold whatever the Fortran does."""
import numpy as np
import pytest

from oracle import ref as R
from oracle import spectrum as S

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (make -C oracle)")


def _field(N, seed):
    rng = np.random.default_rng(seed)
    s = (N // 2 + 1, N, N)
    return (rng.standard_normal(s) + 1j * rng.standard_normal(s)).astype(np.complex64)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30)


@needs_ref
@pytest.mark.parametrize("N", [16, 24])
def test_five_delta2g_matches_reference_fortran(N):
    xx, yy, zz, xy, yz, zx, g = (_field(N, 30 + k) for k in range(7))
    ours = S.five_delta2g_1(xx, yy, zz, N)
    theirs = R.five_delta2g_1(xx, yy, zz, N)
    assert np.array_equal(ours, theirs)
    ours2 = S.five_delta2g_2(g, xx, xy, yz, zx, N)
    theirs2 = R.five_delta2g_2(g, xx, xy, yz, zx, N)
    assert np.array_equal(ours2, theirs2)


@needs_ref
@pytest.mark.parametrize("irsd", [1, 2, 3])
def test_build_quad_matches_reference_fortran(irsd):
    N = 20
    d1, d2 = _field(N, 40), _field(N, 41)
    assert np.array_equal(S.build_quad(d1, d2, irsd, N), R.build_quad(d1, d2, irsd, N))


@needs_ref
@pytest.mark.parametrize("iabcd", [(1, 1, 0, 0), (1, 2, 0, 0), (3, 1, 0, 0), (1, 2, 3, 3)])
def test_quad_weights_match_reference_assignment(iabcd):
    # the Fortran weights inside assign_quad; the oracle assigns its weights
    rng = np.random.default_rng(42)
    N, Lbox = 16, 200.0
    xyz = rng.uniform(0.0, Lbox * 0.999, size=(3, 2000)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=2000).astype(np.float32)
    kf = np.float32(N / Lbox)
    wq = S.quad_weights(xyz, w, *iabcd)
    A, B = S.assign_quad(xyz, wq.astype(np.float64), N, kf)
    _, Ar, Br = R.assign_quad(xyz, w, N, kf, iabcd=iabcd)
    scale = np.max(np.abs(A))
    assert np.max(np.abs(A - Ar)) / scale < 1e-5
    assert np.max(np.abs(B - Br)) / scale < 1e-5


def test_build_quad_closed_form():
    # amu = +-1 on the line-of-sight axis: factor 5; perpendicular plane: -2.5;
    # k = 0 keeps dclr2
    N = 12
    d1, d2 = _field(N, 43), _field(N, 44)
    out = S.build_quad(d1, d2, 3, N)
    assert np.array_equal(out[0, 0, 1:], (5 * d1[0, 0, 1:]).astype(np.complex64))
    assert np.allclose(out[1:, :, 0], -2.5 * d1[1:, :, 0], rtol=1e-6)
    assert out[0, 0, 0] == d2[0, 0, 0]


def test_five_delta2g_closed_forms():
    N = 12
    F, G = _field(N, 45), _field(N, 46)
    # real unit vectors: kx^2 + ky^2 + kz^2 = 1, so equal inputs give 7.5 F
    out = S.five_delta2g_1(F, F, F, N, unit=True)
    assert np.allclose(out[1:], 7.5 * F[1:], rtol=2e-6)
    # the Fortran's integer unit vectors: 7.5 x the axis component, 0 off the axes
    out = S.five_delta2g_1(F, G, G, N)
    assert np.array_equal(out[3, 0, 0], np.complex64(7.5 * F[3, 0, 0]))
    assert np.array_equal(out[0, 2, 0], np.complex64(7.5 * G[0, 2, 0]))
    assert np.all(out[1, 1, :] == 0) and np.all(out[2, 3, 1:] == 0)
    assert out[0, 0, 0] == F[0, 0, 0]
    # no cross terms: dcgxx - 2.5 dcg (integer unit vectors never pair up)
    Z = np.zeros_like(F)
    out2 = S.five_delta2g_2(G, F, F, F, F, N)
    assert np.allclose(out2.ravel()[1:], (F - 2.5 * G).ravel()[1:], rtol=1e-6, atol=1e-6)
    assert np.allclose(S.five_delta2g_2(G, F, Z, Z, Z, N, unit=True).ravel()[1:],
                       (F - 2.5 * G).ravel()[1:], rtol=1e-6, atol=1e-6)


def test_quad_weights_trace_is_the_weight():
    rng = np.random.default_rng(47)
    xyz = rng.uniform(-100, 100, size=(3, 500)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=500).astype(np.float32)
    tr = sum(S.quad_weights(xyz, w, i, i, 0, 0).astype(np.float64) for i in (1, 2, 3))
    assert np.allclose(tr, w, rtol=1e-6)
    tr4 = sum(S.quad_weights(xyz, w, i, i, j, j).astype(np.float64) for i in (1, 2, 3) for j in (1, 2, 3))
    assert np.allclose(tr4, w, rtol=1e-5)
    assert np.array_equal(S.quad_weights(xyz, w, 0, 0, 0, 0), w)
