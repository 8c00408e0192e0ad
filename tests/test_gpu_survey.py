"""GPU parity tests of the survey-geometry bispectrum (B0_survey,
FFT_survey_mono, _B0_survey; pyspectrum.py:13-282, :731-826) against the
oracle, through the C ABI."""
import numpy as np
import pytest
import torch

from oracle import spectrum as S
from pyspectrum_amd import _lib as L
from pyspectrum_amd import pyspectrum as P
from pyspectrum_amd import util

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda", 0)


def _say(*a):
    print("[case]", *a, flush=True)


def _catalogue(n, seed, zlo=0.05, zhi=0.12):
    rng = np.random.default_rng(seed)
    radecz = np.array([rng.uniform(0, 60, n), rng.uniform(-20, 20, n), rng.uniform(zlo, zhi, n)])
    nb = rng.uniform(1e-4, 5e-4, n)
    w = rng.uniform(0.8, 1.2, n)
    return radecz, nb, w


def _rel(a, b, floor):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_offset_assignment_matches_oracle(prec):
    _say("survey offset assignment", prec)
    N, Lbox = 32, 400.0
    rng = np.random.default_rng(21)
    xyz = rng.uniform(-0.5 * Lbox * 0.999, 0.5 * Lbox * 0.999, size=(3, 4000)).astype(np.float32)
    xyz[:, :4] = [[-199.5, 0.0, 199.5, -0.01]] * 3   # faces and the centre
    w = rng.uniform(0.5, 1.5, size=4000).astype(np.float32)
    kf_ks, off = np.float32(N / Lbox), np.float32(0.5 * N)
    A, B = S.assign_quad(xyz, w.astype(np.float64), N, kf_ks, offset=off)
    x = torch.from_numpy(np.ascontiguousarray(xyz)).to(DEV)
    wd = torch.from_numpy(w).to(DEV)
    g = torch.empty((N, N, N), dtype=torch.complex128 if prec == "fp64" else torch.complex64,
                    device=DEV)
    L.call("psx_assign_quad", L.context(0).handle, L.ptr(x), L.ptr(wd), x.shape[1], N,
           float(kf_ks), float(off), float(np.float32(-0.5 * Lbox)), float(np.float32(0.5 * Lbox)),
           float(np.abs(w).sum()), L.ptr(g), L.PREC[prec])
    out = g.cpu().numpy().transpose(2, 1, 0)
    tol = 1e-11 if prec == "fp64" else 5e-7
    scale = np.max(np.abs(A))
    assert np.max(np.abs(out.real - A)) / scale < tol and np.max(np.abs(out.imag - B)) / scale < tol


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_fft_survey_mono_matches_oracle(prec):
    _say("FFT_survey_mono", prec)
    N, Lbox = 32, 1000.0
    radecz, nb, w = _catalogue(2000, 22)
    got = P.FFT_survey_mono(radecz.copy(), nb, w=w, P0_fkp=1e4, Lbox=Lbox, Ngrid=N, precision=prec)
    ref = S.FFT_survey_mono(radecz, nb, w=w, P0_fkp=1e4, Lbox=Lbox, Ngrid=N)
    scale = np.max(np.abs(ref[0]))
    tol = 1e-10 if prec == "fp64" else 2e-6
    assert np.max(np.abs(got[0] - ref[0])) / scale < tol
    for a, b in zip(got[1:], ref[1:]):
        assert abs(a - b) <= 1e-12 * abs(b)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "bf16x3"])
def test_b0_survey_matches_oracle(prec):
    _say("B0_survey", prec)
    N, Lbox, Nmax, Ncut, step = 32, 1000.0, 5, 3, 3
    d, nb, w = _catalogue(3000, 23)
    r, nbr, _ = _catalogue(12000, 24)
    got = P.B0_survey(d.copy(), nb, w=w, radecz_r=r.copy(), nbar_r=nbr, P0_fkp=1e4, Lbox=Lbox,
                      Ngrid=N, step=step, Ncut=Ncut, Nmax=Nmax, precision=prec)
    counts = S.counts_Bk123(N, Nmax, Ncut, step)
    ref = S.B0_survey(d, nb, w, r, nbr, None, counts, P0_fkp=1e4, Lbox=Lbox, Ngrid=N, step=step,
                      Ncut=Ncut, Nmax=Nmax)
    for k in ("i_k1", "i_k2", "i_k3"):
        assert np.array_equal(got[k], ref[k]), k
    assert np.array_equal(got["counts"] > 0, ref["counts"] > 0)
    # delta = delta_d - alpha delta_r is a small difference of two large
    # fields, and p0k, b123 subtract shot-noise terms of their own size; the
    # float32 pipelines (like the reference's complex64 one) carry the float32
    # rounding of those large terms, so errors are measured against them
    _, _, Nd, _ = S.survey_catalogue(d, nb, w, 1e4, Lbox, util.default_cosmology())
    _, _, Nr, Ir = S.survey_catalogue(r, nbr, None, 1e4, Lbox, util.default_cosmology())
    alpha = Nd / Nr
    sn_p = (1. + alpha) * Ir[0] / Ir[2]
    floors = {"p0k1": sn_p, "b123": np.max(np.abs(ref["b123"])),
              "q123": np.max(np.abs(ref["q123"])), "counts": np.max(ref["counts"])}
    tol = {"fp64": 1e-9, "fp32": 5e-5, "bf16x3": 5e-5}[prec]
    for k in ("p0k1", "b123", "counts") + (("q123",) if prec == "fp64" else ()):
        e = _rel(got[k], ref[k], floors[k])
        assert e < tol, (prec, k, e)
    if prec != "fp64":
        # Q = B / (P1 P2 + P2 P3 + P3 P1): the denominator mixes signs here
        # (p0k crosses zero after shot noise), so Q is compared where it is
        # well conditioned and is otherwise checked to be exactly that ratio
        p1, p2, p3 = got["p0k1"], got["p0k2"], got["p0k3"]
        den = p1 * p2 + p2 * p3 + p3 * p1
        v = got["counts"] > 0
        assert np.allclose(got["q123"][v], got["b123"][v] / den[v], rtol=1e-14, atol=0)
        rden = ref["p0k1"] * ref["p0k2"] + ref["p0k2"] * ref["p0k3"] + ref["p0k3"] * ref["p0k1"]
        good = v & (np.abs(rden) > 0.1 * np.max(np.abs(rden)))
        assert good.sum() >= 5
        e = _rel(got["q123"][good], ref["q123"][good], floors["q123"])
        assert e < 1e-4, (prec, "q123", e)
    assert got["meta"]["Ngrid"] == N and got["meta"]["N"] == d.shape[1]


def test_b0_survey_identical_catalogues_closed_form():
    # randoms = data: alpha = 1, delta = 0 exactly; p0k = -2 I12/I22,
    # b123 = 6 I12 I23/(I22 I33) on every counted triangle
    _say("B0_survey delta = 0")
    N, Lbox, Nmax = 32, 1000.0, 5
    d, nb, w = _catalogue(3000, 25)
    got = P.B0_survey(d.copy(), nb, w=w, radecz_r=d.copy(), nbar_r=nb, w_r=w, P0_fkp=1e4, Lbox=Lbox,
                      Ngrid=N, Nmax=Nmax, precision="fp32")
    _, _, _, I = S.survey_catalogue(d, nb, w, 1e4, Lbox, util.default_cosmology())
    I12, I13, I22, I23, I33 = I
    valid = got["counts"] > 0
    assert np.allclose(got["p0k1"][valid], -2 * I12 / I22, rtol=1e-13)
    assert np.allclose(got["b123"][valid], 6 * I12 * I23 / (I22 * I33), rtol=1e-12)


def test_b0_survey_full_field_entry_matches():
    # _B0_survey on delta_d - alpha delta_r given as a full field (the
    # reference's own composition, pyspectrum.py:120-129) equals B0_survey
    _say("_B0_survey full field")
    N, Lbox, Nmax = 32, 1000.0, 5
    d, nb, w = _catalogue(3000, 26)
    r, nbr, _ = _catalogue(12000, 27)
    kw = dict(P0_fkp=1e4, Lbox=Lbox, Ngrid=N, precision="fp64")
    dd = P.FFT_survey_mono(d.copy(), nb, w=w, **kw)
    rr = P.FFT_survey_mono(r.copy(), nbr, **kw)
    alpha = dd[1] / rr[1]
    delta = P.reflect_delta(dd[0], Ngrid=N) - alpha * P.reflect_delta(rr[0], Ngrid=N)
    Is = [alpha * v for v in rr[2:]]
    a = P._B0_survey(delta, alpha, *Is, Nmax=Nmax, precision="fp64")
    b = P.B0_survey(d.copy(), nb, w=w, radecz_r=r.copy(), nbar_r=nbr, Nmax=Nmax, **kw)
    for k in ("p0k1", "b123", "q123", "counts"):
        floor = 1e-6 * np.max(np.abs(b[k]))
        assert _rel(a[k], b[k], floor) < 1e-10, k


def test_device_catalogue_matches_host():
    # psx_survey_catalogue (float64 on the device) against the host numpy
    # restatement of pyspectrum.py:767-812: float32 positions agree to one
    # float32 ulp (float64 cos/sin may differ by an ulp between libraries),
    # weights bit for bit, sums to float64 round-off of the summation order
    _say("device survey catalogue")
    d, nb, w = _catalogue(50000, 28, zlo=0.0, zhi=0.6)
    c = util.default_cosmology()
    xyz_h, w_h, Ntot_h, I_h = S.survey_catalogue(d, nb, w, 1e4, 4000.0, c)
    dev = torch.device("cuda", 0)
    rd = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
    nbd = torch.from_numpy(nb).to(dev)
    wd = torch.from_numpy(w).to(dev)
    zmax = float(d[2].max())
    Dn, sl = util.distance_table(c, zmax)
    xyz = torch.empty((3, d.shape[1]), dtype=torch.float32, device=dev)
    wf = torch.empty(d.shape[1], dtype=torch.float32, device=dev)
    sums = np.zeros(L.SURVEY_SUMS)
    L.call("psx_survey_catalogue", L.context(0).handle, L.ptr(rd), L.ptr(nbd), L.ptr(wd), d.shape[1],
           Dn.ctypes.data_as(L._p), sl.ctypes.data_as(L._p), len(Dn) - 1, zmax, c.h, 1e4, L.ptr(xyz),
           L.ptr(wf), sums.ctypes.data_as(L._p))
    xg = xyz.cpu().numpy()
    ulp = np.spacing(np.abs(xyz_h).astype(np.float32))
    assert np.all(np.abs(xg - xyz_h) <= ulp)
    assert np.mean(xg != xyz_h) < 1e-3
    assert np.array_equal(wf.cpu().numpy(), w_h.astype(np.float32))
    exp = [Ntot_h] + list(I_h)
    assert np.allclose(sums[:6], exp, rtol=1e-12, atol=0)
    assert np.allclose(sums[7:10], xyz_h.astype(np.float64).max(axis=1), rtol=1e-6)
    assert np.allclose(-sums[10:13], xyz_h.astype(np.float64).min(axis=1), rtol=1e-6)


def test_b0_survey_device_tensor_inputs():
    # catalogues given as CUDA tensors (and unit weights) give the same
    # result as the numpy inputs: counts and triangles bit for bit, values
    # that contain the P(k) shell powers (float64 atomic bin sums) to round-off
    _say("B0_survey device inputs")
    N, Lbox, Nmax = 32, 1000.0, 5
    d, nb, _ = _catalogue(3000, 29)
    r, nbr, _ = _catalogue(9000, 30)
    kw = dict(P0_fkp=1e4, Lbox=Lbox, Ngrid=N, Nmax=Nmax, precision="fp32")
    a = P.B0_survey(d.copy(), nb, radecz_r=r.copy(), nbar_r=nbr, **kw)
    dev = torch.device("cuda", 0)
    b = P.B0_survey(torch.from_numpy(d).to(dev), torch.from_numpy(nb).to(dev),
                    radecz_r=torch.from_numpy(r).to(dev), nbar_r=torch.from_numpy(nbr).to(dev), **kw)
    for k in ("counts", "i_k1", "i_k2", "i_k3"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("p0k1", "b123", "q123"):
        assert np.allclose(a[k], b[k], rtol=1e-11, atol=1e-11 * np.max(np.abs(a[k]))), k
    with pytest.raises(ValueError):
        P.B0_survey(d[:, :0], nb[:0], radecz_r=r, nbar_r=nbr, **kw)
