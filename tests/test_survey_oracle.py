"""CPU tests that pin the survey-geometry oracle and host logic
(oracle/spectrum.py survey section, pyspectrum_amd/util.py).

Checks: the comoving distance against closed forms and adaptive quadrature,
the cartesian conversion on axis points, the offset assignment and
fcomb_survey against the reference's own Fortran (oracle/_ref), and the
survey normalisation against its closed form when data and randoms coincide
(delta = 0)."""
import math

import numpy as np
import pytest

from oracle import ref as R
from oracle import spectrum as S
from pyspectrum_amd import util

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (make -C oracle)")


def test_comoving_distance_closed_forms():
    z = np.linspace(0.0, 4.0, 401)
    # Einstein-de Sitter: chi = 2 (1 - 1/sqrt(1+z)); de Sitter: chi = z
    for Om0, chi in ((1.0, 2.0 * (1.0 - 1.0 / np.sqrt(1.0 + z))), (0.0, z)):
        ref = util.C_KMS / 70.0 * chi
        for d in (util.FlatLambdaCDM(70.0, Om0).comoving_distance(z),
                  S.comoving_distance(z, 70.0, Om0)):
            assert np.max(np.abs(d - ref) / np.maximum(ref, 1e-300)) < 1e-11


def test_comoving_distance_host_matches_quadrature():
    z = np.array([1e-3, 0.05, 0.2, 0.43, 0.57, 0.7, 1.3, 2.5])
    host = util.FlatLambdaCDM(67.6, 0.31).comoving_distance(z)
    quad = S.comoving_distance(z, 67.6, 0.31)
    assert np.max(np.abs(host - quad) / quad) < 1e-11


def test_radecz_axis_points():
    c = util.default_cosmology()
    z = np.array([0.1, 0.1, 0.1, 0.3])
    radecz = np.array([[0.0, 90.0, 0.0, 180.0], [0.0, 0.0, 90.0, -45.0], z])
    r = c.comoving_distance(z) * c.h
    xyz = util.radecz_to_cartesian(radecz.copy(), c)
    exp = np.array([[r[0], 0, 0, -r[3] / math.sqrt(2)], [0, r[1], 0, 0],
                    [0, 0, r[2], -r[3] / math.sqrt(2)]])
    assert np.allclose(xyz, exp, rtol=1e-12, atol=1e-9 * r.max())
    # the oracle's conversion (quadrature distances) agrees
    assert np.allclose(S.radecz_to_cartesian(radecz, 67.6, 0.31), xyz, rtol=1e-11, atol=1e-8)
    # the caller's array is left in degrees
    radecz2 = radecz.copy()
    util.radecz_to_cartesian(radecz2, c)
    assert np.array_equal(radecz2, radecz)


@needs_ref
def test_offset_assignment_matches_reference_fortran():
    # survey coordinates: centred on the observer, offset Ngrid/2 (pyspectrum.py:819)
    rng = np.random.default_rng(14)
    N, Lbox = 24, 300.0
    xyz32 = rng.uniform(-0.5 * Lbox * 0.999, 0.5 * Lbox * 0.999, size=(3, 3000)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, size=3000).astype(np.float32)
    off = np.float32(0.5 * N)
    A, B = S.assign_quad(xyz32, w.astype(np.float64), N, np.float32(N / Lbox), offset=off)
    _, Ar, Br = R.assign_quad(xyz32, w, N, np.float32(N / Lbox), offset=off)
    scale = np.max(np.abs(A))
    assert np.max(np.abs(A - Ar)) / scale < 1e-6
    assert np.max(np.abs(B - Br)) / scale < 1e-6
    # the offset is a shift by half the box: same grid as the periodic
    # assignment of x + Lbox/2 wherever that shift is exact in float32
    shifted = S.assign_quad((xyz32.astype(np.float64) + 0.5 * Lbox).astype(np.float32),
                            w.astype(np.float64), N, np.float32(N / Lbox))
    assert abs(np.sum(A) - np.sum(shifted[0])) < 1e-9 * np.sum(A)


@needs_ref
def test_fcomb_survey_matches_reference_fortran():
    rng = np.random.default_rng(16)
    for N in (16, 24):
        F = (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))).astype(np.complex64)
        ours = S.fcomb_survey(F.astype(np.complex128), N)
        theirs = R.fcomb_survey(F, N)
        scale = np.max(np.abs(ours))
        assert np.max(np.abs(ours - theirs)) / scale < 1e-6, N
        # and it is fcomb_periodic with one object (cf = 1/(6^3*4))
        assert np.max(np.abs(R.fcomb_periodic(F, 1.0, N) - theirs)) == 0.0


def test_survey_normalisation_closed_form_when_delta_vanishes():
    # data = randoms: alpha = 1, delta = 0, so p0k = -2 I12/I22 for every shell,
    # b = 6 I12 I23 / (I22 I33) and q = I22 I23 / (2 I12 I33) on every triangle
    Ngrid, Nmax, Ncut, step = 24, 6, 3, 3
    rng = np.random.default_rng(17)
    I12, I13, I22, I23, I33 = rng.uniform(0.5, 2.0, size=5)
    counts = np.zeros((Nmax, Nmax, Nmax))
    I, J, Lx, F = S.triangles(Nmax, Ncut, step)
    for i, j, l in zip(I, J, Lx):
        counts[i - 1, j - 1, l - 1] = 5.0 * Ngrid ** 3
    out = S.B0_survey_from_sums(lambda i, j, l: 0.0, np.zeros(Nmax), counts, 1.0, I12, I13, I22,
                                I23, I33, Nmax, Ncut, step, Ngrid)
    p = -2.0 * I12 / I22
    assert np.allclose(out["p0k1"], p, rtol=1e-14)
    assert np.allclose(out["b123"], 6.0 * I12 * I23 / (I22 * I33), rtol=1e-13)
    assert np.allclose(out["q123"], I22 * I23 / (2.0 * I12 * I33), rtol=1e-13)
    assert np.allclose(out["counts"], 5.0 / F, rtol=1e-14)
    assert len(out["i_k1"]) == len(I) and np.array_equal(out["i_k1"], I * step)


def test_survey_catalogue_sums_match_oracle():
    # host-side FKP weights and normalisation sums (pyspectrum.py:796-805)
    from pyspectrum_amd import pyspectrum as P
    rng = np.random.default_rng(18)
    n = 500
    radecz = np.array([rng.uniform(0, 60, n), rng.uniform(-20, 20, n), rng.uniform(0.05, 0.12, n)])
    nb = rng.uniform(1e-4, 5e-4, n)
    w = rng.uniform(0.8, 1.2, n)
    xyzs, wf, Ntot, I = S.survey_catalogue(radecz, nb, w, 1e4, 1000.0, util.default_cosmology())
    ref_w = w * (1. / (1. + nb * 1e4))
    assert Ntot == np.sum(w)
    assert np.array_equal(wf, ref_w)
    exp = (np.sum(ref_w ** 2), np.sum(ref_w ** 3), np.sum(nb * ref_w ** 2), np.sum(nb * ref_w ** 3),
           np.sum(nb ** 2 * ref_w ** 3))
    assert tuple(I) == exp
    assert xyzs.dtype == np.float32
    with pytest.raises(AssertionError):
        S.survey_catalogue(radecz, nb, w, 1e4, 200.0, util.default_cosmology())


def test_ijl_order_sorts_by_l_then_j_then_i():
    I, J, Lx, _ = S.triangles(8, 3, 3)
    order = util.ijl_order(I, J, Lx).ravel()
    keys = list(zip(Lx[order], J[order], I[order]))
    assert keys == sorted(keys) and len(set(order)) == len(I)


def test_driver_text_tables_round_trip(tmp_path):
    # run/glam/glam.py:104-106, :127-133: same columns, formats and headers
    rng = np.random.default_rng(19)
    I, J, Lx, _ = S.triangles(6, 3, 3)
    bk = {"i_k1": I * 3, "i_k2": J * 3, "i_k3": Lx * 3}
    for k in ("p0k1", "p0k2", "p0k3", "b123", "q123", "counts", "b123_sn"):
        bk[k] = rng.standard_normal(len(I)) * 1e4
    path = tmp_path / "b123.dat"
    util.write_bk_text(str(path), bk, 2500.0)
    lines = path.read_text().splitlines()
    assert lines[0] == "# pyspectrum bispectrum calculation test. k_f = 2pi/2500.0"
    first = lines[1].split()
    assert first[:3] == [str(I[0] * 3), str(J[0] * 3), str(Lx[0] * 3)]
    assert first[3] == "%.5e" % bk["p0k1"][0]
    back = util.read_bk_text(str(path))
    assert np.array_equal(back["i_k1"], bk["i_k1"])
    assert np.allclose(back["b123"], bk["b123"], rtol=1e-5)
    spec = {"k": np.linspace(0.01, 0.5, 20), "p0k": rng.uniform(1e3, 1e4, 20),
            "counts": np.arange(20) * 10.0}
    util.write_pk_text(str(tmp_path / "pk.dat"), spec, 2500.0, 1e-4)
    hdr = (tmp_path / "pk.dat").read_text().splitlines()[0]
    assert hdr == "# pyspectrum P_l=0(k) calculation. Lbox=2500.0, k_f=%.5e, SN=1.00000e+04" % (2 * np.pi / 2500)
    k, p, c = util.read_pk_text(str(tmp_path / "pk.dat"))
    assert np.allclose(k, spec["k"], rtol=1e-5) and np.array_equal(c, spec["counts"])
    bk["i_k1"] = bk["i_k1"][:-1]
    with pytest.raises(ValueError):
        util.write_bk_text(str(tmp_path / "bad.dat"), bk, 2500.0)


def test_ijl_order_matches_reference_grouping():
    # bit-level: the lexsort form against the oracle's restatement of the
    # reference's grouping loop, on a shuffled triangle list and on a list in
    # which every triangle occurs twice
    I, J, Lx, _ = S.triangles(12, 3, 3)
    rng = np.random.default_rng(5)
    p = rng.permutation(len(I))
    for a, b, c in ((I[p], J[p], Lx[p]), (np.r_[I, I][::-1], np.r_[J, J][::-1], np.r_[Lx, Lx][::-1])):
        got = util.ijl_order(a, b, c)
        want = S.ijl_order(a, b, c)
        assert got.shape == want.shape and np.array_equal(got, want)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("los", ["x", "y", "z"])
def test_apply_rsd_matches_reference_arithmetic(dtype, los):
    # util.applyRSD bit for bit against the oracle's restatement of
    # util.py:54-75 (float64 shift factor, x + (f v + Lbox), then % Lbox)
    rng = np.random.default_rng(6)
    x = rng.uniform(0, 1000, (3, 20000)).astype(dtype)
    x[:, :3] = [[0.0], [999.999], [500.0]]
    v = rng.normal(0, 300, (3, 20000)).astype(dtype)
    v[:, 0] = -3e5  # a shift longer than the box: negative before the wrap
    got = util.applyRSD(x, v, 0.5, h=0.7, omega0_m=0.3, LOS=los, Lbox=1000.0)
    want = S.apply_rsd(x, v, 0.5, 0.7, 0.3, los, 1000.0)
    assert got.dtype == x.dtype and np.array_equal(got, want)
    other = [a for a in range(3) if a != "xyz".index(los)]
    assert np.array_equal(got[other], x[other])
    with pytest.raises(ValueError):
        util.applyRSD(x, v, 0.5, LOS=los)
