"""CPU tests that pin the oracle (oracle/spectrum.py).

Each oracle function is checked against something independent of its own
formula: the reference's FP64 N_tri caches (committed fixtures), the
reference's own Fortran kernels compiled from its source (oracle/_ref),
closed forms, and brute force on tiny grids.
"""
import math
import os

import numpy as np
import pytest

from oracle import ref as R
from oracle import spectrum as S

needs_ref = pytest.mark.skipif(not R.available(), reason="oracle/_ref not built (make -C oracle)")


def _load_ntri(golden_dir, nmax):
    return np.load(os.path.join(golden_dir, "ntri_Ngrid360_Ncut3_step3_Nmax%d.npz" % nmax))


# ------------------------------------------------------------ shell bins ----
def test_shell_index_matches_float_formula():
    # pyspectrum.py:370-378 evaluated literally in float64
    for N in (24, 96, 360):
        for step in (1, 2, 3, 5):
            m = S.kmag2(N)
            ref = (np.sqrt(m.astype(np.float64)) / step + 0.5).astype(int)
            assert np.array_equal(S.shell_index(N, step), ref), (N, step)


def test_shell_sizes_step3():
    # Nk[1..3] = 370, 1402, 3154 at step 3 for any Ngrid >= 24
    for N in (24, 96, 360):
        assert list(S.shell_counts(N, 3, 3)[1:4]) == [370, 1402, 3154]


def test_pk_bins_match_float_formula():
    # pyspectrum.py:691-703 with kf = 2 pi / Lbox
    N, Lbox = 96, 1000.0
    kf = 2 * np.pi / Lbox
    rk = kf * np.sqrt(S.kmag2(N).astype(np.float64))
    ref = ((N // 2) * rk / (kf * N / 2.0) + 0.5).astype(int)
    assert np.array_equal(S.pk_bins(N), ref)


def test_triangle_enumeration_matches_fixtures(golden_dir):
    for nmax, n in ((10, 150), (40, 6350), (50, 12000)):
        I, J, L, F = S.triangles(nmax, 3, 3)
        assert len(I) == n
        g = _load_ntri(golden_dir, nmax)
        nz = g["ntri"] != 0
        tri = set(zip(I, J, L))
        assert tri == set(zip(g["i"][nz], g["j"][nz], g["l"][nz]))
        # fac of pyspectrum.py:418-422
        assert np.all(F[(I == J) & (J == L)] == 6)
        assert np.all(F[(I == J) ^ (J == L)] == 2)


# -------------------------------------------------------------- N_tri ------
def test_ntri_fixture_is_integer(golden_dir):
    for nmax in (10, 40, 50):
        g = _load_ntri(golden_dir, nmax)
        G = 360.0 ** 3
        assert np.max(np.abs(g["counts"] / G - g["ntri"])) < 5e-4


def test_counts_oracle_matches_fixture_small_grid(golden_dir):
    # no aliasing for shells up to 10 (|k| < 31.5) on a 96^3 grid: N_tri
    # must equal the reference's 360^3 cache exactly
    counts = S.counts_Bk123(96, 10, 3, 3)
    I, J, L, _ = S.triangles(10, 3, 3)
    ntri = np.rint(counts[I - 1, J - 1, L - 1] / 96.0 ** 3).astype(np.int64)
    g = _load_ntri(golden_dir, 10)
    ref = {(a, b, c): n for a, b, c, n in zip(g["i"], g["j"], g["l"], g["ntri"])}
    assert [ref[(a, b, c)] for a, b, c in zip(I, J, L)] == list(ntri)


def test_counts_oracle_matches_bruteforce():
    N, nmax, ncut, step = 16, 5, 1, 1
    counts = S.counts_Bk123(N, nmax, ncut, step)
    brute = S.ntri_bruteforce(N, nmax, ncut, step)
    I, J, L, _ = S.triangles(nmax, ncut, step)
    got = np.rint(counts[I - 1, J - 1, L - 1] / N ** 3).astype(np.int64)
    assert np.array_equal(got, brute[I - 1, J - 1, L - 1])
    assert np.all(np.abs(counts[I - 1, J - 1, L - 1] / N ** 3 - got) < 1e-9)


# ---------------------------------------------------------- assignment ----
def test_assign_invariants():
    rng = np.random.default_rng(3)
    N, Lbox = 16, 100.0
    xyz = rng.uniform(0, Lbox, size=(3, 500))
    w = rng.uniform(0.5, 2.0, size=500)
    A, B = S.assign_quad(S.clip_positions(xyz, Lbox), w, N, np.float32(N / Lbox))
    # the x6 PCS weights sum to 6 per axis
    assert A.sum() == pytest.approx(216 * w.sum(), rel=1e-12)
    assert B.sum() == pytest.approx(216 * w.sum(), rel=1e-12)


def test_assign_node_stencil():
    # a particle exactly on node n of grid A: weights 1, 4, 1 per axis
    N, Lbox = 8, 8.0
    xyz = np.array([[3.0], [5.0], [0.0]])
    A, _ = S.assign_quad(S.clip_positions(xyz, Lbox), np.ones(1), N, np.float32(1.0))
    w1 = np.zeros(N); w1[[2, 3, 4]] = [1, 4, 1]
    w2 = np.zeros(N); w2[[4, 5, 6]] = [1, 4, 1]
    w3 = np.zeros(N); w3[[7, 0, 1]] = [1, 4, 1]
    assert np.array_equal(A, np.einsum("i,j,k->ijk", w1, w2, w3))


@needs_ref
def test_assign_matches_reference_fortran():
    rng = np.random.default_rng(4)
    N, Lbox = 24, 300.0
    xyz = rng.uniform(0, Lbox, size=(3, 3000))
    w = rng.uniform(0.5, 1.5, size=3000).astype(np.float32)
    xyz32 = S.clip_positions(xyz, Lbox)
    A, B = S.assign_quad(xyz32, w.astype(np.float64), N, np.float32(N / Lbox))
    _, Ar, Br = R.assign_quad(xyz32, w, N, np.float32(N / Lbox))
    scale = np.max(np.abs(A))
    assert np.max(np.abs(A - Ar)) / scale < 1e-6
    assert np.max(np.abs(B - Br)) / scale < 1e-6


# ---------------------------------------------------------------- FFT -----
def test_fft_packed_is_textbook_dft():
    rng = np.random.default_rng(5)
    N = 6
    A = rng.standard_normal((N, N, N))
    B = rng.standard_normal((N, N, N))
    F = S.fft_packed(A, B)
    n = np.arange(N)
    E = np.exp(2j * np.pi * np.outer(n, n) / N)  # backward, unnormalised
    direct = np.einsum("ai,bj,ck,ijk->abc", E, E, E, A + 1j * B)
    assert np.allclose(F, direct, rtol=1e-12, atol=1e-10)


# -------------------------------------------------------------- fcomb -----
def _one_particle_closed_form(N, node):
    """delta(k) of one unit particle on grid-A node n (closed form of the
    interlaced PCS estimate divided by the PCS window): e^{+i theta.n}
    [prod(4 + 2cos t) + prod(5.75 cos(t/2) + 0.25 cos(3t/2))] / (432 W(k))."""
    f = S.signed_freq(N)
    out = np.ones((N // 2 + 1, N, N), dtype=np.complex128)
    th = [2 * np.pi * f[: N // 2 + 1] / N, 2 * np.pi * f / N, 2 * np.pi * f / N]
    A = [4 + 2 * np.cos(t) for t in th]
    Bf = [5.75 * np.cos(t / 2) + 0.25 * np.cos(3 * t / 2) for t in th]
    W = [np.where(t == 0, 1.0, (np.sin(t / 2) / np.where(t == 0, 1, t / 2)) ** 4) for t in th]
    ph = [np.exp(1j * t * n) for t, n in zip(th, node)]
    pa = A[0][:, None, None] * A[1][None, :, None] * A[2][None, None, :]
    pb = Bf[0][:, None, None] * Bf[1][None, :, None] * Bf[2][None, None, :]
    w = W[0][:, None, None] * W[1][None, :, None] * W[2][None, None, :]
    phase = ph[0][:, None, None] * ph[1][None, :, None] * ph[2][None, None, :]
    out = phase * (pa + pb) / (432.0 * w)
    return out


def test_fft_periodic_one_particle_closed_form():
    N, Lbox = 16, 16.0
    node = (3, 7, 12)
    xyz = np.array([[float(node[0])], [float(node[1])], [float(node[2])]])
    half = S.FFT_periodic(xyz, None, Lbox, N)
    cf = _one_particle_closed_form(N, node)
    # away from the aliasing planes (|k_d| = N/2) the combination is exact up
    # to the float32-seeded twiddle recurrence (~1e-6 relative)
    f = np.abs(S.signed_freq(N))
    inner = (f[: N // 2 + 1][:, None, None] < N // 2) & (f[None, :, None] < N // 2) & (f[None, None, :] < N // 2)
    assert np.allclose(half[inner], cf[inner], rtol=2e-5, atol=2e-5)
    assert abs(half[0, 0, 0] - 1.0) < 1e-6


@needs_ref
def test_fcomb_matches_reference_fortran():
    rng = np.random.default_rng(6)
    for N in (16, 24):
        F = (rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))).astype(np.complex64)
        ours = S.fcomb_periodic(F.astype(np.complex128), 123.0, N)
        theirs = R.fcomb_periodic(F, 123.0, N)
        scale = np.max(np.abs(ours))
        assert np.max(np.abs(ours - theirs)) / scale < 1e-6, N


def test_fcomb_hermitian_planes():
    # x = 0 and x = N/2 planes come out exactly conjugate-symmetric except at
    # their self-conjugate points (the Nyquist corners reflect_delta makes real)
    rng = np.random.default_rng(7)
    N = 12
    F = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    out = S.fcomb_periodic(F, 10.0, N)
    neg = (-np.arange(N)) % N
    selfc = np.zeros((N, N), dtype=bool)
    for y in (0, N // 2):
        for z in (0, N // 2):
            selfc[y, z] = True
    for x in (0, N // 2):
        P = out[x]
        assert np.array_equal(P[~selfc], np.conj(P[neg][:, neg])[~selfc])


# ------------------------------------------------------------- reflect ----
def test_reflect_is_hermitian_extension():
    rng = np.random.default_rng(8)
    N = 10
    F = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    half = S.fcomb_periodic(F, 5.0, N)[: N // 2 + 1]
    full = S.reflect_delta(half, N)
    for x in range(N):
        for y in range(N):
            for z in range(N):
                if x <= N // 2:
                    ref = half[x, y, z]
                    hg = N // 2
                    if all(c in (0, hg) for c in (x, y, z)) and (x, y, z) != (0, 0, 0):
                        ref = ref.real
                else:
                    ref = np.conj(half[N - x, (-y) % N, (-z) % N])
                assert full[x, y, z] == ref


# -------------------------------------------------------- shells / B ------
def _random_hermitian_full(N, seed):
    rng = np.random.default_rng(seed)
    r = rng.standard_normal((N, N, N)) + rng.standard_normal((N, N, N)) ** 2
    return np.fft.fftn(r)  # Hermitian by construction


def test_shell_power_parseval():
    N, step = 20, 2
    full = _random_hermitian_full(N, 9)
    D, p0k = S.shell_fields(full, step, 2, 4)
    irk = S.shell_index(N, step)
    for j in range(1, 5):
        direct = np.sum(np.abs(full[irk == j]) ** 2) / np.sum(irk == j)
        assert p0k[j - 1] == pytest.approx(direct, rel=1e-12)


def test_triangle_sum_matches_kspace_bruteforce():
    # sum_x d_i d_j d_l = N^3 sum_{k1+k2+k3 = 0 mod N} delta(k1) delta(k2) delta(k3)
    N, step, ncut, nmax = 12, 2, 2, 3
    full = _random_hermitian_full(N, 10)
    D, p0k = S.shell_fields(full, step, ncut, nmax)
    irk = S.shell_index(N, step)
    I, J, L, _ = S.triangles(nmax, ncut, step)
    for i, j, l in zip(I, J, L):
        real_space = np.einsum("i,i,i", D[i].ravel(), D[j].ravel(), D[l].ravel())
        mi, mj = np.argwhere(irk == i), np.argwhere(irk == j)
        k3 = (-(mi[:, None, :] + mj[None, :, :])) % N
        sel = irk[k3[..., 0], k3[..., 1], k3[..., 2]] == l
        vals = (full[tuple(mi.T)][:, None] * full[tuple(mj.T)][None, :] *
                full[k3[..., 0], k3[..., 1], k3[..., 2]])
        kspace = N ** 3 * np.sum(vals[sel]).real
        assert real_space == pytest.approx(kspace, rel=1e-9, abs=1e-6 * abs(kspace) + 1e-6)


def test_bk_zero_count_quirk():
    # zero-count triangles: values appended as 0, no i_k entry; after the
    # unit conversion b123 = 2/nbar^2 and q123 = 2/3 (pyspectrum.py:423-445, :346-355)
    N, step, ncut, nmax = 12, 1, 1, 9  # shells beyond the grid are empty
    rng = np.random.default_rng(11)
    xyz = rng.uniform(0, 100.0, size=(3, 400))
    b = S.Bk_periodic(xyz, Lbox=100.0, Ngrid=N, step=step, Ncut=ncut, Nmax=nmax)
    I, J, L, _ = S.triangles(nmax, ncut, step)
    assert len(b["b123"]) == len(I)
    zero = b["counts"] == 0
    assert zero.any() and len(b["i_k1"]) == np.sum(~zero)
    nbar = b["meta"]["nbar"]
    assert np.allclose(b["b123"][zero], 2 / nbar ** 2)
    assert np.allclose(b["q123"][zero], 2 / 3)


# -------------------------------------------------------------- P(k) -----
@needs_ref
@pytest.mark.parametrize("N", [16, 24, 32])
def test_pk_pbox_rsd_matches_reference_fortran(N):
    rng = np.random.default_rng(12 + N)
    F = rng.standard_normal((N, N, N)) + 1j * rng.standard_normal((N, N, N))
    half = S.fcomb_periodic(F, 77.0, N)[: N // 2 + 1].astype(np.complex64)
    for irsd in (0, 1, 2):
        for nmu in (5, 10):
            a = S.pk_pbox_rsd(half, irsd, 1000, N // 2, nmu, N)
            b = R.pk_pbox_rsd(half, irsd, 1000, N // 2, nmu, N)
            for name, x, y in zip("k p0 p2 p4 nk km mk pkm nkm".split(), a, b):
                if name in ("nk", "nkm"):
                    assert np.array_equal(x, y), (irsd, nmu, name)
                elif name in ("p2", "p4"):
                    # Legendre-weighted sums cancel: scale the error by the monopole
                    # (the reference squares |delta| in float32, the oracle in float64)
                    assert np.all(np.abs(x - y) <= 2e-6 * np.abs(y) + 1e-5 * b[1]), (irsd, nmu, name)
                else:
                    assert np.allclose(x, y, rtol=2e-6, atol=1e-12 * np.max(np.abs(y))), (irsd, nmu, name)


def test_pk_periodic_parseval_against_shells():
    # p0k of Pk_periodic (step 1 bins) equals the shell power of the B(k) path
    rng = np.random.default_rng(13)
    N, Lbox = 16, 200.0
    xyz = rng.uniform(0, Lbox, size=(3, 800))
    pk = S.Pk_periodic(xyz, Lbox=Lbox, Ngrid=N)
    half = S.FFT_periodic(xyz, None, Lbox, N)
    full = S.reflect_delta(half, N)
    _, p0k = S.shell_fields(full, 1, 1, N // 2)
    kf = 2 * np.pi / Lbox
    nbar = 800 / Lbox ** 3
    want = p0k * (2 * np.pi) ** 3 / kf ** 3 - 1 / nbar
    assert np.allclose(pk["p0k"], want, rtol=1e-10)


# --------------------------------------------------- units of B(k), P(k) -----
def test_bk_units_scale_with_box_size():
    """bk_finalize's unit conversion (pyspectrum.py:346-353) by dimensions:
    positions and Lbox scaled by s = 2 leave every float32 grid coordinate
    unchanged, so B must scale as s^6, P as s^3 and Q not at all.  An
    exponent slip in (2 pi)^6/kf^6 or (2 pi)^3/kf^3 breaks this."""
    rng = np.random.default_rng(41)
    N, Lbox = 24, 600.0
    xyz = rng.uniform(0, Lbox, size=(3, 30000))
    counts = S.counts_Bk123(N, 6, 2, 2)
    counts[np.rint(counts / N ** 3) == 0] = 0.0
    a = S.Bk_periodic(xyz, None, Lbox, N, 2, 2, 6, counts=counts)
    b = S.Bk_periodic(2.0 * xyz, None, 2.0 * Lbox, N, 2, 2, 6, counts=counts)
    ok = a["counts"] > 0
    assert np.allclose(b["b123"][ok], 64.0 * a["b123"][ok], rtol=1e-12, atol=0)
    assert np.allclose(b["b123_sn"][ok], 64.0 * a["b123_sn"][ok], rtol=1e-12, atol=0)
    assert np.allclose(b["p0k1"][ok], 8.0 * a["p0k1"][ok], rtol=1e-12, atol=0)
    assert np.allclose(b["q123"][ok], a["q123"][ok], rtol=1e-12, atol=0)


def test_bk_units_uniform_randoms_are_shot_noise():
    """The constants of the unit conversion: for Poisson-sampled uniform
    points the shot-noise-subtracted P(k) and B(k) scatter around zero, i.e.
    P_raw (2 pi)^3/kf^3 = 1/nbar and B_raw (2 pi)^6/kf^6 = B_sn on average.
    A wrong power of 2 pi would shift P/P_sn by >= 5 and B/B_sn by >= 5."""
    rng = np.random.default_rng(42)
    N, Lbox = 48, 1000.0
    xyz = rng.uniform(0, Lbox, size=(3, 400000))
    counts = S.counts_Bk123(N, 8, 3, 3)
    counts[np.rint(counts / N ** 3) == 0] = 0.0
    b = S.Bk_periodic(xyz, None, Lbox, N, 3, 3, 8, counts=counts)
    nbar = b["meta"]["nbar"]
    assert abs(np.mean(b["p0k1"]) * nbar) < 0.05
    assert abs(np.mean(b["b123"] / b["b123_sn"])) < 0.5
