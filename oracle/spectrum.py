"""FP64 numpy restatement of pySpectrum's periodic-box estimator (the oracle).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Array convention: numpy arrays indexed ``[ix, iy, iz]`` as in the reference
(``pyspectrum/pyspectrum.py:1060-1080``); index n of an axis of length Ngrid is
the signed frequency n for n <= Ngrid/2 and n - Ngrid above.

Precision: all arithmetic is float64 except where the reference's float32
arithmetic decides an *integer* (a cell index, a k or mu bin) or defines a
constant the reference hard-codes in float32; there the float32 operation
sequence is replayed literally, because those integers must match bit for bit.
"""
import math

import numpy as np

f32 = np.float32
f64 = np.float64


# ----------------------------------------------------------------------------
# integer helpers
# ----------------------------------------------------------------------------
def abs_freq(Ngrid):
    """|k| per axis index: min(i, Ngrid - i)  (pyspectrum.py:367, :696)."""
    n = np.arange(Ngrid, dtype=np.int64)
    return np.minimum(n, Ngrid - n)


def signed_freq(Ngrid):
    """Signed frequency of index n (n, or n - Ngrid above Ngrid/2)."""
    n = np.arange(Ngrid, dtype=np.int64)
    return np.where(n <= Ngrid // 2, n, n - Ngrid)


def kmag2(Ngrid):
    """m = a_x^2 + a_y^2 + a_z^2 on the full grid, int64 [ix, iy, iz]."""
    a = abs_freq(Ngrid)
    return a[:, None, None] ** 2 + a[None, :, None] ** 2 + a[None, None, :] ** 2


def _isqrt_arr(x):
    """floor(sqrt(x)) for a non-negative int64 array, exact."""
    r = np.floor(np.sqrt(x.astype(np.float64))).astype(np.int64)
    r = np.where(r * r > x, r - 1, r)
    r = np.where((r + 1) * (r + 1) <= x, r + 1, r)
    return r


def shell_of_m(m, step):
    """Shell index irk = int(sqrt(m)/step + 0.5)  (pyspectrum.py:370-378).

    Exact integer form: the largest n with (2n-1)*step <= 2*sqrt(m), i.e.
    n = (isqrt(4m // step^2) + 1) // 2.  Ties (even step, sqrt(m) = step(n-1/2))
    go to n, which is what the float64 expression does since those ties are
    exactly representable.
    """
    m = np.asarray(m, dtype=np.int64)
    q = _isqrt_arr((4 * m) // (step * step))
    return (q + 1) // 2


def shell_index(Ngrid, step):
    """irk on the full [ix, iy, iz] grid (pyspectrum.py:370-378)."""
    return shell_of_m(kmag2(Ngrid), step)


def shell_counts(Ngrid, step, Nmax):
    """Nk[j] = #(irk == j) for j = 0..Nmax  (pyspectrum.py:380)."""
    irk = shell_index(Ngrid, step).ravel()
    return np.bincount(irk, minlength=Nmax + 1)[: Nmax + 1].astype(np.int64)


def triangles(Nmax, Ncut, step):
    """Triangle loop order of _Bk_periodic (pyspectrum.py:415-417):
    i from Ncut//step..Nmax, j from Ncut//step..i, l from max(i-j, Ncut//step)..j.
    Returns int64 arrays (i, j, l) and fac (pyspectrum.py:418-422)."""
    j0 = Ncut // step
    I, J, L, F = [], [], [], []
    for i in range(j0, Nmax + 1):
        for j in range(j0, i + 1):
            for l in range(max(i - j, j0), j + 1):
                fac = 1.0
                if i == j == l:
                    fac = 6.0
                elif i == j or j == l:
                    fac = 2.0
                I.append(i); J.append(j); L.append(l); F.append(fac)
    return (np.array(I, dtype=np.int64), np.array(J, dtype=np.int64),
            np.array(L, dtype=np.int64), np.array(F, dtype=np.float64))


# ----------------------------------------------------------------------------
# mesh assignment (estimator.f:284-512, branch ia=ib=ic=id=0)
# ----------------------------------------------------------------------------
def clip_positions(xyz, Lbox):
    """float32 (3, N) positions clipped to [0, Lbox(1-1e-6)]  (pyspectrum.py:937-941)."""
    xyz = np.asarray(xyz)
    out = np.zeros((3, xyz.shape[1]), dtype=np.float32)
    for d in range(3):
        out[d] = np.clip(xyz[d], 0.0, Lbox * (1.0 - 1e-6))
    return out


def _pcs_axis(r32, kf_ks32, Ngrid, shift, offset=0.0):
    """Cell indices (0-based, wrapped) and the 4 PCS weights x6 along one axis.

    Follows estimator.f:302-340 (grid A, shift 0) and :305-307, :342-378 (grid B,
    shift 0.5): rx = kf_ks*x + 1 + offset (float32, evaluated left to right;
    1-based grid coordinate); the base cell is int(rx) and h = rx - int(rx);
    weights (1-h)^3, 4+(3h-6)h^2, 6 - (others), h^3 go to base-1, base,
    base+1, base+2.  The float32 coordinate is replayed exactly (it selects the
    cells); weights are float64.
    """
    rx = (f32(kf_ks32) * r32).astype(f32) + f32(1.0)
    rx = (rx + f32(offset)).astype(f32)
    if shift:
        rx = (rx + f32(0.5)).astype(f32)
    base1 = rx.astype(np.int64)                  # 1-based, int() truncation, rx >= 0
    h = rx.astype(f64) - base1.astype(f64)
    c = (base1 - 1) % Ngrid                      # 0-based base cell (grid B wraps, :353)
    h2 = h * h
    wm2 = (1.0 - h) ** 3
    wm1 = 4.0 + (3.0 * h - 6.0) * h2
    wp2 = h2 * h
    wp1 = 6.0 - wm2 - wm1 - wp2
    idx = np.stack([(c - 1) % Ngrid, c % Ngrid, (c + 1) % Ngrid, (c + 2) % Ngrid])
    wts = np.stack([wm2, wm1, wp1, wp2])
    return idx, wts


def assign_quad(xyz32, w, Ngrid, kf_ks32, offset=0.0):
    """PCS assignment with interlacing: returns float64 grids (A, B), [ix, iy, iz].

    Grid A holds the odd (1-based) slots dtl(2i-1, j, k), grid B the even slots
    dtl(2i, j, k) (estimator.f:380-508).  Each particle adds, to each grid, the
    64-cell tensor product of its per-axis weights times w.
    """
    xyz32 = np.asarray(xyz32, dtype=np.float32)
    w = np.asarray(w, dtype=np.float64)
    G = Ngrid ** 3
    grids = []
    for shift in (False, True):
        ix, wx = _pcs_axis(xyz32[0], kf_ks32, Ngrid, shift, offset)
        iy, wy = _pcs_axis(xyz32[1], kf_ks32, Ngrid, shift, offset)
        iz, wz = _pcs_axis(xyz32[2], kf_ks32, Ngrid, shift, offset)
        idx, val = [], []
        for a in range(4):
            for b in range(4):
                wab = wx[a] * wy[b]
                iab = (ix[a] * Ngrid + iy[b]) * Ngrid
                for c in range(4):
                    idx.append(iab + iz[c])
                    val.append(wab * wz[c] * w)
        flat = np.bincount(np.concatenate(idx), weights=np.concatenate(val), minlength=G)
        grids.append(flat.reshape(Ngrid, Ngrid, Ngrid))
    return grids[0], grids[1]


# ----------------------------------------------------------------------------
# FFT of the packed grid (pyspectrum.py:1060-1080)
# ----------------------------------------------------------------------------
def fft_packed(A, B):
    """Unnormalised backward (e^{+i}) 3D c2c FFT of A + iB, complex128."""
    G = A.size
    return np.fft.ifftn(A + 1j * B) * G


# ----------------------------------------------------------------------------
# fcomb_periodic (estimator.f:605-675)
# ----------------------------------------------------------------------------
def fcomb_twiddle_table(Ngrid):
    """xrec = recx**n, n = 0..Ngrid/2, by the reference's recurrence.

    estimator.f:607,617-623: tpi is an implicitly-REAL parameter (float32),
    tpiL = tpi/float(Ngrid) is a float32 quotient stored in REAL*8,
    recx = cmplx(dcos(piL), dsin(piL)) is a *single* precision complex, and the
    recurrence xrec = xrec*recx (:667) runs in COMPLEX*16 from xrec = 1.
    """
    tpi32 = f32(6.283185307)
    tpiL = f64(f32(tpi32 / f32(Ngrid)))
    piL = -tpiL / 2.0
    rec = complex(f64(f32(math.cos(piL))), f64(f32(math.sin(piL))))
    T = np.empty(Ngrid // 2 + 1, dtype=np.complex128)
    x = complex(1.0, 0.0)
    for n in range(Ngrid // 2 + 1):
        T[n] = x
        x = x * rec
    return T


def fcomb_window(Ngrid):
    """Per-axis PCS window W(n) = (sin(rk/2)/(rk/2))^4, n = 0..Ngrid/2.

    rk = tpiL*(n) is rounded to float32 as in estimator.f:626-645 (rkz is an
    implicitly-REAL variable); the sinc is evaluated in float64.
    """
    tpi32 = f32(6.283185307)
    tpiL = f64(f32(tpi32 / f32(Ngrid)))
    n = np.arange(Ngrid // 2 + 1)
    rk = (tpiL * n).astype(f32).astype(f64)
    W = np.ones_like(rk)
    nz = rk != 0.0
    W[nz] = (np.sin(rk[nz] / 2.0) / (rk[nz] / 2.0)) ** 4
    return W


def fcomb_periodic(F, sumw, Ngrid):
    """Interlacing combination + PCS deconvolution + 1/(6^3*4*sum w), in place.

    Literal restatement of estimator.f:605-675 on the full [ix, iy, iz] grid:
    for every (ix, iy, iz) in [0, Ngrid/2]^3 the eight values at {ix, -ix} x
    {iy, -iy} x {iz, -iz} are combined and the eight stores happen in the
    reference's order (:658-665), so that where indices alias (0 and Ngrid/2)
    the same value survives.  Iterations touch disjoint element sets, so the
    loop is vectorised over all iterations store by store.
    ``sumw`` is cast to float32 as f2py does for the REAL argument N.
    """
    F = np.array(F, dtype=np.complex128, copy=True)
    Nh = Ngrid // 2 + 1
    cf = 1.0 / (216.0 * 4.0 * f64(f32(sumw)))
    T = fcomb_twiddle_table(Ngrid)
    W = fcomb_window(Ngrid)
    n = np.arange(Nh)
    cn = (Ngrid - n) % Ngrid
    IX, IY, IZ = np.meshgrid(n, n, n, indexing="ij")
    CX, CY, CZ = np.meshgrid(cn, cn, cn, indexing="ij")
    xr, yr, zr = T[IX], T[IY], T[IZ]
    cfac = cf / (W[IX] * W[IY] * W[IZ])
    ci = 1j
    cma = ci * xr * yr * zr
    cmb = ci * xr * yr * np.conj(zr)
    cmc = ci * xr * np.conj(yr) * zr
    cmd = ci * xr * np.conj(yr * zr)
    c000 = F[IX, IY, IZ] * (1 - cma) + np.conj(F[CX, CY, CZ]) * (1 + cma)
    c001 = F[IX, IY, CZ] * (1 - cmb) + np.conj(F[CX, CY, IZ]) * (1 + cmb)
    c010 = F[IX, CY, IZ] * (1 - cmc) + np.conj(F[CX, IY, CZ]) * (1 + cmc)
    c011 = F[IX, CY, CZ] * (1 - cmd) + np.conj(F[CX, IY, IZ]) * (1 + cmd)
    F[IX, IY, IZ] = c000 * cfac
    F[IX, IY, CZ] = c001 * cfac
    F[IX, CY, IZ] = c010 * cfac
    F[IX, CY, CZ] = c011 * cfac
    F[CX, IY, IZ] = np.conj(F[IX, CY, CZ])
    F[CX, IY, CZ] = np.conj(F[IX, CY, IZ])
    F[CX, CY, IZ] = np.conj(F[IX, IY, CZ])
    F[CX, CY, CZ] = np.conj(F[IX, IY, IZ])
    return F


def FFT_periodic(xyz, w=None, Lbox=2600.0, Ngrid=360):
    """Half field delta(k), ix in [0, Ngrid/2]  (pyspectrum.py:909-959)."""
    N = xyz.shape[1]
    if w is None:
        w = np.ones(N)
    kf_ks = f32(float(Ngrid) / Lbox)
    xyz32 = clip_positions(xyz, Lbox)
    w32 = np.asarray(w, dtype=np.float32).astype(np.float64)  # f2py REAL cast
    A, B = assign_quad(xyz32, w32, Ngrid, kf_ks)
    F = fft_packed(A, B)
    F = fcomb_periodic(F, np.sum(w), Ngrid)
    return F[: Ngrid // 2 + 1]


# ----------------------------------------------------------------------------
# survey geometry (B0_survey, pyspectrum.py:13-282; FFT_survey_mono, :731-826)
# ----------------------------------------------------------------------------
C_KMS = 299792.458


def comoving_distance(z, H0=67.6, Om0=0.31):
    """Flat LCDM comoving distance [Mpc], adaptive quadrature per redshift
    (the astropy FlatLambdaCDM(H0, Om0) background: no radiation)."""
    from scipy.integrate import quad
    z = np.atleast_1d(np.asarray(z, dtype=np.float64))
    inv_e = lambda x: 1.0 / math.sqrt(Om0 * (1.0 + x) ** 3 + (1.0 - Om0))
    out = np.empty_like(z)
    cache = {}
    for k, zz in enumerate(z):
        v = cache.get(zz)
        if v is None:
            v = cache[zz] = quad(inv_e, 0.0, zz, epsabs=1e-15, epsrel=1e-13, limit=200)[0]
        out[k] = v
    return C_KMS / H0 * out


def radecz_to_cartesian(radecz, H0=67.6, Om0=0.31):
    """util.py:27-51: degrees -> radians, r = comoving distance * h [Mpc/h]."""
    ra = np.array(radecz[0], dtype=np.float64) * (np.pi / 180.)
    dec = np.array(radecz[1], dtype=np.float64) * (np.pi / 180.)
    rad = comoving_distance(radecz[2], H0, Om0) * (H0 / 100.0)
    return np.array([rad * np.cos(dec) * np.cos(ra), rad * np.cos(dec) * np.sin(ra),
                     rad * np.sin(dec)])


def fcomb_survey(F, Ngrid):
    """estimator.f:677-745: the fcomb_periodic combination with
    cf = 1/(6^3*4) (no division by the number of objects, :686)."""
    return fcomb_periodic(F, 1.0, Ngrid)


def FFT_survey_mono(radecz, nb, w=None, P0_fkp=1e6, Lbox=2600.0, Ngrid=360, H0=67.6, Om0=0.31):
    """pyspectrum.py:731-826: cartesian positions, FKP weights
    w/(1 + nb P0), the normalisation sums I12..I33, assignment with offset
    0.5*Ngrid (no clipping), backward FFT, fcomb_survey; returns
    (half field, Ntot, I12, I13, I22, I23, I33)."""
    N = np.asarray(radecz).shape[1]
    w = np.ones(N) if w is None else np.array(w, dtype=np.float64)
    nb = np.asarray(nb, dtype=np.float64)
    xyz = radecz_to_cartesian(radecz, H0, Om0)
    assert np.sum(xyz.max(axis=1) >= 0.5 * Lbox) + np.sum(np.abs(xyz.min(axis=1)) >= 0.5 * Lbox) == 0
    xyzs = xyz.astype(np.float32)
    Ntot = np.sum(w)
    w = w * (1. / (1. + nb * P0_fkp))
    I12 = np.sum(w ** 2)
    I13 = np.sum(w ** 3)
    I22 = np.sum(nb * w ** 2)
    I23 = np.sum(nb * w ** 3)
    I33 = np.sum(nb ** 2 * w ** 3)
    w32 = w.astype(np.float32).astype(np.float64)  # f2py REAL cast
    A, B = assign_quad(xyzs, w32, Ngrid, f32(float(Ngrid) / Lbox), offset=f32(0.5 * Ngrid))
    F = fcomb_survey(fft_packed(A, B), Ngrid)
    return F[: Ngrid // 2 + 1], Ntot, I12, I13, I22, I23, I33


def B0_survey_from_sums(braw_fn, p0k_raw, counts, alpha, I12, I13, I22, I23, I33, Nmax, Ncut,
                        step, Ngrid):
    """pyspectrum.py:246-282 on given shell powers and triangle sums:
    p0k -> p0k/I22 - (1+alpha) I12/I22; for counts > 0,
    b = (sum_x d_i d_j d_l / counts - (p_i+p_j+p_l) I23 - (1-alpha^2) I13) / I33,
    q = b / (p_i p_j + p_j p_l + p_l p_i), counts/(fac Ngrid^3); zeros (and
    no i_k entries) where counts = 0.  ``braw_fn(i, j, l)`` gives the sum."""
    p0k = np.array(p0k_raw, dtype=np.float64)
    p0k /= I22
    p0k -= (1. + alpha) * I12 / I22
    out = {k: [] for k in ("i_k1", "i_k2", "i_k3", "p0k1", "p0k2", "p0k3", "b123", "q123",
                           "counts")}
    for i in range(Ncut // step, Nmax + 1):
        for j in range(Ncut // step, i + 1):
            for l in range(max(i - j, Ncut // step), j + 1):
                fac = 1.
                if (j == l) and (i == j): fac = 6.
                if (i == j) and (j != l): fac = 2.
                if (i == l) and (l != j): fac = 2.
                if (j == l) and (l != i): fac = 2.
                c = counts[i - 1, j - 1, l - 1]
                if c > 0:
                    out["i_k1"].append(i); out["i_k2"].append(j); out["i_k3"].append(l)
                    b = braw_fn(i, j, l) / c
                    b -= (p0k[i - 1] + p0k[j - 1] + p0k[l - 1]) * I23 + (1. - alpha ** 2) * I13
                    b /= I33
                    out["p0k1"].append(p0k[i - 1]); out["p0k2"].append(p0k[j - 1])
                    out["p0k3"].append(p0k[l - 1])
                    out["b123"].append(b)
                    out["q123"].append(b / (p0k[i - 1] * p0k[j - 1] + p0k[j - 1] * p0k[l - 1]
                                            + p0k[l - 1] * p0k[i - 1]))
                    out["counts"].append(c / (fac * float(Ngrid ** 3)))
                else:
                    for k in ("p0k1", "p0k2", "p0k3", "b123", "q123", "counts"):
                        out[k].append(0.)
    res = {k: np.array(v, dtype=np.float64) for k, v in out.items()}
    for k in ("i_k1", "i_k2", "i_k3"):
        res[k] = np.array(out[k], dtype=np.int64) * step
    return res


def B0_survey(radecz, nbar, w, radecz_r, nbar_r, w_r, counts, P0_fkp=1e6, Lbox=2600.0, Ngrid=360,
              step=3, Ncut=3, Nmax=40, H0=67.6, Om0=0.31):
    """pyspectrum.py:13-132 + 135-282: delta = delta_d - alpha delta_r with
    alpha = Ntot_d/Ntot_r and the random catalogue's I's times alpha; shells
    and triangle sums as in the periodic path.  ``counts``: the N_tri cache."""
    d, Nd, *_ = FFT_survey_mono(radecz, nbar, w, P0_fkp, Lbox, Ngrid, H0, Om0)
    r, Nr, I12r, I13r, I22r, I23r, I33r = FFT_survey_mono(radecz_r, nbar_r, w_r, P0_fkp, Lbox,
                                                            Ngrid, H0, Om0)
    alpha = Nd / Nr
    I12, I13, I22, I23, I33 = (alpha * v for v in (I12r, I13r, I22r, I23r, I33r))
    delta = reflect_delta(d, Ngrid) - alpha * reflect_delta(r, Ngrid)
    D, p0k = shell_fields(delta, step, Ncut, Nmax)
    fn = lambda i, j, l: np.einsum("i,i,i", D[i].ravel(), D[j].ravel(), D[l].ravel())
    return B0_survey_from_sums(fn, p0k, counts, alpha, I12, I13, I22, I23, I33, Nmax, Ncut, step,
                               Ngrid)


def reflect_delta(delt, Ngrid):
    """Full Hermitian field from the half field  (pyspectrum.py:1134-1157).

    delta[ix] = delt[ix] for ix <= Ngrid/2; for ix > Ngrid/2
    delta[ix, iy, iz] = conj(delt[Ngrid-ix, -iy, -iz]); then the seven Nyquist
    corners keep only their real part.
    """
    Nh = Ngrid // 2 + 1
    delta = np.zeros((Ngrid, Ngrid, Ngrid), dtype=np.result_type(delt.dtype, np.complex64))
    delta[:Nh] = delt
    neg = (-np.arange(Ngrid)) % Ngrid
    ix = np.arange(Nh, Ngrid)
    delta[Nh:] = np.conj(delt[Ngrid - ix][:, neg][:, :, neg])
    hg = Ngrid // 2
    for c in [(hg, 0, 0), (0, hg, 0), (0, 0, hg), (0, hg, hg), (hg, 0, hg), (hg, hg, 0), (hg, hg, hg)]:
        delta[c] = delt[c].real
    return delta


# ----------------------------------------------------------------------------
# shell fields and the triangle sums (_Bk_periodic, pyspectrum.py:359-457)
# ----------------------------------------------------------------------------
def shell_fields(delta_full, step, Ncut, Nmax):
    """d_j(x) = Re FFT_forward[delta * 1(irk == j)], j = Ncut//step..Nmax.

    pyspectrum.py:387-404.  Returns (D, p0k) with D[j] float64 (unset rows
    zero) and p0k[j-1] = sum_x d_j^2 / G / Nk[j] (NaN when Nk[j] = 0).
    """
    Ngrid = delta_full.shape[0]
    G = Ngrid ** 3
    irk = shell_index(Ngrid, step)
    Nk = shell_counts(Ngrid, step, Nmax)
    j0 = Ncut // step
    D = np.zeros((Nmax + 1, Ngrid, Ngrid, Ngrid), dtype=np.float64)
    p0k = np.zeros(Nmax)
    for j in range(j0, Nmax + 1):
        D[j] = shell_field(delta_full, irk, j)
        with np.errstate(invalid="ignore", divide="ignore"):
            p0k[j - 1] = np.sum(D[j] * D[j]) / G / Nk[j]
    return D, p0k


def shell_field(delta_full, irk, j):
    """One shell of pyspectrum.py:390-400: Re FFT_forward[delta * 1(irk == j)]."""
    tmp = np.where(irk == j, delta_full, 0).astype(np.complex128)
    return np.fft.fftn(tmp).real


def unit_shell_fields(Ngrid, step, Ncut, Nmax):
    """Real FFT of the shell indicator I_j (delta = 1), FP64 (pyspectrum.py:975-1013)."""
    D, _ = shell_fields(np.ones((Ngrid,) * 3, dtype=np.complex128), step, Ncut, Nmax)
    return D


def counts_Bk123(Ngrid, Nmax, Ncut, step):
    """counts[i-1, j-1, l-1] = sum_x I_i I_j I_l on the triangle loop, FP64.

    pyspectrum.py:1015-1023 (cache-miss path of _counts_Bk123).  Equals
    Ngrid^3 * N_tri up to FFT round-off.
    """
    D = unit_shell_fields(Ngrid, step, Ncut, Nmax)
    I, J, L, _ = triangles(Nmax, Ncut, step)
    counts = np.zeros((Nmax, Nmax, Nmax), dtype=np.float64)
    for i, j, l in zip(I, J, L):
        counts[i - 1, j - 1, l - 1] = np.einsum("i,i,i", D[i].ravel(), D[j].ravel(), D[l].ravel())
    return counts


def ntri_bruteforce(Ngrid, Nmax, Ncut, step):
    """N_tri by integer enumeration: #{k1 in S_i, k2 in S_j, k3 in S_l,
    k1+k2+k3 = 0 (mod Ngrid)}.  Tiny grids only."""
    irk = shell_index(Ngrid, step)
    I, J, L, _ = triangles(Nmax, Ncut, step)
    modes = {}
    for s in range(Nmax + 1):
        modes[s] = np.argwhere(irk == s)
    out = np.zeros((Nmax, Nmax, Nmax), dtype=np.int64)
    for i, j, l in zip(I, J, L):
        mi, mj = modes[i], modes[j]
        if len(mi) == 0 or len(mj) == 0:
            continue
        k3 = (-(mi[:, None, :] + mj[None, :, :])) % Ngrid
        out[i - 1, j - 1, l - 1] = int(np.sum(irk[k3[..., 0], k3[..., 1], k3[..., 2]] == l))
    return out


def bk_from_shells(D, p0k, counts, Ngrid, step, Ncut, Nmax):
    """The triangle loop of _Bk_periodic (pyspectrum.py:412-457).

    ``counts`` is the (Nmax, Nmax, Nmax) normalisation.  Zero-count triangles
    append 0 to the value arrays but nothing to i_k* (reference behaviour).
    """
    G = float(Ngrid ** 3)
    I, J, L, FAC = triangles(Nmax, Ncut, step)
    out = {k: [] for k in ("i_k1", "i_k2", "i_k3", "p0k1", "p0k2", "p0k3", "b123", "q123", "counts")}
    for i, j, l, fac in zip(I, J, L, FAC):
        c = counts[i - 1, j - 1, l - 1]
        if c > 0:
            out["i_k1"].append(i); out["i_k2"].append(j); out["i_k3"].append(l)
            b = np.einsum("i,i,i", D[i].ravel(), D[j].ravel(), D[l].ravel())
            out["p0k1"].append(p0k[i - 1]); out["p0k2"].append(p0k[j - 1]); out["p0k3"].append(p0k[l - 1])
            out["b123"].append(b / c)
            out["q123"].append(b / c / (p0k[i - 1] * p0k[j - 1] + p0k[j - 1] * p0k[l - 1] + p0k[l - 1] * p0k[i - 1]))
            out["counts"].append(c / (fac * G))
        else:
            for k in ("p0k1", "p0k2", "p0k3", "b123", "q123", "counts"):
                out[k].append(0.0)
    res = {k: np.array(v) for k, v in out.items()}
    for k in ("i_k1", "i_k2", "i_k3"):
        res[k] = res[k].astype(np.int64) * step
    return res


def bk_finalize(bispec, Lbox, nbar):
    """Unit conversion and shot noise of Bk_periodic (pyspectrum.py:346-355)."""
    kf = 2 * np.pi / Lbox
    b = dict(bispec)
    for key in ("p0k1", "p0k2", "p0k3"):
        b[key] = b[key] * (2 * np.pi) ** 3 / kf ** 3 - 1.0 / nbar
    b["p0k_sn"] = 1.0 / nbar
    sn = (b["p0k1"] + b["p0k2"] + b["p0k3"]) / nbar + 1.0 / nbar ** 2
    b["b123"] = b["b123"] * (2 * np.pi) ** 6 / kf ** 6 - sn
    b["b123_sn"] = sn
    b["q123"] = b["b123"] / (b["p0k1"] * b["p0k2"] + b["p0k1"] * b["p0k3"] + b["p0k2"] * b["p0k3"])
    return b


def Bk_periodic(xyz, w=None, Lbox=2600, Ngrid=360, step=3, Ncut=3, Nmax=40, counts=None,
                return_shells=False):
    """pyspectrum.py:285-356 with the Ngrid == 360 assert lifted.

    ``counts`` defaults to the FP64 unit-shell computation (cache-miss path).
    """
    N = xyz.shape[1]
    if w is None:
        w = np.ones(N)
    nbar = np.sum(w) / Lbox ** 3
    kf = 2 * np.pi / Lbox
    delta = FFT_periodic(xyz, w=w, Lbox=Lbox, Ngrid=Ngrid)
    full = reflect_delta(delta, Ngrid)
    D, p0k = shell_fields(full, step, Ncut, Nmax)
    if counts is None:
        counts = counts_Bk123(Ngrid, Nmax, Ncut, step)
    b = bk_from_shells(D, p0k, counts, Ngrid, step, Ncut, Nmax)
    b["meta"] = {"Lbox": Lbox, "Ngrid": Ngrid, "step": step, "Ncut": Ncut, "Nmax": Nmax,
                 "N": N, "nbar": nbar, "kf": kf}
    raw = {k: b[k].copy() for k in ("b123", "p0k1", "p0k2", "p0k3")}
    b = bk_finalize(b, Lbox, nbar)
    b["_raw"] = raw
    if return_shells:
        return b, D, p0k
    return b


# ----------------------------------------------------------------------------
# P(k)
# ----------------------------------------------------------------------------
def pk_bins(Ngrid):
    """int(Nbins*rk/phys_nyq + 0.5) of pyspectrum.py:691-703 in exact integers
    (= round(sqrt(m)); no ties are possible)."""
    return shell_of_m(kmag2(Ngrid), 1)


def Pk_periodic(xyz, w=None, Lbox=2600, Ngrid=360):
    """pyspectrum.py:644-728, binning the full field delta_fft.

    The reference indexes the half field with a full-grid mask at :713, which
    numpy rejects; the evident intent (bin the reflected full field) is used.
    """
    N = xyz.shape[1]
    if w is None:
        w = np.ones(N)
    nbar = np.sum(w) / Lbox ** 3
    delta = FFT_periodic(xyz, w=w, Lbox=Lbox, Ngrid=Ngrid)
    full = reflect_delta(delta, Ngrid)
    return pk_from_full(full, Lbox, Ngrid, nbar, N)


def pk_from_full(full, Lbox, Ngrid, nbar, N):
    Nbins = Ngrid // 2
    kf = 2 * np.pi / float(Lbox)
    m = kmag2(Ngrid).ravel()
    irk = shell_of_m(m, 1)
    rk = kf * np.sqrt(m.astype(np.float64))
    p2 = np.abs(full.ravel().astype(np.complex128)) ** 2
    nk = np.bincount(irk, minlength=Nbins + 1)[1:Nbins + 1].astype(np.float64)
    ksum = np.bincount(irk, weights=rk, minlength=Nbins + 1)[1:Nbins + 1]
    psum = np.bincount(irk, weights=p2, minlength=Nbins + 1)[1:Nbins + 1]
    k = np.zeros(Nbins)
    p0k = np.zeros(Nbins)
    ok = nk > 0
    k[ok] = ksum[ok] / nk[ok]
    p0k[ok] = psum[ok] / nk[ok] / kf ** 3
    p0k *= (2.0 * np.pi) ** 3
    return {"meta": {"Lbox": Lbox, "Ngrid": Ngrid, "N": N, "nbar": nbar, "kf": kf},
            "k": k, "p0k": p0k - 1.0 / nbar, "counts": nk, "p0k_sn": 1.0 / nbar}


def _rsd_mu32(kx, ky, kz, irsd):
    """mu in float32 with the reference's operation order (estimator.f:165-212).

    Returns (rk32, mu32); rk32 = sqrt(kx^2+ky^2+kz^2) in float32.
    """
    kx = kx.astype(f32); ky = ky.astype(f32); kz = kz.astype(f32)
    pi32 = f32(3.141592654)
    if irsd == 0:
        th, ph = f32(f32(0.5) * pi32), f32(0.0)
    elif irsd == 1:
        th, ph = f32(f32(0.5) * pi32), f32(f32(0.5) * pi32)
    else:
        th, ph = f32(0.0), f32(0.0)
    rk = np.sqrt((kx * kx + ky * ky).astype(f32) + kz * kz).astype(f32)
    with np.errstate(invalid="ignore", divide="ignore"):
        cot1 = (kz / rk).astype(f32)
        sit1 = np.sqrt((f32(1.0) - (cot1 * cot1).astype(f32)).astype(f32)).astype(f32)
        pos = sit1 > 0
        den = (rk * sit1).astype(f32)
        cp = np.where(pos, (kx / den).astype(f32), f32(0.0)).astype(f32)
        sp = np.where(pos, (ky / den).astype(f32), f32(0.0)).astype(f32)
    # libm sinf/cosf of these constants are correctly rounded: round the f64 value
    sph, cph = f32(math.sin(f64(ph))), f32(math.cos(f64(ph)))
    sth, cth = f32(math.sin(f64(th))), f32(math.cos(f64(th)))
    cc = np.where(pos, ((sph * sp).astype(f32) + (cph * cp).astype(f32)).astype(f32), f32(0.0)).astype(f32)
    mu = ((cth * cot1).astype(f32) + ((sth * sit1).astype(f32) * cc).astype(f32)).astype(f32)
    return rk, mu


def pk_pbox_rsd(dtl, irsd, Lbox, Nbin, Nmu, Ngrid):
    """Restatement of estimator.f:155-264 on the half field dtl[ix<=Ngrid/2, iy, iz].

    Bins (k: nint(Nbin*rk/(Ngrid/2)) in float32; mu: int((|mu|+dmu)/dmu) in
    float64 of float32 values) replay the reference's float32 sequence exactly;
    the accumulations are float64 with per-mode |delta|^2 in float64.
    Returns (k, p0, p2, p4, nk, km, mk, pkm, nkm) like the Fortran routine.
    """
    Lbox_i = int(Lbox)
    tpi = f32(f32(2.0) * f32(3.141592654))
    kf = f32(tpi / f32(Lbox_i))
    mubin = f32(f32(1.0) / f32(Nmu))
    n = np.arange(Ngrid)
    # estimator.f:197-203: index iz -> rkz = mod(iz+Ngrid/2-2, Ngrid) - Ngrid/2 + 1 (1-based iz)
    sf = ((n + 1 + Ngrid // 2 - 2) % Ngrid) - Ngrid // 2 + 1
    KX, KY, KZ = np.meshgrid(sf, sf, sf, indexing="ij")
    rk, mu = _rsd_mu32(KX, KY, KZ, irsd)
    # nint: round half away from zero (no half-integers occur, see shell_of_m)
    imk = np.floor(((f32(Nbin) * rk).astype(f32) / f32(Ngrid // 2)).astype(f32).astype(f64) + 0.5)
    imk = np.where(np.isfinite(imk), imk, 0).astype(np.int64)
    mu64 = mu.astype(f64)
    md = f64(mubin)
    with np.errstate(invalid="ignore"):
        imu = np.floor((np.abs(mu64) + md) / md)
    imu = np.where(np.isfinite(imu), imu, -1).astype(np.int64)
    # amplitude: ix <= Ngrid/2 reads dtl directly, else the mirrored element (:223-227)
    ix = np.arange(Ngrid)
    cx = np.where(ix <= Ngrid // 2, ix, (Ngrid - ix) % Ngrid)
    cyz = np.arange(Ngrid)
    ncyz = (Ngrid - cyz) % Ngrid
    amp = np.empty((Ngrid, Ngrid, Ngrid), dtype=np.float64)
    low = ix <= Ngrid // 2
    amp[low] = np.abs(dtl[ix[low]].astype(np.complex128)) ** 2
    hi = ~low
    amp[hi] = np.abs(dtl[cx[hi]][:, ncyz][:, :, ncyz].astype(np.complex128)) ** 2
    Le2 = -0.5 + 1.5 * mu64 ** 2
    Le4 = 0.375 - 3.75 * mu64 ** 2 + 4.375 * mu64 ** 4
    sel = (imk <= Nbin) & (imk != 0)
    kb = imk[sel] - 1
    kval = (kf * rk[sel]).astype(f32).astype(f64)
    a = amp[sel]
    nk = np.bincount(kb, minlength=Nbin).astype(f64)
    k = np.bincount(kb, weights=kval, minlength=Nbin)
    p0 = np.bincount(kb, weights=a, minlength=Nbin)
    p2 = np.bincount(kb, weights=a * 5.0 * Le2[sel], minlength=Nbin)
    p4 = np.bincount(kb, weights=a * 9.0 * Le4[sel], minlength=Nbin)
    im = imu[sel]
    ms = (im <= Nmu) & (im > 0)
    flat = kb[ms] * Nmu + (im[ms] - 1)
    nkm = np.bincount(flat, minlength=Nbin * Nmu).astype(f64).reshape(Nbin, Nmu)
    km = np.bincount(flat, weights=kval[ms], minlength=Nbin * Nmu).reshape(Nbin, Nmu)
    mk = np.bincount(flat, weights=np.abs(mu64[sel][ms]), minlength=Nbin * Nmu).reshape(Nbin, Nmu)
    pkm = np.bincount(flat, weights=a[ms], minlength=Nbin * Nmu).reshape(Nbin, Nmu)
    kf3 = f64(f32(f32(kf * kf) * kf))
    ok = nk > 0
    k[ok] /= nk[ok]
    p0[ok] = p0[ok] / nk[ok] / kf3
    p2[ok] = p2[ok] / nk[ok] / kf3
    p4[ok] = p4[ok] / nk[ok] / kf3
    okm = nkm > 0
    km[okm] /= nkm[okm]
    mk[okm] /= nkm[okm]
    pkm[okm] = pkm[okm] / nkm[okm] / kf3
    return k, p0, p2, p4, nk, km, mk, pkm, nkm


def Pk_periodic_rsd(xyz, w=None, Lbox=2600, Ngrid=360, rsd=2, Nmubin=10):
    """pyspectrum.py:460-538 with the Fortran branch :628-641."""
    N = xyz.shape[1]
    nbar = float(N) / Lbox ** 3
    kf = 2 * np.pi / Lbox
    delta = FFT_periodic(xyz, w=w, Lbox=Lbox, Ngrid=Ngrid)
    k, p0, p2, p4, nk, km, mk, pkm, nkm = pk_pbox_rsd(delta, rsd, Lbox, Ngrid // 2, Nmubin, Ngrid)
    norm = (2.0 * np.pi) ** 3
    p0 = p0 * norm; p2 = p2 * norm; p4 = p4 * norm; pkm = pkm * norm
    return {"meta": {"Lbox": Lbox, "Ngrid": Ngrid, "N": N, "nbar": nbar, "kf": kf},
            "k": k, "p0k": p0 - 1.0 / nbar, "p2k": p2, "p4k": p4,
            "p_sn": np.repeat(1.0 / nbar, len(k)), "counts": nk,
            "k_kmu": km, "mu_kmu": mk, "p_kmu": pkm - 1.0 / nbar, "counts_kmu": nkm}


# ------------------------------------------------- quadrupole machinery -----
# estimator.f:514-604 and the weight branch of assign_quad (:291-299).  Half
# fields are indexed [ix, iy, iz] with ix = 0..N/2 (the Fortran (N/2+1, N, N)).
# Float32 semantics as in the Fortran: rk = sqrt(float(m)), unit components
# float(k)/rk, in FiveDelta2g_1/2 truncated to INTEGER (implicit typing of the
# undeclared kxh, kyh, kzh) unless unit=True.  Complex values are combined
# part by part in float32 in the Fortran's order; k = 0 is left unchanged.
def _quad_modes(Ngrid):
    n = np.arange(Ngrid)
    f = np.where(n <= Ngrid // 2, n, n - Ngrid)
    kx = np.arange(Ngrid // 2 + 1)[:, None, None]
    ky = f[None, :, None]
    kz = f[None, None, :]
    m = (kx * kx + ky * ky + kz * kz).astype(np.int64)
    rk = np.sqrt(m.astype(f32))
    return kx, ky, kz, rk


def _quad_units(Ngrid, unit):
    kx, ky, kz, rk = _quad_modes(Ngrid)
    ok = rk > 0
    with np.errstate(invalid="ignore", divide="ignore"):
        u = [np.where(ok, (np.broadcast_to(k, rk.shape).astype(f32) / rk).astype(f32), f32(0))
             for k in (kx, ky, kz)]
    if not unit:
        u = [np.trunc(v).astype(f32) for v in u]
    return ok, u


def five_delta2g_1(xx, yy, zz, Ngrid, unit=False):
    """estimator.f:514-540: dcgxx = 7.5*(dcgxx*kxh**2 + dcgyy*kyh**2 + dcgzz*kzh**2)."""
    ok, (ux, uy, uz) = _quad_units(Ngrid, unit)
    kx2, ky2, kz2 = ux * ux, uy * uy, uz * uz
    out = np.array(xx, dtype=np.complex64, copy=True)
    parts = []
    for part in (np.real, np.imag):
        a, b, c = (part(np.asarray(v, np.complex64)).astype(f32) for v in (xx, yy, zz))
        parts.append(f32(7.5) * ((a * kx2 + b * ky2) + c * kz2))
    new = (parts[0] + 1j * parts[1]).astype(np.complex64)
    out[ok] = new[ok]
    return out


def five_delta2g_2(g, xx, xy, yz, zx, Ngrid, unit=False):
    """estimator.f:542-572: dcgxx = dcgxx + 7.5*(2*dcgxy*kxh*kyh + 2*dcgyz*kyh*kzh
    + 2*dcgzx*kzh*kxh) - 2.5*dcg."""
    ok, (ux, uy, uz) = _quad_units(Ngrid, unit)
    out = np.array(xx, dtype=np.complex64, copy=True)
    parts = []
    for part in (np.real, np.imag):
        d, a, p, q, r = (part(np.asarray(v, np.complex64)).astype(f32) for v in (g, xx, xy, yz, zx))
        two = f32(2)
        s = ((((two * p) * ux) * uy + ((two * q) * uy) * uz) + ((two * r) * uz) * ux)
        parts.append((a + f32(7.5) * s) - f32(2.5) * d)
    new = (parts[0] + 1j * parts[1]).astype(np.complex64)
    out[ok] = new[ok]
    return out


def build_quad(d1, d2, irsd, Ngrid):
    """estimator.f:574-604: dclr2 = (7.5*amu**2 - 2.5)*dclr1, amu = k_los/rk with
    irsd 1, 2, 3 = x, y, z (the Fortran STOPs otherwise)."""
    if irsd not in (1, 2, 3):
        raise ValueError("irsd must be 1, 2 or 3")
    kx, ky, kz, rk = _quad_modes(Ngrid)
    kl = (kx, ky, kz)[irsd - 1]
    ok = rk > 0
    with np.errstate(invalid="ignore", divide="ignore"):
        amu = (np.broadcast_to(kl, rk.shape).astype(f32) / rk).astype(f32)
    fac = f32(7.5) * (amu * amu) - f32(2.5)
    a = np.asarray(d1, np.complex64)
    new = ((fac * a.real.astype(f32)) + 1j * (fac * a.imag.astype(f32))).astype(np.complex64)
    out = np.array(d2, dtype=np.complex64, copy=True)
    out[ok] = new[ok]
    return out


def quad_weights(xyz32, w, ia, ib, ic, id_):
    """estimator.f:291-299: w, w*r(ia)*r(ib)/rnorm or w*r(ia)*r(ib)*r(ic)*r(id)/rnorm**2
    with rnorm = r(1)**2 + r(2)**2 + r(3)**2, all float32, left to right."""
    r = np.asarray(xyz32, dtype=f32)
    ww = np.ones(r.shape[1], f32) if w is None else np.asarray(w, dtype=f32)
    if ia == 0 and ib == 0 and ic == 0 and id_ == 0:
        return ww.copy()
    rnorm = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]
    if ic == 0 and id_ == 0:
        return ((ww * r[ia - 1]) * r[ib - 1]) / rnorm
    return ((((ww * r[ia - 1]) * r[ib - 1]) * r[ic - 1]) * r[id_ - 1]) / (rnorm * rnorm)


# ------------------------------------------------------- host utilities -----
def ijl_order(i_k, j_k, l_k):
    """util.py:8-24 ('GM' order): for each distinct l ascending, each j
    ascending among the triangles with that l, each i ascending among those
    with (l, j): the indices of the triangles equal to (i, j, l), in index
    order.  Restated as a plain grouping loop (test use only)."""
    groups = {}
    for n, key in enumerate(zip(np.ravel(l_k), np.ravel(j_k), np.ravel(i_k))):
        groups.setdefault(tuple(int(v) for v in key), []).append(n)
    return np.array([groups[k] for k in sorted(groups)])


def apply_rsd(xyz, vxyz, redshift, h, omega0_m, los, Lbox):
    """util.py:54-75: the line-of-sight coordinate becomes
    (x + (f v + Lbox)) % Lbox with f = (1+z)/(100 E(z)), E of flat LCDM with
    H0 = 100 h, a numpy float64 scalar (the cosmology object's efunc returns
    one, :67-70).  numpy 2 arithmetic: f v + Lbox is float64 for float32
    velocities; the in-place add at :73 adds in float64 and rounds to the
    positions' dtype; the wrap at :74 runs in that dtype (a Python float
    Lbox takes the array's dtype)."""
    ez = math.sqrt(omega0_m * (1.0 + redshift) ** 3 + (1.0 - omega0_m))
    f = (1 + redshift) / (100 * np.float64(ez))
    out = np.array(xyz, copy=True)
    a = {"x": 0, "y": 1, "z": 2}[los]
    col = out[a]
    col += f * np.asarray(vxyz)[a] + Lbox
    out[a] = col % Lbox
    return out


def survey_catalogue(radecz, nb, w, P0_fkp, Lbox, cosmo):
    """Host part of FFT_survey_mono (pyspectrum.py:767-806), float64, with a
    given cosmology object (comoving_distance, h): cartesian positions
    (util.py:27-51), the 'box not big enough' check (an AssertionError),
    Ntot = sum w before the FKP factor, FKP weights w/(1 + nb P0) and the sums
    I12, I13, I22, I23, I33.  Returns (float32 positions, weights, Ntot,
    (I12, I13, I22, I23, I33)); the inputs are not modified."""
    radecz = np.asarray(radecz, dtype=np.float64)
    nb = np.asarray(nb, dtype=np.float64)
    n = radecz.shape[1]
    w = np.ones(n) if w is None else np.array(w, dtype=np.float64)
    ra = radecz[0] * (np.pi / 180.)
    dec = radecz[1] * (np.pi / 180.)
    rad = np.asarray(cosmo.comoving_distance(radecz[2]) * cosmo.h, dtype=np.float64)
    xyz = np.array([rad * np.cos(dec) * np.cos(ra), rad * np.cos(dec) * np.sin(ra),
                    rad * np.sin(dec)])
    if np.sum(xyz.max(axis=1) >= 0.5 * Lbox) + np.sum(np.abs(xyz.min(axis=1)) >= 0.5 * Lbox):
        raise AssertionError("box not big enough!")
    Ntot = np.sum(w)
    w = w * (1. / (1. + nb * P0_fkp))
    sums = (np.sum(w ** 2), np.sum(w ** 3), np.sum(nb * w ** 2), np.sum(nb * w ** 3),
            np.sum(nb ** 2 * w ** 3))
    return np.ascontiguousarray(xyz.astype(np.float32)), w, Ntot, sums
