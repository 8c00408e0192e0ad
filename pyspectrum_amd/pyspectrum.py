"""Drop-in periodic-box estimator: pySpectrum's call signatures on gfx950 HIP kernels.

Mirrors ``pyspectrum/pyspectrum.py`` of the reference for the periodic-box
path -- same function names, keyword defaults, returned dict keys, dtypes and
ordering -- with every stage running in ``libpsx.so``:

=====================  ============================================  ==========================
function               reference                                      device entry point
=====================  ============================================  ==========================
FFT_periodic           pyspectrum.py:909-959                          psx_fft_periodic
reflect_delta          pyspectrum.py:1134-1157                        psx_reflect_delta
Pk_periodic            pyspectrum.py:644-728                          psx_bin_power
Pk_periodic_rsd        pyspectrum.py:460-538, :628-641                psx_pk_pbox_rsd
Bk_periodic            pyspectrum.py:285-356                          psx_bk_periodic
_Bk_periodic           pyspectrum.py:359-457                          psx_hermitian_half + psx_bk_raw
_counts_Bk123          pyspectrum.py:962-1030                         psx_bk_counts (cache miss)
=====================  ============================================  ==========================

Extra keyword ``precision`` selects the pipeline: ``'fp32'`` (default: FP32
FFTs, FP32 MFMA contraction, FP64 cross-tile sums), ``'bf16x3'`` (split-BF16
three-term MFMA contraction) or ``'fp64'``.  The split-BF16 bound (B within
1e-3 of FP64, relative to max(|B|, B_sn)) holds for triangles whose shells
lie inside the Nyquist sphere ((j + 1/2) step <= Ngrid/2, e.g. every shell
at Ngrid 360, step 3, Nmax 40).  Shells past it hold only modes near the
cube's corners, whose triangle sums cancel by orders of magnitude; there the
16 significant bits of each split operand give errors up to about 1e-2
(six times the FP32 pipeline's, measured at Ngrid 128), so such grids want
``'fp32'`` or ``'fp64'``.  ``fft`` and ``nthreads`` are
accepted and ignored (the reference's ``'fortran'``/``'numpy'`` FFT branches
are broken, pyspectrum.py:1076-1079).

Deliberate deviations (documented reference quirks):

* the ``Ngrid == 360`` assert of Bk_periodic (:332) is lifted;
* Pk_periodic bins the reflected full field; the reference indexes the half
  field with a full-grid mask at :713, which numpy rejects;
* a triangle whose count rounds to N_tri = 0 is treated as empty even if the
  FP64 count carries round-off of either sign;
* N_tri caches that are missing are computed on the GPU and written to a user
  cache directory, never into the package (reference: :1026-1028).

Inputs may be numpy arrays or torch tensors; torch CUDA tensors are used in
place (no host round trip).
"""
import os
import threading

import numpy as np
import torch

from . import _lib as L
from . import dat_dir, util

__all__ = ["FFT_periodic", "reflect_delta", "Pk_periodic", "Pk_periodic_rsd", "Bk_periodic",
           "_Bk_periodic", "_counts_Bk123", "FFTperiodic", "Bk123_periodic", "cache_dir",
           "FFT_survey_mono", "B0_survey", "_B0_survey"]


# --------------------------------------------------------------- helpers ----
def _prec(precision):
    try:
        return L.PREC[precision]
    except KeyError:
        raise ValueError("precision must be one of %s" % sorted(L.PREC)) from None


def _cdtype(prec):
    return torch.complex128 if prec == L.PREC["fp64"] else torch.complex64


def _device(device):
    return torch.device("cuda", device)


def _positions(xyz, device):
    """float32 [3, N] contiguous device tensor (the reference's float32 cast)."""
    if isinstance(xyz, torch.Tensor):
        t = xyz.to(device=_device(device), dtype=torch.float32)
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(xyz, dtype=np.float32))).to(_device(device))
    if t.dim() != 2 or t.shape[0] != 3:
        raise ValueError("xyz must have shape (3, N)")
    return t.contiguous()


def _weights(w, n, device):
    """(device float32 weights or None, sum w (float64), sum |w| (float64))."""
    if w is None:
        return None, float(n), float(n)
    if isinstance(w, torch.Tensor):
        wd = w.to(device=_device(device), dtype=torch.float32).contiguous()
        w64 = w.detach().to(dtype=torch.float64)
        return wd, float(w64.sum().item()), float(w64.abs().sum().item())
    w64 = np.asarray(w, dtype=np.float64)
    if w64.shape != (n,):
        raise ValueError("w must have shape (N,)")
    wd = torch.from_numpy(np.ascontiguousarray(w64.astype(np.float32))).to(_device(device))
    return wd, float(np.sum(w64)), float(np.sum(np.abs(w64)))


def _half_on_device(xyz, w, Lbox, Ngrid, prec, device):
    ctx = L.context(device)
    x = _positions(xyz, device)
    n = x.shape[1]
    wd, sumw, sumabs = _weights(w, n, device)
    Nh = Ngrid // 2 + 1
    half = torch.empty((Ngrid, Ngrid, Nh), dtype=_cdtype(prec), device=_device(device))
    ctx.call("psx_fft_periodic", L.ptr(x), L.ptr(wd), n, Ngrid, float(Lbox), sumw, sumabs,
           L.ptr(half), prec)
    return half, n, sumw


def _as_numpy_half(half):
    """Device half field [z][y][x<=N/2] -> numpy (N/2+1, N, N) indexed [ix, iy, iz]."""
    return half.permute(2, 1, 0).cpu().numpy()


def _check_ngrid(Ngrid):
    if Ngrid % 2:
        raise ValueError("Ngrid must be even")


# ------------------------------------------------------------- front end ----
def FFT_periodic(xyz, w=None, Lbox=2600., Ngrid=360, fft='pyfftw', silent=True, precision='fp32',
                 device=0):
    """Half field delta(k) of particles in a periodic box (pyspectrum.py:909-959).

    PCS assignment with interlacing, backward 3D FFT and the interlacing /
    window / normalisation combination of fcomb_periodic.  Returns a complex
    numpy array of shape (Ngrid//2+1, Ngrid, Ngrid) indexed [ix, iy, iz]
    (complex64, complex128 for precision='fp64').
    """
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    half, n, sumw = _half_on_device(xyz, w, Lbox, Ngrid, prec, device)
    if not silent:
        print('%i positions, Ntot=%.f' % (n, sumw))
    return _as_numpy_half(half)


def reflect_delta(delt, Ngrid=360, silent=True, device=0):
    """Full Hermitian field from the half field (pyspectrum.py:1134-1157).

    ``delt`` is (Ngrid//2+1, Ngrid, Ngrid) [ix, iy, iz]; returns the C-order
    (Ngrid, Ngrid, Ngrid) complex array (complex64 unless ``delt`` is complex128).
    """
    _check_ngrid(Ngrid)
    if not silent:
        print('reflecting the half field')
    d = delt if isinstance(delt, torch.Tensor) else torch.from_numpy(np.asarray(delt))
    prec = L.PREC["fp64"] if d.dtype == torch.complex128 else L.PREC["fp32"]
    Nh = Ngrid // 2 + 1
    if tuple(d.shape) != (Nh, Ngrid, Ngrid):
        raise ValueError("delt must have shape (Ngrid//2+1, Ngrid, Ngrid)")
    half = d.to(device=_device(device), dtype=_cdtype(prec)).permute(2, 1, 0).contiguous()
    full = torch.empty((Ngrid, Ngrid, Ngrid), dtype=_cdtype(prec), device=_device(device))
    L.context(device).call("psx_reflect_delta", L.ptr(half), L.ptr(full), Ngrid, prec)
    return full.cpu().numpy()


# ------------------------------------------------------------------ P(k) ----
def Pk_periodic(xyz, w=None, Lbox=2600, Ngrid=360, fft='pyfftw', silent=True, precision='fp32',
                device=0):
    """Power spectrum monopole in kf-wide shells (pyspectrum.py:644-728).

    Keys: k (mean |k| of the bin, h/Mpc), p0k (shot-noise subtracted), counts,
    p0k_sn, meta.
    """
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    N = xyz.shape[1]
    if w is None:
        w = np.ones(N)
    w_np = w.detach().cpu().numpy() if isinstance(w, torch.Tensor) else np.asarray(w)
    nbar = np.sum(w_np) / Lbox ** 3
    if not silent:
        print('------------------')
        print('%i positions in %i box' % (N, Lbox))
        print('nbar = %f' % nbar)
    half, _, _ = _half_on_device(xyz, w, Lbox, Ngrid, prec, device)
    Nbins = Ngrid // 2
    kf = 2 * np.pi / float(Lbox)
    sums = torch.empty(3 * Nbins, dtype=torch.float64, device=_device(device))
    L.context(device).call("psx_bin_power", L.ptr(half), Ngrid, 1, Nbins, kf,
           L.ptr(sums), prec)
    ksum, psum, nk = sums.cpu().numpy().reshape(3, Nbins)
    k = np.zeros(Nbins)
    p0k = np.zeros(Nbins)
    counts = np.zeros(Nbins)
    ok = nk > 0
    k[ok] = ksum[ok] / nk[ok]
    p0k[ok] = psum[ok] / nk[ok] / kf ** 3
    counts[ok] = nk[ok]
    p0k *= (2. * np.pi) ** 3
    meta = {'Lbox': Lbox, 'Ngrid': Ngrid, 'N': N, 'nbar': nbar, 'kf': kf}
    return {'meta': meta, 'k': k, 'p0k': p0k - 1. / nbar, 'counts': counts, 'p0k_sn': 1. / nbar}


def Pk_periodic_rsd(xyz, w=None, Lbox=2600, Ngrid=360, rsd=2, Nmubin=10, fft='pyfftw',
                    code='fortran', silent=True, precision='fp32', device=0):
    """Multipoles P0/P2/P4 and P(k, mu) with the line of sight along ``rsd``
    (pyspectrum.py:460-538; the estimator of estimator.f:155-264).

    Keys: k, p0k, p2k, p4k, p_sn, counts, k_kmu, mu_kmu, p_kmu, counts_kmu, meta.
    ``code`` is accepted for signature compatibility; the reference's
    'python' branch (pedagogical, :545-626) is not reproduced.
    """
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    N = xyz.shape[1]
    nbar = float(N) / Lbox ** 3
    kf = 2 * np.pi / Lbox
    if not silent:
        print('------------------')
        print('%i positions in %i box' % (N, Lbox))
        print('nbar = %f' % nbar)
    half, _, _ = _half_on_device(xyz, w, Lbox, Ngrid, prec, device)
    k, p0k, p2k, p4k, nk, k_kmu, mu_kmu, p_kmu, n_kmu = _pk_pbox_rsd(half, rsd, Lbox, Ngrid // 2,
                                                                      Nmubin, Ngrid, prec, device)
    pk_norm = (2. * np.pi) ** 3
    p0k = p0k * pk_norm
    p2k = p2k * pk_norm
    p4k = p4k * pk_norm
    p_kmu = p_kmu * pk_norm
    meta = {'Lbox': Lbox, 'Ngrid': Ngrid, 'N': N, 'nbar': nbar, 'kf': kf}
    return {'meta': meta, 'k': k, 'p0k': p0k - 1. / nbar, 'p2k': p2k, 'p4k': p4k,
            'p_sn': np.repeat(1. / nbar, len(k)), 'counts': nk, 'k_kmu': k_kmu,
            'mu_kmu': mu_kmu, 'p_kmu': p_kmu - 1. / nbar, 'counts_kmu': n_kmu}


def _pk_pbox_rsd(half, irsd, Lbox, Nbin, Nmu, Ngrid, prec, device, ctx=None):
    """pk_pbox_rsd on a device half field; the final divisions as estimator.f:246-263."""
    ctx = ctx or L.context(device)
    out = torch.empty(5 * Nbin + 4 * Nbin * Nmu, dtype=torch.float64, device=_device(device))
    ctx.call("psx_pk_pbox_rsd", L.ptr(half), int(irsd), int(Lbox), Nbin,
           Nmu, Ngrid, L.ptr(out), prec)
    o = out.cpu().numpy()
    k, p0, p2, p4, nk = [o[i * Nbin:(i + 1) * Nbin].copy() for i in range(5)]
    rest = o[5 * Nbin:].reshape(4, Nbin, Nmu)
    km, mk, pkm, nkm = [np.array(rest[i]) for i in range(4)]
    f32 = np.float32
    kf = f32(f32(f32(2.0) * f32(3.141592654)) / f32(int(Lbox)))
    kf3 = np.float64(f32(f32(kf * kf) * kf))
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


# ------------------------------------------------------------------ B(k) ----
def cache_dir():
    """Directory of computed N_tri caches (env PYSPECTRUM_AMD_CACHE, else ~/.cache)."""
    d = os.environ.get("PYSPECTRUM_AMD_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "pyspectrum_amd"))
    return d


_COUNTS_MEM = {}
_COUNTS_LOCK = threading.Lock()


def _counts_name(Ngrid, Nmax, Ncut, step, fft='pyfftw'):
    return ''.join(['counts', '.Ngrid', str(Ngrid), '.Nmax', str(Nmax), '.Ncut', str(Ncut),
                    '.step', str(step), '.', fft])


def _read_fortran_record(path, n):
    raw = open(path, "rb").read()
    head = int(np.frombuffer(raw[:4], dtype="<i4")[0])
    if head != 8 * n or len(raw) != 8 * n + 8:
        raise ValueError("%s: not a FortranFile record of %d float64" % (path, n))
    return np.frombuffer(raw[4:-4], dtype="<f8").copy()


def _write_fortran_record(path, arr):
    a = np.ascontiguousarray(arr, dtype="<f8")
    marker = np.array([a.nbytes], dtype="<i4").tobytes()
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "wb") as f:
        f.write(marker + a.tobytes() + marker)
    os.replace(tmp, path)


def _counts_Bk123(Ngrid=360, Nmax=40, Ncut=3, step=3, fft='pyfftw', silent=True, device=0):
    """Bispectrum normalisation counts[i-1, j-1, l-1] = sum_x I_i I_j I_l
    (pyspectrum.py:962-1030).

    Same FortranFile format and file name as the reference's cache
    (``counts.Ngrid{N}.Nmax{M}.Ncut{c}.step{s}.{fft}``, one record of Nmax^3
    float64, C order).  Looked up in the package data directory, then in
    ``cache_dir()``; on a miss it is computed on the GPU from FP64 unit-shell
    fields and written to ``cache_dir()``.
    """
    name = _counts_name(Ngrid, Nmax, Ncut, step, fft)
    with _COUNTS_LOCK:
        hit = _COUNTS_MEM.get(name)
        if hit is not None:
            return hit
        for d in (dat_dir(), cache_dir()):
            path = os.path.join(d, name)
            if os.path.isfile(path):
                counts = _read_fortran_record(path, Nmax ** 3).reshape(Nmax, Nmax, Nmax)
                _COUNTS_MEM[name] = counts
                return counts
        if not silent:
            print("--- calculating %s ---" % name)
        counts = _counts_on_gpu(Ngrid, Nmax, Ncut, step, device)
        _COUNTS_MEM[name] = counts
        try:
            os.makedirs(cache_dir(), exist_ok=True)
            _write_fortran_record(os.path.join(cache_dir(), name), counts)
        except OSError:
            pass  # a read-only cache only costs a recomputation
        return counts


def _counts_on_gpu(Ngrid, Nmax, Ncut, step, device=0):
    ijl = L.triangles(step, Ncut, Nmax)
    out = torch.zeros(max(len(ijl), 1), dtype=torch.float64, device=_device(device))
    L.context(device).call("psx_bk_counts", Ngrid, step, Ncut, Nmax, L.ptr(out))
    vals = out.cpu().numpy()[: len(ijl)]
    counts = np.zeros((Nmax, Nmax, Nmax), dtype=np.float64)
    counts[ijl[:, 0] - 1, ijl[:, 1] - 1, ijl[:, 2] - 1] = vals
    return counts


def _tri_counts(counts, ijl, Ngrid):
    """Per-triangle normalisation; counts that round to N_tri = 0 become 0."""
    c = np.asarray(counts)[ijl[:, 0] - 1, ijl[:, 1] - 1, ijl[:, 2] - 1].astype(np.float64)
    G = float(Ngrid) ** 3
    c[np.rint(c / G) == 0] = 0.0
    return c


def _assemble(ijl, step, fin, valid):
    """Reference dict layout: i_k* only for counted triangles (pyspectrum.py:423-445)."""
    keys = ('p0k1', 'p0k2', 'p0k3', 'b123', 'q123', 'counts', 'b123_sn')
    out = {k: fin[n] for n, k in enumerate(keys)}
    out['i_k1'] = ijl[valid, 0].astype(np.int64) * step
    out['i_k2'] = ijl[valid, 1].astype(np.int64) * step
    out['i_k3'] = ijl[valid, 2].astype(np.int64) * step
    return out


def Bk_periodic(xyz, w=None, Lbox=2600, Ngrid=360, step=3, Ncut=3, Nmax=40, fft='pyfftw', nthreads=1,
                silent=True, precision='fp32', device=0):
    """Bispectrum of particles in a periodic box (pyspectrum.py:285-356).

    Keys: i_k1, i_k2, i_k3 (triangle sides in units of kf), p0k1, p0k2, p0k3,
    b123, q123, counts, p0k_sn, b123_sn, meta -- same order and dtypes as the
    reference (arrays over the triangle loop of pyspectrum.py:415-417).
    """
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    N = xyz.shape[1]
    w_np = None
    if w is not None:
        w_np = w.detach().cpu().numpy() if isinstance(w, torch.Tensor) else np.asarray(w)
    sumw = float(N) if w_np is None else np.sum(w_np)
    nbar = sumw / Lbox ** 3
    kf = 2 * np.pi / Lbox
    if not silent:
        print('------------------')
        print('%i positions in %i box' % (N, Lbox))
        print('sum w_i = %f' % sumw)
        print('nbar = %f' % nbar)
    ijl = L.triangles(step, Ncut, Nmax)
    counts = _counts_Bk123(Ngrid=Ngrid, Nmax=Nmax, Ncut=Ncut, step=step, fft=fft, silent=silent,
                           device=device)
    tc = _tri_counts(counts, ijl, Ngrid)
    dev = _device(device)
    ctx = L.context(device)
    x = _positions(xyz, device)
    wd, sw, sabs = _weights(w, N, device)
    ntri = len(ijl)
    cnt_d = torch.from_numpy(tc).to(dev)
    ijl_d = torch.from_numpy(np.ascontiguousarray(ijl)).to(dev)
    out7 = torch.empty((7, max(ntri, 1)), dtype=torch.float64, device=dev)
    p0k = torch.empty(Nmax, dtype=torch.float64, device=dev)
    ctx.call("psx_bk_periodic", L.ptr(x), L.ptr(wd), N, Ngrid, float(Lbox), sw, sabs, step,
           Ncut, Nmax, L.ptr(cnt_d), L.ptr(ijl_d), float(nbar), (2 * np.pi) ** 3, kf ** 3,
           (2 * np.pi) ** 6, kf ** 6, L.ptr(out7), L.ptr(p0k), prec)
    fin = out7.cpu().numpy()[:, :ntri]
    bispec = _assemble(ijl, step, fin, tc > 0)
    out = {k: bispec[k] for k in ('i_k1', 'i_k2', 'i_k3', 'p0k1', 'p0k2', 'p0k3', 'b123', 'q123',
                                  'counts')}
    out['meta'] = {'Lbox': Lbox, 'Ngrid': Ngrid, 'step': step, 'Ncut': Ncut, 'Nmax': Nmax, 'N': N,
                   'nbar': nbar, 'kf': kf}
    out['p0k_sn'] = 1. / nbar
    out['b123_sn'] = bispec['b123_sn']
    return out


def _Bk_periodic(delta, Nmax=40, Ncut=3, step=3, fft='pyfftw', nthreads=1, silent=True,
                 precision='fp32', device=0):
    """Bispectrum of a full delta(k) field, before units and shot noise
    (pyspectrum.py:359-457).  ``delta``: (Ngrid, Ngrid, Ngrid) complex [ix, iy, iz].
    """
    d = delta if isinstance(delta, torch.Tensor) else torch.from_numpy(np.asarray(delta))
    Ngrid = d.shape[0]
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    dev = _device(device)
    ctx = L.context(device)
    in_prec = L.PREC["fp64"] if d.dtype == torch.complex128 else L.PREC["fp32"]
    full = d.to(device=dev, dtype=_cdtype(in_prec))
    sx, sy, sz = full.stride()
    Nh = Ngrid // 2 + 1
    half = torch.empty((Ngrid, Ngrid, Nh), dtype=_cdtype(prec), device=dev)
    ctx.call("psx_hermitian_half", L.ptr(full), in_prec, sx, sy, sz, L.ptr(half), Ngrid,
           prec)
    ijl = L.triangles(step, Ncut, Nmax)
    ntri = len(ijl)
    braw = torch.empty(max(ntri, 1), dtype=torch.float64, device=dev)
    p0k_d = torch.empty(Nmax, dtype=torch.float64, device=dev)
    ctx.call("psx_bk_raw", L.ptr(half), Ngrid, step, Ncut, Nmax, L.ptr(braw), L.ptr(p0k_d),
           prec)
    counts = _counts_Bk123(Ngrid=Ngrid, Nmax=Nmax, Ncut=Ncut, step=step, fft=fft, silent=silent,
                           device=device)
    tc = _tri_counts(counts, ijl, Ngrid)
    b = braw.cpu().numpy()[:ntri]
    p0k = p0k_d.cpu().numpy()
    G = float(Ngrid ** 3)
    i, j, l = ijl[:, 0] - 1, ijl[:, 1] - 1, ijl[:, 2] - 1
    fac = np.ones(ntri)
    fac[(ijl[:, 0] == ijl[:, 1]) | (ijl[:, 1] == ijl[:, 2])] = 2.0
    fac[(ijl[:, 0] == ijl[:, 1]) & (ijl[:, 1] == ijl[:, 2])] = 6.0
    valid = tc > 0
    z = np.zeros(ntri)
    p1 = np.where(valid, p0k[i], z)
    p2 = np.where(valid, p0k[j], z)
    p3 = np.where(valid, p0k[l], z)
    with np.errstate(divide='ignore', invalid='ignore'):
        b123 = np.where(valid, b / np.where(valid, tc, 1.0), z)
        q123 = np.where(valid, b123 / (p1 * p2 + p2 * p3 + p3 * p1), z)
    cnt = np.where(valid, tc / (fac * G), z)
    return {'i_k1': ijl[valid, 0].astype(np.int64) * step, 'i_k2': ijl[valid, 1].astype(np.int64) * step,
            'i_k3': ijl[valid, 2].astype(np.int64) * step, 'p0k1': p1, 'p0k2': p2, 'p0k3': p3,
            'b123': b123, 'q123': q123, 'counts': cnt}


# ------------------------------------------------------- survey geometry ----
def _survey_half_on_device(radecz, nb, w, P0_fkp, Lbox, Ngrid, cosmo, prec, device):
    """FFT_survey_mono on the device: psx_survey_catalogue (positions, FKP
    weights, normalisation sums), the box check, then psx_fft_survey_mono."""
    if cosmo is None:
        cosmo = util.default_cosmology()
    dev = _device(device)

    def dev64(a, shape):
        t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a, dtype=np.float64))
        t = t.to(device=dev, dtype=torch.float64).contiguous()
        if tuple(t.shape) != shape:
            raise ValueError("expected shape %s, got %s" % (shape, tuple(t.shape)))
        return t

    n = radecz.shape[1]
    if n == 0:
        raise ValueError("empty catalogue")  # the reference fails in np.max (pyspectrum.py:779)
    rd = dev64(radecz, (3, n))
    nbd = dev64(nb, (n,))
    wd_in = None if w is None else dev64(w, (n,))
    z = rd[2]
    zmax = float(z.max().item()) if n else 0.0
    if n and (not bool(torch.isfinite(z).all().item()) or float(z.min().item()) < 0):
        raise ValueError("redshifts must be finite and non-negative")
    Dn, sl = util.distance_table(cosmo, zmax)
    xyz = torch.empty((3, n), dtype=torch.float32, device=dev)
    wf = torch.empty(n, dtype=torch.float32, device=dev)
    sums = np.zeros(L.SURVEY_SUMS)
    ctx = L.context(device)
    ctx.call("psx_survey_catalogue", L.ptr(rd), L.ptr(nbd), L.ptr(wd_in), n,
           Dn.ctypes.data_as(L._p), sl.ctypes.data_as(L._p), len(Dn) - 1, zmax,
           float(_value_h(cosmo)), float(P0_fkp), L.ptr(xyz), L.ptr(wf),
           sums.ctypes.data_as(L._p))
    Ntot, I12, I13, I22, I23, I33, sabs = (float(v) for v in sums[:7])
    xyz_max, xyz_min = sums[7:10], -sums[10:13]
    assert np.sum(xyz_max >= 0.5 * Lbox) + np.sum(np.abs(xyz_min) >= 0.5 * Lbox) == 0, \
        'box not big enough!'
    assert np.all([(I12 >= 0), (I13 >= 0), (I22 >= 0), (I23 >= 0), (I33 >= 0)])
    half = torch.empty((Ngrid, Ngrid, Ngrid // 2 + 1), dtype=_cdtype(prec), device=dev)
    ctx.call("psx_fft_survey_mono", L.ptr(xyz), L.ptr(wf), n, Ngrid, float(Lbox), sabs,
           L.ptr(half), prec)
    return half, Ntot, (I12, I13, I22, I23, I33)


def _value_h(cosmo):
    return float(np.asarray(getattr(cosmo.h, "value", cosmo.h)))


def FFT_survey_mono(radecz, nb, w=None, P0_fkp=1e6, Lbox=2600., Ngrid=360, cosmo=None, fft='pyfftw',
                    silent=True, precision='fp32', device=0):
    """Half field delta_0(k) of a survey catalogue (pyspectrum.py:731-826).

    (RA, Dec, z) -> comoving Mpc/h (``cosmo``: an object with
    ``comoving_distance(z)`` [Mpc] and ``h``; default FlatLambdaCDM(67.6,
    0.31)), FKP weights w/(1 + nb P0_fkp), assignment at offset Ngrid/2,
    backward FFT and fcomb_survey.  Returns (delta0, Ntot, I12, I13, I22,
    I23, I33) with delta0 of shape (Ngrid//2+1, Ngrid, Ngrid) [ix, iy, iz].
    """
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    half, Ntot, I = _survey_half_on_device(radecz, nb, w, P0_fkp, Lbox, Ngrid, cosmo, prec, device)
    if not silent:
        print('Ntot=%.2f' % Ntot)
        for name, v in zip(('I12', 'I13', 'I22', 'I23', 'I33'), I):
            print('%s=%.2e' % (name, v))
    return (_as_numpy_half(half), Ntot) + tuple(I)


def _B0_from_half(half, alpha, I12, I13, I22, I23, I33, Nmax, Ncut, step, fft, silent, prec,
                  device):
    """Shell fields + triangle sums on the device (psx_bk_raw), then the
    survey normalisation and shot noise of pyspectrum.py:246-282 (float64)."""
    Ngrid = half.shape[0]
    dev = half.device
    ijl = L.triangles(step, Ncut, Nmax)
    ntri = len(ijl)
    braw_d = torch.empty(max(ntri, 1), dtype=torch.float64, device=dev)
    p0k_d = torch.empty(Nmax, dtype=torch.float64, device=dev)
    L.context(device).call("psx_bk_raw", L.ptr(half), Ngrid, step, Ncut, Nmax,
           L.ptr(braw_d), L.ptr(p0k_d), prec)
    counts = _counts_Bk123(Ngrid=Ngrid, Nmax=Nmax, Ncut=Ncut, step=step, fft=fft, silent=silent,
                           device=device)
    tc = _tri_counts(counts, ijl, Ngrid)
    braw = braw_d.cpu().numpy()[:ntri]
    p0k = p0k_d.cpu().numpy()
    p0k /= I22
    p0k -= (1. + alpha) * I12 / I22
    G = float(Ngrid ** 3)
    i, j, l = ijl[:, 0] - 1, ijl[:, 1] - 1, ijl[:, 2] - 1
    fac = np.ones(ntri)
    fac[(ijl[:, 0] == ijl[:, 1]) | (ijl[:, 1] == ijl[:, 2])] = 2.0
    fac[(ijl[:, 0] == ijl[:, 1]) & (ijl[:, 1] == ijl[:, 2])] = 6.0
    valid = tc > 0
    z = np.zeros(ntri)
    p1 = np.where(valid, p0k[i], z)
    p2 = np.where(valid, p0k[j], z)
    p3 = np.where(valid, p0k[l], z)
    with np.errstate(divide='ignore', invalid='ignore'):
        b = braw / np.where(valid, tc, 1.0)
        b -= (p1 + p2 + p3) * I23 + (1. - alpha ** 2) * I13
        b /= I33
        b123 = np.where(valid, b, z)
        q123 = np.where(valid, b123 / (p1 * p2 + p2 * p3 + p3 * p1), z)
    cnt = np.where(valid, tc / (fac * G), z)
    return {'i_k1': ijl[valid, 0].astype(np.int64) * step, 'i_k2': ijl[valid, 1].astype(np.int64) * step,
            'i_k3': ijl[valid, 2].astype(np.int64) * step, 'p0k1': p1, 'p0k2': p2, 'p0k3': p3,
            'b123': b123, 'q123': q123, 'counts': cnt}


def _B0_survey(delta, alpha, I12, I13, I22, I23, I33, Nmax=40, Ncut=3, step=3, fft='pyfftw',
               nthreads=1, silent=True, precision='fp32', device=0):
    """Survey bispectrum monopole of a full delta(k) = delta_data - alpha
    delta_random (pyspectrum.py:135-282).  ``delta``: (Ngrid, Ngrid, Ngrid)
    complex [ix, iy, iz]; the I's are the random catalogue's sums times alpha."""
    d = delta if isinstance(delta, torch.Tensor) else torch.from_numpy(np.asarray(delta))
    Ngrid = d.shape[0]
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    dev = _device(device)
    in_prec = L.PREC["fp64"] if d.dtype == torch.complex128 else L.PREC["fp32"]
    full = d.to(device=dev, dtype=_cdtype(in_prec))
    sx, sy, sz = full.stride()
    half = torch.empty((Ngrid, Ngrid, Ngrid // 2 + 1), dtype=_cdtype(prec), device=dev)
    L.context(device).call("psx_hermitian_half", L.ptr(full), in_prec, sx, sy, sz,
           L.ptr(half), Ngrid, prec)
    return _B0_from_half(half, alpha, I12, I13, I22, I23, I33, Nmax, Ncut, step, fft, silent, prec,
                         device)


def B0_survey(radecz, nbar, w=None, radecz_r=None, nbar_r=None, w_r=None, P0_fkp=1e6, Lbox=2600,
              Ngrid=360, step=3, Ncut=3, Nmax=40, cosmo=None, fft='pyfftw', nthreads=1, silent=True,
              precision='fp32', device=0):
    """Bispectrum monopole in survey geometry, Scoccimarro (2015) estimator
    (pyspectrum.py:13-132).

    Data and random catalogues (RA, Dec, z) with number densities ``nbar``,
    ``nbar_r`` and optional weights; delta = delta_data - alpha delta_random
    with alpha = sum w / sum w_r.  Keys: i_k1, i_k2, i_k3, p0k1, p0k2, p0k3,
    b123, q123, counts, meta.
    """
    if radecz_r is not None and nbar_r is None:
        raise ValueError('specify nbar for random catalog')
    if radecz_r is None:
        raise ValueError('a random catalog (radecz_r, nbar_r) is required')
    if cosmo is None:
        cosmo = util.default_cosmology()
    _check_ngrid(Ngrid)
    prec = _prec(precision)
    kf = 2 * np.pi / Lbox
    N = radecz.shape[1]
    half_d, Ngtot, _ = _survey_half_on_device(radecz, nbar, w, P0_fkp, Lbox, Ngrid, cosmo, prec,
                                              device)
    half_r, Nrtot, Ir = _survey_half_on_device(radecz_r, nbar_r, w_r, P0_fkp, Lbox, Ngrid, cosmo,
                                               prec, device)
    alpha = Ngtot / Nrtot
    if not silent:
        print('alpha=%e' % alpha)
    I12, I13, I22, I23, I33 = (alpha * v for v in Ir)
    # delta_d - alpha delta_r on the half fields (reflect_delta is linear); the
    # product and the difference are rounded separately, as in numpy
    half = half_d - alpha * half_r
    del half_d, half_r
    bispec = _B0_from_half(half, alpha, I12, I13, I22, I23, I33, Nmax, Ncut, step, fft, True, prec,
                           device)
    bispec['meta'] = {'Lbox': Lbox, 'Ngrid': Ngrid, 'step': step, 'Ncut': Ncut, 'Nmax': Nmax, 'N': N,
                      'nbar': nbar, 'kf': kf}
    return bispec


# names used by older drivers of the reference (run/_test_*.py)
FFTperiodic = FFT_periodic
Bk123_periodic = Bk_periodic
