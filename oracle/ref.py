"""ctypes binding to the reference's own Fortran kernels (oracle/_ref/libestimator.so).

TEST INFRASTRUCTURE ONLY.  The library is compiled by ``oracle/Makefile`` from
``pyspectrum/estimator.f`` where it lies in the reference checkout; it is used
to pin the numpy restatement in ``oracle.spectrum`` and never by the product.

Fortran calling convention: every argument by reference, symbol name with a
trailing underscore, arrays column-major (numpy order='F').  The argument
order is the Fortran one (estimator.f:155, :284, :514, :542, :574, :605), not
f2py's.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_ref", "libestimator.so")
_lib = None


def available():
    return os.path.isfile(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        if not available():
            raise OSError("reference Fortran library not built: run `make -C oracle`")
        # sfftw_* (bk_counts/ffting only) are unresolved: lazy binding
        _lib = ctypes.CDLL(LIB_PATH, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _i(v):
    return ctypes.byref(ctypes.c_int32(int(v)))


def _f(v):
    return ctypes.byref(ctypes.c_float(float(v)))


def assign_quad(xyz32, w, Ngrid, kf_ks, offset=0.0, iabcd=(0, 0, 0, 0)):
    """estimator.f:284 assign_quad(r, w, dtl, Np, Ngrid, kf_ks, offset, ia, ib, ic, id).

    Returns the interleaved float32 grid dtl(2*Ngrid, Ngrid, Ngrid) as an F-order
    array, plus its (A, B) views indexed [ix, iy, iz]."""
    r = np.asfortranarray(np.asarray(xyz32, dtype=np.float32))
    ww = np.asfortranarray(np.asarray(w, dtype=np.float32))
    dtl = np.zeros((2 * Ngrid, Ngrid, Ngrid), dtype=np.float32, order="F")
    ia, ib, ic, id_ = iabcd
    lib().assign_quad_(_p(r), _p(ww), _p(dtl), _i(r.shape[1]), _i(Ngrid), _f(kf_ks), _f(offset),
                       _i(ia), _i(ib), _i(ic), _i(id_))
    return dtl, dtl[0::2], dtl[1::2]


def _half(F, Ngrid):
    d = np.asfortranarray(np.asarray(F, dtype=np.complex64)).copy(order="F")
    assert d.shape == (Ngrid // 2 + 1, Ngrid, Ngrid)
    return d


def five_delta2g_1(xx, yy, zz, Ngrid):
    """estimator.f:514 FiveDelta2g_1(dcgxx, dcgyy, dcgzz, Ngrid); returns the new dcgxx."""
    x, y, z = _half(xx, Ngrid), _half(yy, Ngrid), _half(zz, Ngrid)
    lib().fivedelta2g_1_(_p(x), _p(y), _p(z), _i(Ngrid))
    return x


def five_delta2g_2(g, xx, xy, yz, zx, Ngrid):
    """estimator.f:542 FiveDelta2g_2(dcg, dcgxx, dcgxy, dcgyz, dcgzx, Ngrid); returns dcgxx."""
    a = [_half(v, Ngrid) for v in (g, xx, xy, yz, zx)]
    lib().fivedelta2g_2_(*[_p(v) for v in a], _i(Ngrid))
    return a[1]


def build_quad(d1, d2, irsd, Ngrid):
    """estimator.f:574 build_quad(dclr1, dclr2, irsd, Ngrid); returns the new dclr2."""
    a, b = _half(d1, Ngrid), _half(d2, Ngrid)
    assert irsd in (1, 2, 3)  # the Fortran STOPs otherwise
    lib().build_quad_(_p(a), _p(b), _i(irsd), _i(Ngrid))
    return b


def fcomb_periodic(F, sumw, Ngrid):
    """estimator.f:605 fcomb_periodic(dcl, N, Ngrid) on a complex64 F-order copy."""
    dcl = np.asfortranarray(np.asarray(F, dtype=np.complex64)).copy(order="F")
    lib().fcomb_periodic_(_p(dcl), _f(sumw), _i(Ngrid))
    return dcl


def fcomb_survey(F, Ngrid):
    """estimator.f:677 fcomb_survey(dcl, Ngrid) on a complex64 F-order copy."""
    dcl = np.asfortranarray(np.asarray(F, dtype=np.complex64)).copy(order="F")
    lib().fcomb_survey_(_p(dcl), _i(Ngrid))
    return dcl


def pk_pbox_rsd(dtl, irsd, Lbox, Nbin, Nmu, Ngrid):
    """estimator.f:155 pk_pbox_rsd(dtl,k,p0,p2,p4,nk,km,mk,pkm,nkm,irsd,Lbox,Nbin,Nmu,Ngrid)."""
    d = np.asfortranarray(np.asarray(dtl, dtype=np.complex64))
    assert d.shape == (Ngrid // 2 + 1, Ngrid, Ngrid)
    outs1 = [np.zeros(Nbin, dtype=np.float64) for _ in range(5)]
    outs2 = [np.zeros((Nbin, Nmu), dtype=np.float64, order="F") for _ in range(4)]
    k, p0, p2, p4, nk = outs1
    km, mk, pkm, nkm = outs2
    lib().pk_pbox_rsd_(_p(d), _p(k), _p(p0), _p(p2), _p(p4), _p(nk), _p(km), _p(mk), _p(pkm), _p(nkm),
                       _i(irsd), _i(int(Lbox)), _i(Nbin), _i(Nmu), _i(Ngrid))
    return k, p0, p2, p4, nk, km, mk, pkm, nkm
