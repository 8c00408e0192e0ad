"""Quadrupole machinery of the reference (its unfinished survey quadrupole).

The reference's Python side of this (pyspectrum.py:828-900 and the helpers
``_FiveDelta2g_1``/``_FiveDelta2g_2`` at :1083-1131) cannot run: it calls
``assign_quad`` with the wrong arguments, a nonexistent ``fEstimate.fcomb`` and
names that are never defined.  What it was meant to call is the Fortran, and
that is what this module provides, on the device:

=================  ==========================  ======================
function           reference                   entry point
=================  ==========================  ======================
quad_weights       estimator.f:291-299         psx_quad_weights
FiveDelta2g_1      estimator.f:514-540         psx_five_delta2g_1
FiveDelta2g_2      estimator.f:542-572         psx_five_delta2g_2
build_quad         estimator.f:574-604         psx_build_quad
=================  ==========================  ======================

Half fields are the Fortran arrays (Ngrid//2+1, Ngrid, Ngrid) [ix, iy, iz]:
numpy arrays of that shape (complex64 or complex128) are updated in place,
as f2py's intent(inout) does, or torch CUDA tensors in this library's
layout (Ngrid, Ngrid, Ngrid//2+1) [iz, iy, ix] are used in place.

``unit_vectors=False`` reproduces the Fortran exactly: FiveDelta2g_1/2 never
declare kxh, kyh, kzh, so they are INTEGER and float(k)/|k| is truncated
(+-1 on the axes, 0 elsewhere; the cross terms of FiveDelta2g_2 vanish).
``unit_vectors=True`` uses k_i/|k| as Scoccimarro (2015) eq. 20 intends.
"""
import numpy as np
import torch

from . import _lib as L

__all__ = ["quad_weights", "FiveDelta2g_1", "FiveDelta2g_2", "build_quad"]


def _half_arg(a, Ngrid, device):
    """(device tensor [iz, iy, ix], precision, write-back function)."""
    Nh = Ngrid // 2 + 1
    if isinstance(a, torch.Tensor):
        if a.device.type != "cuda" or tuple(a.shape) != (Ngrid, Ngrid, Nh) or not a.is_contiguous():
            raise ValueError("a torch half field must be a contiguous CUDA tensor (Ngrid, Ngrid, Ngrid//2+1)")
        if a.dtype not in (torch.complex64, torch.complex128):
            raise ValueError("half fields are complex64 or complex128")
        prec = L.PREC["fp64"] if a.dtype == torch.complex128 else L.PREC["fp32"]
        return a, prec, lambda: None
    arr = np.asarray(a)
    if arr.shape != (Nh, Ngrid, Ngrid):
        raise ValueError("a numpy half field must have shape (Ngrid//2+1, Ngrid, Ngrid)")
    if arr.dtype not in (np.complex64, np.complex128):
        raise ValueError("half fields are complex64 or complex128")
    prec = L.PREC["fp64"] if arr.dtype == np.complex128 else L.PREC["fp32"]
    t = torch.from_numpy(np.ascontiguousarray(arr.transpose(2, 1, 0))).to(torch.device("cuda", device))

    def back():
        arr[...] = t.cpu().numpy().transpose(2, 1, 0)
    return t, prec, back


def _same_precision(precs):
    if len(set(precs)) != 1:
        raise ValueError("all half fields must have the same dtype")
    return precs[0]


def quad_weights(xyz, w=None, ia=0, ib=0, ic=0, id=0, device=0):
    """Particle weights of assign_quad (estimator.f:291-299), float32:
    w (all indices 0), w r_ia r_ib / |r|^2 (ic = id = 0) or
    w r_ia r_ib r_ic r_id / |r|^4, components 1..3 = x, y, z.  Returns a
    CUDA tensor when ``xyz`` is one, else a numpy array; pass it as ``w`` of
    FFT_periodic / psx_assign_quad for the Q_ij fields."""
    dev = torch.device("cuda", device)
    host = not isinstance(xyz, torch.Tensor)
    x = (torch.from_numpy(np.ascontiguousarray(np.asarray(xyz, np.float32))) if host
         else xyz.to(dtype=torch.float32)).to(dev).contiguous()
    if x.dim() != 2 or x.shape[0] != 3:
        raise ValueError("xyz must have shape (3, N)")
    n = x.shape[1]
    wt = None
    if w is not None:
        wt = (w if isinstance(w, torch.Tensor) else torch.from_numpy(np.asarray(w, np.float32)))
        wt = wt.to(device=dev, dtype=torch.float32).contiguous()
        if wt.numel() != n:
            raise ValueError("w must have N entries")
    out = torch.empty(n, dtype=torch.float32, device=dev)
    L.context(device).call("psx_quad_weights", L.ptr(x), L.ptr(wt), n, ia, ib, ic, id, L.ptr(out))
    return out.cpu().numpy() if host else out


def FiveDelta2g_1(dcgxx, dcgyy, dcgzz, Ngrid, unit_vectors=False, device=0):
    """dcgxx = 7.5 (dcgxx kxh^2 + dcgyy kyh^2 + dcgzz kzh^2), in place
    (estimator.f:514-540); returns dcgxx."""
    args = [_half_arg(a, Ngrid, device) for a in (dcgxx, dcgyy, dcgzz)]
    prec = _same_precision([p for _, p, _ in args])
    (xx, _, back), (yy, _, _), (zz, _, _) = args
    L.context(device).call("psx_five_delta2g_1", L.ptr(xx), L.ptr(yy), L.ptr(zz), Ngrid,
                           1 if unit_vectors else 0, prec)
    back()
    return dcgxx


def FiveDelta2g_2(dcg, dcgxx, dcgxy, dcgyz, dcgzx, Ngrid, unit_vectors=False, device=0):
    """dcgxx = dcgxx + 7.5 (2 dcgxy kxh kyh + 2 dcgyz kyh kzh + 2 dcgzx kzh kxh)
    - 2.5 dcg, in place (estimator.f:542-572); returns dcgxx."""
    args = [_half_arg(a, Ngrid, device) for a in (dcg, dcgxx, dcgxy, dcgyz, dcgzx)]
    prec = _same_precision([p for _, p, _ in args])
    t = [a for a, _, _ in args]
    L.context(device).call("psx_five_delta2g_2", *[L.ptr(v) for v in t], Ngrid,
                           1 if unit_vectors else 0, prec)
    args[1][2]()
    return dcgxx


def build_quad(dclr1, dclr2, irsd, Ngrid, device=0):
    """dclr2 = (7.5 mu^2 - 2.5) dclr1 with mu = k_los/|k|, line of sight
    irsd = 1, 2, 3 (x, y, z), in place (estimator.f:574-604); returns dclr2.
    The Fortran STOPs for any other irsd; this raises."""
    if irsd not in (1, 2, 3):
        raise ValueError("irsd must be 1, 2 or 3")
    (d1, p1, _), (d2, p2, back) = (_half_arg(a, Ngrid, device) for a in (dclr1, dclr2))
    prec = _same_precision([p1, p2])
    L.context(device).call("psx_build_quad", L.ptr(d1), L.ptr(d2), irsd, Ngrid, prec)
    back()
    return dclr2
