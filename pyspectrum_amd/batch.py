"""Bispectra of many boxes of one shape with one launch per stage.

``Bk_periodic_batch`` is Bk_periodic (pyspectrum.py:285-356) for a list of
boxes that share (Ngrid, step, Ncut, Nmax, precision): every stage --
assignment, FFT, fcomb, shell passes, contraction, shell powers and the
normalisation -- runs once for up to 16 boxes (psx_bk_periodic_batch), and
each box's dict is bit-identical to Bk_periodic on that box.  The boxes may
differ in particle count, weights and Lbox.  The reference's campaign runs
the boxes one after another (run/glam/glam.py:98-116).
"""
import numpy as np
import torch

from . import _lib as L
from . import pyspectrum as P

MAX_BATCH = 16


def run_batch(ctx, xyzs, ws, Lboxes, Ngrid, step, Ncut, Nmax, precision, device, half=None):
    """One psx_bk_periodic_batch call (at most 16 boxes).  Returns the device
    outputs (out7 [nbox][7][ntri], p0k [nbox][Nmax]), the per-box weight sums
    and N_tri mask, and the half fields ([nbox][N][N][N/2+1])."""
    nbox = len(xyzs)
    if not 1 <= nbox <= MAX_BATCH:
        raise ValueError("a batch holds 1 to %d boxes" % MAX_BATCH)
    prec = P._prec(precision)
    dev = P._device(device)
    ijl = L.triangles(step, Ncut, Nmax)
    ntri = len(ijl)
    counts = P._counts_Bk123(Ngrid=Ngrid, Nmax=Nmax, Ncut=Ncut, step=step, device=device)
    tc = P._tri_counts(counts, ijl, Ngrid)
    boxes = (L.PsxBox * nbox)()
    keep = []
    sums = []
    for b in range(nbox):
        x = P._positions(xyzs[b], device)
        n = x.shape[1]
        wd, sw, sabs = P._weights(None if ws is None else ws[b], n, device)
        keep += [x, wd]
        Lb = float(Lboxes[b])
        kf = 2 * np.pi / Lb
        nbar = sw / Lb ** 3
        boxes[b] = L.PsxBox(x.data_ptr(), None if wd is None else wd.data_ptr(), n, Lb, sw, sabs,
                            nbar, kf ** 3, kf ** 6)
        sums.append((n, sw, nbar, kf))
    if half is None:
        half = torch.empty((nbox, Ngrid, Ngrid, Ngrid // 2 + 1), dtype=P._cdtype(prec), device=dev)
    cnt_d = torch.from_numpy(tc).to(dev)
    ijl_d = torch.from_numpy(np.ascontiguousarray(ijl)).to(dev)
    out7 = torch.empty((nbox, 7, max(ntri, 1)), dtype=torch.float64, device=dev)
    p0k = torch.empty((nbox, Nmax), dtype=torch.float64, device=dev)
    ctx.call("psx_bk_periodic_batch", nbox, boxes, Ngrid, step, Ncut, Nmax, L.ptr(cnt_d),
             L.ptr(ijl_d), (2 * np.pi) ** 3, (2 * np.pi) ** 6, L.ptr(out7), L.ptr(p0k),
             L.ptr(half), prec)
    return out7, p0k, half, sums, ijl, tc, keep


def Bk_periodic_batch(xyzs, ws=None, Lbox=2600., Ngrid=360, step=3, Ncut=3, Nmax=40,
                      precision='fp32', device=0):
    """Bk_periodic of every box in ``xyzs`` (a list of (3, N) arrays or
    tensors); ``ws`` a list of weights or None, ``Lbox`` a number or a list.
    Returns the list of Bk_periodic dicts."""
    P._check_ngrid(Ngrid)
    nb = len(xyzs)
    Lboxes = list(Lbox) if np.ndim(Lbox) else [Lbox] * nb
    if len(Lboxes) != nb or (ws is not None and len(ws) != nb):
        raise ValueError("one Lbox and one weight array (or None) per box")
    ctx = L.context(device)
    out = []
    for s in range(0, nb, MAX_BATCH):
        sl = slice(s, min(nb, s + MAX_BATCH))
        out7, _, _, sums, ijl, tc, _ = run_batch(ctx, xyzs[sl], None if ws is None else ws[sl],
                                                 Lboxes[sl], Ngrid, step, Ncut, Nmax, precision,
                                                 device)
        fin = out7.cpu().numpy()
        for b, (n, sw, nbar, kf) in enumerate(sums):
            bispec = P._assemble(ijl, step, fin[b][:, :len(ijl)], tc > 0)
            d = {k: bispec[k] for k in ('i_k1', 'i_k2', 'i_k3', 'p0k1', 'p0k2', 'p0k3', 'b123',
                                        'q123', 'counts')}
            d['meta'] = {'Lbox': Lboxes[s + b], 'Ngrid': Ngrid, 'step': step, 'Ncut': Ncut,
                         'Nmax': Nmax, 'N': n, 'nbar': nbar, 'kf': kf}
            d['p0k_sn'] = 1. / nbar
            d['b123_sn'] = bispec['b123_sn']
            out.append(d)
    return out
