"""Regenerate the committed golden fixtures from the data files the reference ships.

Run in the build container (it reads ``/root/reference``, which the GPU host does
not have):

    /opt/conda/bin/python3.9 tests/golden/make_golden.py

(the hdf5 part needs h5py, which only that interpreter has; the counts part runs
under any python with scipy).

Outputs, all plain data:

* ``ntri_Ngrid360_Ncut3_step3_Nmax{10,40,50}.npz`` -- the reference's FP64
  triangle-count caches ``pyspectrum/dat/counts.Ngrid360.Nmax*.Ncut3.step3.pyfftw``
  (FortranFile: one record of Nmax**3 float64, C order, read by
  ``pyspectrum/pyspectrum.py:969-973``), reduced to the entries with i>=j>=l:
  arrays ``i, j, l`` (1-based shell indices as in ``counts[i-1, j-1, l-1]``),
  ``counts`` (the raw float64 value = Ngrid**3 * N_tri + FFT round-off) and
  ``ntri = rint(counts / Ngrid**3)``.
* ``test_box.npz`` -- ``pyspectrum/dat/test_box.hdf5`` (73,442 halos; the input
  of ``tests/test_pyspec.py:14-16``) as float64 ``xyz`` and ``vxyz`` of shape (N, 3).
"""
import os
import sys

import numpy as np

REF_DAT = "/root/reference/pyspectrum/dat"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_fortran_record_f64(path):
    """One unformatted sequential record: int32 length, payload, int32 length."""
    raw = open(path, "rb").read()
    n0 = np.frombuffer(raw[:4], dtype="<i4")[0]
    n1 = np.frombuffer(raw[-4:], dtype="<i4")[0]
    assert n0 == n1 == len(raw) - 8, (path, n0, n1, len(raw))
    return np.frombuffer(raw[4:-4], dtype="<f8").copy()


def make_counts():
    G = 360 ** 3
    for nmax in (10, 40, 50):
        path = os.path.join(REF_DAT, "counts.Ngrid360.Nmax%d.Ncut3.step3.pyfftw" % nmax)
        vals = read_fortran_record_f64(path).reshape(nmax, nmax, nmax)
        ii, jj, ll = np.nonzero(np.ones_like(vals, dtype=bool))
        keep = (ii >= jj) & (jj >= ll)
        ii, jj, ll = ii[keep], jj[keep], ll[keep]
        c = vals[ii, jj, ll]
        # entries off the i>=j>=l simplex are all exactly zero in the file
        off = np.ones_like(vals, dtype=bool)
        off[ii, jj, ll] = False
        assert np.all(vals[off] == 0.0)
        out = os.path.join(HERE, "ntri_Ngrid360_Ncut3_step3_Nmax%d.npz" % nmax)
        np.savez_compressed(out, i=(ii + 1).astype(np.int32), j=(jj + 1).astype(np.int32),
                            l=(ll + 1).astype(np.int32), counts=c,
                            ntri=np.rint(c / G).astype(np.int64))
        print(out, "entries", len(c), "nonzero", int(np.sum(np.rint(c / G) != 0)),
              "max |counts/G - rint|", float(np.max(np.abs(c / G - np.rint(c / G)))))


def make_box():
    try:
        import h5py
    except ImportError:
        print("h5py not importable here; skipping test_box.npz", file=sys.stderr)
        return
    with h5py.File(os.path.join(REF_DAT, "test_box.hdf5"), "r") as f:
        xyz = f["xyz"][...]
        vxyz = f["vxyz"][...]
    out = os.path.join(HERE, "test_box.npz")
    np.savez_compressed(out, xyz=xyz.astype(np.float64), vxyz=vxyz.astype(np.float64))
    print(out, xyz.shape, vxyz.shape)


if __name__ == "__main__":
    make_counts()
    make_box()
