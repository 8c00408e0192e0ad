"""pyspectrum_amd: pySpectrum's periodic-box P(k)/B(k) estimator on MI355X (gfx950).

``from pyspectrum_amd import pyspectrum as pySpec`` gives the reference's
periodic-box API (``Bk_periodic``, ``Pk_periodic``, ``Pk_periodic_rsd``,
``FFT_periodic``, ``reflect_delta``, ``_Bk_periodic``, ``_counts_Bk123``)
backed by the HIP library ``libpsx.so``.
"""
import os

_DAT_DIR = os.path.join(os.path.dirname(os.path.realpath(__file__)), 'dat')


def dat_dir():
    """Package data directory (pyspectrum/__init__.py:4-7)."""
    return _DAT_DIR
