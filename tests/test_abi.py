"""CPU tests of the C-ABI library: it loads, exports every symbol include/psx.h
declares, and its host-only logic (triangle list, argument checks) is right.
No kernel is launched here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import spectrum as S
from pyspectrum_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "psx.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(psx_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = _declared()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    # and the Python binding covers all of them
    assert set(names) == set(L.SIGNATURES) | {"psx_last_error"}


def test_abi_version():
    assert L.lib().psx_abi_version() == 2


@pytest.mark.parametrize("step,ncut,nmax", [(3, 3, 10), (3, 3, 40), (2, 2, 60), (1, 1, 5), (3, 3, 30)])
def test_triangle_list_matches_reference_loop(step, ncut, nmax):
    I, J, Lx, _ = S.triangles(nmax, ncut, step)
    got = L.triangles(step, ncut, nmax)
    assert got.shape == (len(I), 3)
    assert np.array_equal(got[:, 0], I) and np.array_equal(got[:, 1], J) and np.array_equal(got[:, 2], Lx)


def test_errors_are_reported_not_raised_across_the_abi():
    lib = L.lib()
    # null output pointer -> PSX_ERR_ARG with a message
    st = lib.psx_bk_num_triangles(3, 3, 40, None)
    assert st == 1
    assert b"null" in lib.psx_last_error()
    with pytest.raises(L.PsxError):
        L.call("psx_bk_triangles", 0, 3, 40, None)


def test_ctx_create_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    st = L.lib().psx_ctx_create(ctypes.byref(h), 0)
    assert st != 0
    assert len(L.lib().psx_last_error()) > 0


def test_product_path_has_no_cpu_fallback():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from pyspectrum_amd import pyspectrum as P
    with pytest.raises(L.PsxError):
        P.Pk_periodic(np.zeros((3, 10)), Lbox=100.0, Ngrid=16)


def test_product_package_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "pyspectrum_amd")
    for dp, _, fs in os.walk(pkg):
        for f in fs:
            if f.endswith((".py", ".hip", ".h")):
                text = open(os.path.join(dp, f)).read()
                assert "oracle" not in text.replace("the oracle", ""), f
