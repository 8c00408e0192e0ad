"""CPU tests of the contraction's host-side triangle planning
(pyspectrum_amd/csrc/psx_plan.h): the triangle order and the shell-pair
cover that decides which pair's products read each triangle.

A small C++ driver is compiled against the header with g++ and prints, per
triangle, (i, j, l) and the chosen (pair, remaining shell).  Checked:
  * the triangle list is the reference loop (pyspectrum.py:415-417), as the
    oracle restates it;
  * every triangle is read by exactly one of its own pairs, with the third
    shell as the column (the multiset {a, b, c} equals {i, j, l});
  * the cover is deterministic (two runs print the same plan);
  * the local search never ends above the greedy cover it starts from
    (339 pairs at Nmax 40 / step 3, 734 at Nmax 60 / step 2).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import spectrum as S

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "pyspectrum_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "psx_plan.h"
int main(int argc, char** argv) {
  const int step = atoi(argv[1]), ncut = atoi(argv[2]), nmax = atoi(argv[3]);
  std::vector<int> ijl, tp, ts;
  psx::triangle_list(step, ncut, nmax, &ijl);
  psx::triangle_cover(ijl, nmax, ncut / step, &tp, &ts);
  for (size_t t = 0; t < tp.size(); ++t)
    printf("%d %d %d %d %d\n", ijl[3 * t], ijl[3 * t + 1], ijl[3 * t + 2], tp[t], ts[t]);
  return 0;
}
"""

needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("plan")
    src = d / "plan_driver.cpp"
    src.write_text(DRIVER)
    exe = d / "plan_driver"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, str(src), "-o", str(exe)], check=True)
    return str(exe)


def run_plan(exe, step, ncut, nmax):
    out = subprocess.run([exe, str(step), str(ncut), str(nmax)], check=True,
                         capture_output=True, text=True).stdout
    return np.array([[int(v) for v in line.split()] for line in out.splitlines()], dtype=np.int64)


@needs_gxx
@pytest.mark.parametrize("step,ncut,nmax,greedy_pairs", [
    (3, 3, 40, 339), (2, 2, 60, 734), (3, 3, 12, None), (1, 0, 5, None), (3, 3, 46, None),
])
def test_cover_reads_every_triangle_once(driver, step, ncut, nmax, greedy_pairs):
    p = run_plan(driver, step, ncut, nmax)
    i_, j_, l_, _ = S.triangles(nmax, ncut, step)
    assert p.shape[0] == len(i_)
    np.testing.assert_array_equal(p[:, 0], i_)
    np.testing.assert_array_equal(p[:, 1], j_)
    np.testing.assert_array_equal(p[:, 2], l_)
    nsh = nmax + 1
    a, b, c = p[:, 3] // nsh, p[:, 3] % nsh, p[:, 4]
    assert np.all(p[:, 3] >= 0)
    assert np.all(a >= b)
    np.testing.assert_array_equal(np.sort(np.stack([a, b, c], 1), 1),
                                  np.sort(p[:, :3], 1))
    npairs = len(np.unique(p[:, 3]))
    if greedy_pairs is not None:
        assert npairs <= greedy_pairs


@needs_gxx
def test_cover_is_deterministic(driver):
    np.testing.assert_array_equal(run_plan(driver, 3, 3, 40), run_plan(driver, 3, 3, 40))
