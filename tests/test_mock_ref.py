"""CPU checks of the numpy restatement the input-side GPU tests compare with
(tests/mockref.py): Philox4x32-10 against the published known-answer
vectors (Salmon et al., SC'11, Random123 kat_vectors), the uniforms'
range, and the sampler's placement of points inside their cells."""
import numpy as np

import mockref as M


def test_philox_known_answers():
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
            (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, want in kat:
        got = tuple(int(x) for x in M.philox4x32_10(ctr, key))
        assert got == want


def test_uniforms_and_normals():
    r = M.draw(12345, np.arange(200000), 0, 0)
    u = M.uniform53(r[0], r[1])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 5e-3
    z = M.normal(r)
    assert abs(z.mean()) < 1e-2 and abs(z.std() - 1.0) < 1e-2
    # extreme words: the largest uniform is 1 - 2^-53, the smallest 0
    top = np.array([0xFFFFFFFF], dtype=np.uint32)
    assert M.uniform53(top, top)[0] == 1.0 - 2.0 ** -53
    assert M.uniform53(top * 0, top * 0)[0] == 0.0


def test_sampler_places_points_in_weighted_cells():
    mesh, L = 8, 80.0
    g = np.full((mesh,) * 3, -50.0, dtype=np.float32)
    g[3, 5, 1] = 2.0           # one cell holds almost all the weight
    xyz = M.sample(g, 5000, L, 7)
    cells = np.floor(xyz.astype(np.float64) / (L / mesh)).astype(int)
    frac = np.mean((cells[0] == 3) & (cells[1] == 5) & (cells[2] == 1))
    assert frac > 0.999
    assert xyz.min() >= 0 and xyz.max() < L
