"""numpy restatement of the input-side kernels (psx_mock.hip) for the tests:
Philox4x32-10, the 53-bit uniforms, Box-Muller normals, the Gaussian field
and the inverse-CDF sampler, in float64 where the kernels use float64."""
import numpy as np

_MASK = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
    """Philox4x32-10 (Salmon et al., SC'11) on arrays: ctr = 4 uint32 words
    (arrays or scalars), key = 2 uint32 words.  Returns 4 uint32 arrays."""
    c = [np.asarray(x, dtype=np.uint64) for x in ctr]
    k0, k1 = np.uint64(key[0]), np.uint64(key[1])
    for r in range(10):
        if r:
            k0 = (k0 + np.uint64(0x9E3779B9)) & _MASK
            k1 = (k1 + np.uint64(0xBB67AE85)) & _MASK
        p0 = np.uint64(0xD2511F53) * c[0]
        p1 = np.uint64(0xCD9E8D57) * c[2]
        c = [((p1 >> np.uint64(32)) ^ c[1] ^ k0) & _MASK, p1 & _MASK,
             ((p0 >> np.uint64(32)) ^ c[3] ^ k1) & _MASK, p0 & _MASK]
    return [x.astype(np.uint32) for x in c]


def draw(seed, idx, word, stream):
    idx = np.asarray(idx, dtype=np.uint64)
    ctr = (idx & _MASK, idx >> np.uint64(32), np.full(idx.shape, word, np.uint64),
           np.full(idx.shape, stream, np.uint64))
    return philox4x32_10(ctr, (seed & 0xFFFFFFFF, seed >> 32))


def uniform53(a, b):
    return ((a >> np.uint32(5)).astype(np.float64) * 67108864.0
            + (b >> np.uint32(6)).astype(np.float64)) * (1.0 / 9007199254740992.0)


def normal(r):
    u1, u2 = uniform53(r[0], r[1]), uniform53(r[2], r[3])
    return np.sqrt(-2.0 * np.log(1.0 - u1)) * np.cos(6.283185307179586 * u2)


def gaussian_field(mesh, Lbox, seed, slope=-1.5, kcut=0.5):
    """Unit-variance field (float64) and the draw's standard deviation."""
    n = mesh ** 3
    w = normal(draw(seed, np.arange(n), 0, 0)).astype(np.float32).reshape((mesh,) * 3)
    W = np.fft.fftn(w.astype(np.float64))
    f = np.fft.fftfreq(mesh, 1.0 / mesh)
    k = 2 * np.pi / Lbox * np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2
                                   + f[None, None, :] ** 2)
    with np.errstate(divide="ignore"):
        amp = np.where(k > 0, k ** (0.5 * slope) * np.exp(-0.5 * (k / kcut) ** 2), 0.0)
    g = np.fft.ifftn(W * amp.astype(np.float32)).real * n
    sig = g.std(ddof=1)
    return g / sig, sig


def sample(g32, n, Lbox, seed):
    """Points (float32 (3, n)) drawn from exp(g - 1/2) of a float32 field."""
    mesh = g32.shape[0]
    cdf = np.cumsum(np.exp(g32.astype(np.float64).ravel() - 0.5))
    i = np.arange(n)
    r0, r1 = draw(seed, i, 0, 1), draw(seed, i, 1, 1)
    target = uniform53(r0[0], r0[1]) * cdf[-1]
    cell = np.minimum(np.searchsorted(cdf, target, side="left"), mesh ** 3 - 1)
    idx = (cell // mesh ** 2, (cell // mesh) % mesh, cell % mesh)
    off = (uniform53(r0[2], r0[3]), uniform53(r1[0], r1[1]), uniform53(r1[2], r1[3]))
    h = Lbox / mesh
    return np.array([np.fmod((idx[a] + off[a]) * h, Lbox) for a in range(3)]).astype(np.float32)


def velocities(n, seed, sigma):
    i = np.arange(n)
    return np.array([sigma * normal(draw(seed, i, a, 2)) for a in range(3)]).astype(np.float32)
