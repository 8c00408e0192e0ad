"""Host utilities of the reference's ``pyspectrum/util.py`` (float64 numpy).

* ``FlatLambdaCDM`` -- the background the reference takes from
  ``astropy.cosmology.FlatLambdaCDM(H0, Om0)`` (astropy's defaults: no
  radiation, no massive neutrinos), reduced to what the estimator uses:
  ``h``, ``efunc(z)`` and ``comoving_distance(z)`` in Mpc.  Any object with
  those attributes (an astropy cosmology included) can be passed instead.
* ``radecz_to_cartesian`` (util.py:27-51), ``ijl_order`` (util.py:8-24) and
  ``applyRSD`` (util.py:54-75).
* Text outputs of the campaign drivers (run/glam/glam.py:104-106, :127-133):
  ``write_pk_text``/``read_pk_text`` and ``write_bk_text``/``read_bk_text``
  with the same columns, formats and header lines.
"""
import numpy as np

C_KMS = 299792.458  # speed of light, km/s


class FlatLambdaCDM:
    """Flat LCDM: E(z) = sqrt(Om0 (1+z)^3 + 1 - Om0).

    comoving_distance(z) = (c/H0) int_0^z dz'/E(z') [Mpc], evaluated on a
    uniform grid over [0, max z] with 8-point Gauss-Legendre per interval
    (4096 intervals), cumulative sums at the nodes, and cubic Hermite
    interpolation between nodes with the exact derivative 1/E(z) (relative
    error below 1e-11 for z <~ 10).
    """

    NSEG = 4096

    def __init__(self, H0, Om0):
        self.H0 = float(H0)
        self.Om0 = float(Om0)
        self.h = self.H0 / 100.0

    def efunc(self, z):
        z = np.asarray(z, dtype=np.float64)
        return np.sqrt(self.Om0 * (1.0 + z) ** 3 + (1.0 - self.Om0))

    def _chi(self, z):
        """Dimensionless comoving distance int_0^z dz'/E(z')."""
        z = np.asarray(z, dtype=np.float64)
        if z.size == 0:
            return np.zeros_like(z)
        if np.any(z < 0) or not np.all(np.isfinite(z)):
            raise ValueError("redshifts must be finite and non-negative")
        zmax = float(np.max(z))
        if zmax == 0.0:
            return np.zeros_like(z)
        M = self.NSEG
        edges = np.linspace(0.0, zmax, M + 1)
        x, wq = np.polynomial.legendre.leggauss(8)
        a, b = edges[:-1], edges[1:]
        mid, hw = 0.5 * (a + b), 0.5 * (b - a)
        nodes = mid[:, None] + hw[:, None] * x[None, :]
        seg = (hw[:, None] * wq[None, :] / self.efunc(nodes)).sum(axis=1)
        F = np.concatenate([[0.0], np.cumsum(seg)])
        hd = (zmax / M) / self.efunc(edges)        # interval width x derivative at the nodes
        s = z * (M / zmax)                         # uniform grid: interval index directly
        k = np.minimum(s.astype(np.int64), M - 1)
        t = s - k
        t2 = t * t
        t3 = t2 * t
        return ((2 * t3 - 3 * t2 + 1) * F[k] + (t3 - 2 * t2 + t) * hd[k]
                + (-2 * t3 + 3 * t2) * F[k + 1] + (t3 - t2) * hd[k + 1])

    def comoving_distance(self, z):
        """Line-of-sight comoving distance in Mpc (float64 array or scalar)."""
        return (C_KMS / self.H0) * self._chi(z)


def _value(q):
    """Plain float64 array of a number, array or astropy Quantity."""
    return np.asarray(getattr(q, "value", q), dtype=np.float64)


def distance_table(cosmo, zmax, nseg=4096):
    """Comoving distance D [Mpc] of ``cosmo`` at nseg+1 uniform nodes on
    [0, zmax] and the node slopes times the spacing, (zmax/nseg) (c/H0)/E(z),
    for cubic Hermite interpolation (the device catalogue kernel)."""
    zmax = float(zmax)
    edges = np.linspace(0.0, zmax, nseg + 1)
    D = _value(cosmo.comoving_distance(edges))
    H0 = float(_value(cosmo.H0))
    slopes = (zmax / nseg) * (C_KMS / H0) / _value(cosmo.efunc(edges))
    return np.ascontiguousarray(D), np.ascontiguousarray(slopes)


def default_cosmology():
    """The reference's default (pyspectrum.py:85, :772): H0=67.6, Om0=0.31."""
    return FlatLambdaCDM(H0=67.6, Om0=0.31)


def radecz_to_cartesian(radecz, cosmo=None):
    """(RA, Dec, z) in degrees and redshift -> comoving (x, y, z) in Mpc/h
    (util.py:27-51).  Unlike the reference, the input array is not converted
    to radians in place."""
    radecz = np.asarray(radecz)
    if radecz.ndim < 1 or radecz.shape[0] != 3:
        raise AssertionError("radecz must be a (3, N) array of RA, Dec, z")
    cosmo = default_cosmology() if cosmo is None else cosmo
    ra, dec = np.radians(np.asarray(radecz[0], dtype=np.float64)), np.radians(
        np.asarray(radecz[1], dtype=np.float64))
    r = np.asarray(cosmo.comoving_distance(np.asarray(radecz[2], dtype=np.float64)) * cosmo.h,
                   dtype=np.float64)
    rxy = r * np.cos(dec)
    return np.stack([rxy * np.cos(ra), rxy * np.sin(ra), r * np.sin(dec)])


def ijl_order(i_k, j_k, l_k, typ='GM'):
    """Triangle indices sorted by (l, j, i) ascending (util.py:8-24).

    One row per distinct triangle, holding the positions where it occurs in
    increasing order (a single column when every triangle occurs once).
    """
    if typ != 'GM':
        raise NotImplementedError(typ)
    keys = np.stack([np.ravel(i_k), np.ravel(j_k), np.ravel(l_k)])
    if keys.shape[1] == 0:
        return np.zeros(0, dtype=np.int64)
    perm = np.lexsort(keys)  # last key primary: l, then j, then i; stable
    ks = keys[:, perm]
    first = np.r_[True, np.any(ks[:, 1:] != ks[:, :-1], axis=0)]
    sizes = np.diff(np.r_[np.flatnonzero(first), perm.size])
    if np.any(sizes != sizes[0]):
        raise ValueError("triangles repeat an unequal number of times")
    return perm.reshape(-1, int(sizes[0]))


_LOS_AXIS = {'x': 0, 'y': 1, 'z': 2}


def rsd_shift(redshift, h=0.7, omega0_m=0.3):
    """Position shift per unit velocity, (1+z)/(100 E(z)) [Mpc/h per km/s],
    as a numpy float64 scalar (so arrays it multiplies are promoted to
    float64, as with the cosmology object's E(z) in util.py:67-70)."""
    ez = FlatLambdaCDM(H0=100. * h, Om0=omega0_m).efunc(redshift)
    return np.float64((1 + redshift) / (100 * ez))


def applyRSD(xyz, vxyz, redshift, h=0.7, omega0_m=0.3, LOS=None, Lbox=None):
    """Redshift-space positions in a periodic box (util.py:54-75), with the
    reference's floating-point steps: the shift s v + Lbox (s = rsd_shift(z),
    float64) is added to the line-of-sight coordinate in float64 and the sum
    is rounded to the positions' dtype (numpy's in-place add, util.py:73);
    the wrap `% Lbox` then runs in that dtype (util.py:74).  Shapes are
    checked as in the reference; LOS and Lbox are required keywords."""
    pos, vel = np.asarray(xyz), np.asarray(vxyz)
    if pos.shape[0] != 3 or vel.shape[0] != 3:
        raise AssertionError("positions and velocities must be (3, N) arrays")
    if LOS is None:
        raise ValueError("LOS (line of sight: 'x', 'y' or 'z') is required")
    if Lbox is None:
        raise ValueError("Lbox (box size) is required")
    ax = _LOS_AXIS[LOS]
    out = np.array(pos, copy=True)
    moved = (pos[ax] + (rsd_shift(redshift, h, omega0_m) * vel[ax] + Lbox)).astype(out.dtype)
    out[ax] = np.remainder(moved, Lbox)
    return out


# ------------------------------------------------------------ text outputs --
PK_TEXT_FMT = '%.5e %.5e %.5e'
BK_TEXT_FMT = '%i %i %i %.5e %.5e %.5e %.5e %.5e %.5e %.5e'
BK_TEXT_KEYS = ('i_k1', 'i_k2', 'i_k3', 'p0k1', 'p0k2', 'p0k3', 'b123', 'q123', 'counts', 'b123_sn')


def write_pk_text(path, spec, Lbox, nbar):
    """P_0(k) table of run/glam/glam.py:104-106: columns k, p0k, counts."""
    kf = 2. * np.pi / Lbox
    hdr = ('pyspectrum P_l=0(k) calculation. Lbox=%.1f, k_f=%.5e, SN=%.5e' % (Lbox, kf, 1. / nbar))
    np.savetxt(path, np.array([spec['k'], spec['p0k'], spec['counts']]).T, fmt=PK_TEXT_FMT,
               delimiter='\t', header=hdr)


def read_pk_text(path):
    """(k, p0k, counts) from a write_pk_text file."""
    return np.loadtxt(path, skiprows=1, unpack=True, usecols=[0, 1, 2])


def write_bk_text(path, bispec, Lbox):
    """B(k) table of run/glam/glam.py:127-133: i_k, j_k, l_k, p0k1, p0k2, p0k3,
    b123, q123, counts, b123_sn (the i_k columns must be as long as the
    others, i.e. every triangle counted)."""
    cols = [np.asarray(bispec[k]) for k in BK_TEXT_KEYS]
    if len({len(c) for c in cols}) != 1:
        raise ValueError("triangle columns of unequal length (a triangle with N_tri = 0)")
    hdr = 'pyspectrum bispectrum calculation test. k_f = 2pi/%.1f' % Lbox
    np.savetxt(path, np.array(cols).T, fmt=BK_TEXT_FMT, delimiter='\t', header=hdr)


def read_bk_text(path):
    """Dict of the ten write_bk_text columns (i_k as int64)."""
    cols = np.loadtxt(path, skiprows=1, unpack=True, usecols=range(10), ndmin=2)
    out = dict(zip(BK_TEXT_KEYS, cols))
    for k in ('i_k1', 'i_k2', 'i_k3'):
        out[k] = out[k].astype(np.int64)
    return out
