"""HEALPix RING-scheme pixel geometry on the host: face coordinates, RING <-> NEST, pixel centres
and the 8-neighbour table (the roles of healpy.pix2ang / ring2nest / nest2ring /
get_all_neighbours, which the reference calls at weighting.py:119,140 for inpainting and which
the learned preconditioner's 9-point stencil uses).

Built from the HEALPix construction (Gorski et al. 2005, ApJ 622, 759): 12 base faces, face f
with centre ring jrll[f] Nside and centre longitude jpll[f] pi/4, pixel (ix, iy) of a face at
continuous coordinates x = (ix + 1/2) / Nside, y = (iy + 1/2) / Nside, where x + y grows towards
the face's north corner and x - y towards its east corner.  A point (x, y) of face f lies at
    t = jrll[f] - x - y,
    z = 1 - t^2 / 3, phi = (jpll[f] + (x - y) / t) pi / 4       (t < 1, north cap)
    z = (2 - t) 2 / 3, phi = (jpll[f] + x - y) pi / 4           (1 <= t <= 3)
    z = -(1 - (4 - t)^2 / 3), phi = (jpll[f] + (x - y) / (4 - t)) pi / 4   (t > 3)
and the RING index of a pixel follows from its ring t Nside and its position along the ring.

Neighbours: the eight (dx, dy) steps in the order of healpy's get_all_neighbours (SW, W, NW, N,
NE, E, SE, S), continued across face edges by the edge identifications of the 12 faces; a step
through a vertex where only three faces meet (the four corners at z = +-2/3 of each polar face)
has no pixel, -1 as in healpy.
"""
import numpy as np

JRLL = np.array([2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4])
JPLL = np.array([1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7])
# (dx, dy) of SW, W, NW, N, NE, E, SE, S (north corner of a face at +x +y, east at +x -y)
DX = np.array([-1, -1, 0, 1, 1, 1, 0, -1])
DY = np.array([0, 1, 1, 1, 0, -1, -1, -1])


def xyf2ring(nside, f, ix, iy):
    """RING index of pixel (ix, iy) of face f (arrays)."""
    n = int(nside)
    f, ix, iy = np.asarray(f, np.int64), np.asarray(ix, np.int64), np.asarray(iy, np.int64)
    ncap = 2 * n * (n - 1)
    npix = 12 * n * n
    jr = JRLL[f] * n - ix - iy - 1  # ring number 1 .. 4n - 1
    nr = np.where(jr < n, jr, np.where(jr > 3 * n, 4 * n - jr, n))
    before = np.where(jr < n, 2 * jr * (jr - 1),
                      np.where(jr > 3 * n, npix - 2 * nr * (nr + 1), ncap + (jr - n) * 4 * n))
    kshift = np.where((jr >= n) & (jr <= 3 * n), (jr - n) & 1, 0)
    jp = (JPLL[f] * nr + ix - iy + 1 + kshift) // 2
    jp = np.where(jp > 4 * nr, jp - 4 * nr, jp)
    jp = np.where(jp < 1, jp + 4 * nr, jp)
    return before + jp - 1


_XYF = {}


def ring2xyf(nside):
    """(face, ix, iy) of every RING pixel, each (npix,) int64 (cached per Nside)."""
    n = int(nside)
    if n not in _XYF:
        f = np.repeat(np.arange(12), n * n)
        ix = np.tile(np.repeat(np.arange(n), n), 12)
        iy = np.tile(np.arange(n), 12 * n)
        ring = xyf2ring(n, f, ix, iy)
        out = [np.empty(12 * n * n, np.int64) for _ in range(3)]
        for o, v in zip(out, (f, ix, iy)):
            o[ring] = v
        _XYF[n] = tuple(out)
    return _XYF[n]


def _spread(v):
    """Bits of v at the even positions of the result."""
    v = np.asarray(v, np.int64)
    out = np.zeros_like(v)
    for b in range(31):
        out |= ((v >> b) & 1) << (2 * b)
    return out


def _compress(v):
    v = np.asarray(v, np.int64)
    out = np.zeros_like(v)
    for b in range(31):
        out |= ((v >> (2 * b)) & 1) << b
    return out


def ring2nest(nside, pix):
    """NEST index of RING pixels (NEST: face f n^2 + ix on the even bits + iy on the odd bits)."""
    n = int(nside)
    assert n & (n - 1) == 0, "NEST ordering needs Nside a power of two"
    f, ix, iy = (a[np.asarray(pix)] for a in ring2xyf(n))
    return f * n * n + _spread(ix) + (_spread(iy) << 1)


def nest2ring(nside, pix):
    n = int(nside)
    assert n & (n - 1) == 0, "NEST ordering needs Nside a power of two"
    pix = np.asarray(pix, np.int64)
    f, r = pix // (n * n), pix % (n * n)
    return xyf2ring(n, f, _compress(r), _compress(r >> 1))


def face_point(f, x, y):
    """(z, phi) of the continuous face coordinates (x, y) in [0, 1]^2 of face f."""
    f = np.asarray(f)
    x, y = np.asarray(x, np.float64), np.asarray(y, np.float64)
    t = JRLL[f] - x - y
    tn = np.where(t < 1, t, np.where(t > 3, 4 - t, 1.))
    with np.errstate(divide='ignore', invalid='ignore'):
        dphi = np.where(tn > 0, (x - y) / tn, 0.)
    z = np.where(t < 1, 1 - t * t / 3, np.where(t > 3, -(1 - (4 - t) ** 2 / 3), (2 - t) * 2. / 3))
    phi = (JPLL[f] + np.where((t < 1) | (t > 3), dphi, x - y)) * np.pi / 4
    return z, np.mod(phi, 2 * np.pi)


def pix2ang(nside, pix):
    """(theta, phi) of RING pixel centres."""
    n = int(nside)
    f, ix, iy = (a[np.asarray(pix)] for a in ring2xyf(n))
    z, phi = face_point(f, (ix + 0.5) / n, (iy + 0.5) / n)
    return np.arccos(np.clip(z, -1, 1)), phi


def _step(n, f, nx, ny):
    """Face and coordinates of the pixel at (nx, ny) of face f's coordinate frame, one step
    outside the face at most; face -1 where the step passes a three-face vertex."""
    xo = np.where(nx < 0, -1, np.where(nx >= n, 1, 0))
    yo = np.where(ny < 0, -1, np.where(ny >= n, 1, 0))
    band = f // 4  # 0 north, 1 equatorial, 2 south
    k = f % 4
    F, X, Y = f.copy(), nx.copy(), ny.copy()

    def put(sel, ff, xx, yy):
        F[sel], X[sel], Y[sel] = ff[sel], xx[sel], yy[sel]

    m = lambda a: np.mod(a, 4)
    north, eq, south = band == 0, band == 1, band == 2
    # edge steps
    put(north & (xo == 1) & (yo == 0), m(k + 1), ny, 2 * n - 1 - nx)
    put(north & (yo == 1) & (xo == 0), m(k + 3), 2 * n - 1 - ny, nx)
    put(north & (xo == -1) & (yo == 0), 4 + k, nx + n, ny)
    put(north & (yo == -1) & (xo == 0), 4 + m(k + 1), nx, ny + n)
    put(eq & (xo == 1) & (yo == 0), k, nx - n, ny)
    put(eq & (yo == 1) & (xo == 0), m(k + 3), nx, ny - n)
    put(eq & (xo == -1) & (yo == 0), 8 + m(k + 3), nx + n, ny)
    put(eq & (yo == -1) & (xo == 0), 8 + k, nx, ny + n)
    put(south & (xo == 1) & (yo == 0), 4 + m(k + 1), nx - n, ny)
    put(south & (yo == 1) & (xo == 0), 4 + k, nx, ny - n)
    put(south & (yo == -1) & (xo == 0), 8 + m(k + 1), -ny - 1, nx)
    put(south & (xo == -1) & (yo == 0), 8 + m(k + 3), ny, -nx - 1)
    # vertex steps: four faces meet at the poles and on the equator, three at z = +-2/3
    put(north & (xo == 1) & (yo == 1), m(k + 2), 2 * n - 1 - nx, 2 * n - 1 - ny)
    put(north & (xo == -1) & (yo == -1), 8 + k, nx + n, ny + n)
    put(eq & (xo == 1) & (yo == -1), 4 + m(k + 1), nx - n, ny + n)
    put(eq & (xo == -1) & (yo == 1), 4 + m(k + 3), nx + n, ny - n)
    put(south & (xo == 1) & (yo == 1), k, nx - n, ny - n)
    put(south & (xo == -1) & (yo == -1), 8 + m(k + 2), -nx - 1, -ny - 1)
    three = ((north | south) & (xo * yo == -1)) | (eq & (xo * yo == 1))
    F[three] = -1
    return F, X, Y


def get_all_neighbours(nside, pix):
    """The 8 neighbours (SW, W, NW, N, NE, E, SE, S) of RING pixels, (8, len(pix)), -1 where a
    pixel has only seven."""
    n = int(nside)
    pix = np.atleast_1d(np.asarray(pix, np.int64))
    f, ix, iy = (a[pix] for a in ring2xyf(n))
    out = np.empty((8, len(pix)), np.int64)
    for d in range(8):
        F, X, Y = _step(n, f, ix + DX[d], iy + DY[d])
        ok = F >= 0
        r = np.full(len(pix), -1, np.int64)
        r[ok] = xyf2ring(n, F[ok], X[ok], Y[ok])
        out[d] = r
    return out


def stencil9(nside):
    """9-point stencil of every RING pixel: itself, then its 8 neighbours (SW, W, NW, N, NE, E,
    SE, S; -1 where missing).  Returns (nbr (npix, 9), rev (npix, rw)) int32, rev listing
    p * 9 + k for every slot (p, k) that reads the pixel (-1 padding)."""
    n = int(nside)
    npix = 12 * n * n
    pix = np.arange(npix)
    nbr = np.empty((npix, 9), np.int64)
    nbr[:, 0] = pix
    nbr[:, 1:] = get_all_neighbours(n, pix).T
    flat = nbr.reshape(-1)
    ok = flat >= 0
    src = np.nonzero(ok)[0]
    dst = flat[ok]
    order = np.argsort(dst, kind='stable')
    dst_s, src_s = dst[order], src[order]
    cnt = np.bincount(dst_s, minlength=npix)
    rw = int(cnt.max())
    rev = -np.ones((npix, rw), np.int64)
    first = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    rev[dst_s, np.arange(len(dst_s)) - first[dst_s]] = src_s
    return nbr.astype(np.int32), rev.astype(np.int32)
