"""Host-side shape helpers used when building scenes (circumscribed radius and
volume of superquadrics, random convex hulls).  The kernels compute their own
mass properties on the host side of the C-ABI; these helpers only size
lattices and pick damping rates."""
import math

import numpy as np


def _beta(x, y):
    return math.exp(math.lgamma(x) + math.lgamma(y) - math.lgamma(x + y))


def superquadric_volume(a, b, c, n1, n2):
    """Superquadric<T>::computeVolume (Grains/Geometry/src/Superquadric.cpp:81-90)."""
    e1, e2 = 2.0 / n1, 2.0 / n2
    return 2.0 / 3.0 * a * b * c * e1 * e2 * _beta(e1, 0.5 * e1) * _beta(0.5 * e2, 0.5 * e2)


def superquadric_volume_vec(prm):
    """superquadric_volume over rows (a, b, c, n1, n2) of an array."""
    from scipy.special import betaln
    prm = np.asarray(prm, float)
    a, b, c, n1, n2 = prm.T
    e1, e2 = 2.0 / n1, 2.0 / n2
    return 2.0 / 3.0 * a * b * c * e1 * e2 * np.exp(betaln(e1, 0.5 * e1) + betaln(0.5 * e2, 0.5 * e2))


def superquadric_rc(a, b, c, n1, n2):
    """Superquadric<T>::computeCircumscribedRadius (Superquadric.cpp:122-155)."""
    if n1 == 2.0 and n2 == 2.0:
        return max(a, b, c)
    if n1 == 2.0:
        alpha = (b / a) ** (2.0 / (n2 - 2.0))
        xt = 1.0 / (1.0 + alpha ** n2) ** (1.0 / n2)
        return max(c, math.sqrt(a * xt * a * xt + alpha * b * xt * alpha * b * xt))
    if n2 == 2.0:
        m = max(a, b)
        beta = (c * c / (m * m)) ** (1.0 / (n1 - 2.0))
        xt = 1.0 / (1.0 + beta ** n1) ** (1.0 / n1)
        return math.sqrt(m * xt * m * xt + beta * c * xt * beta * c * xt)
    alpha = (b / a) ** (2.0 / (n2 - 2.0))
    gamma = (1.0 + alpha ** n2) ** (n1 / n2 - 1.0)
    beta = (gamma * c * c / (a * a)) ** (1.0 / (n1 - 1.0))
    xt = 1.0 / ((1.0 + alpha ** n2) ** (n1 / n2) + beta ** n1) ** (1.0 / n1)
    return math.sqrt(a * xt * a * xt + alpha * b * xt * alpha * b * xt + beta * c * xt * beta * c * xt)


def _planar_max(a, b, n):
    r = max(a, b)
    if n > 2.0:
        with np.errstate(all="ignore"):
            t = np.float64(b / a) ** (n / (n - 2.0))
            x = a * (1.0 + (t * a / b) ** n) ** (-1.0 / n)
            c = np.sqrt(x * x + t * x * t * x)
        if np.isfinite(c):
            r = max(r, float(c))
    return r


def superquadric_rc_exact(a, b, c, n1, n2):
    """Largest |p| on the superquadric: maximum over the critical points of
    |p|^2 of every coordinate family (axes, coordinate planes, interior)."""
    r = max(_planar_max(a, b, n2), _planar_max(a, c, n1), _planar_max(b, c, n1))
    if n1 > 2.0 and n2 > 2.0:
        with np.errstate(all="ignore"):
            al = np.float64(b / a) ** (2.0 / (n2 - 2.0))
            T = 1.0 + al ** n2
            be = (c * c / (a * a) * T ** (n1 / n2 - 1.0)) ** (1.0 / (n1 - 2.0))
            xt = (T ** (n1 / n2) + be ** n1) ** (-1.0 / n1)
            q = np.sqrt((a * xt) ** 2 + (al * b * xt) ** 2 + (be * c * xt) ** 2)
        if np.isfinite(q):
            r = max(r, float(q))
    return r


def random_hull_points(rng, n_points=8, radius=1.0):
    """Points sampled uniformly on a sphere, shifted so that the volume
    centroid of their convex hull is the origin (bodies rotate about it)."""
    p = rng.standard_normal((n_points, 3))
    p /= np.linalg.norm(p, axis=1, keepdims=True)
    p *= radius
    c = hull_centroid(p)
    return p - c


def random_hull_points_batch(rng, n_hulls, n_points=8):
    """n_hulls sets of n_points points uniform on the unit sphere, each
    shifted so that the volume centroid of its convex hull is the origin:
    (n_hulls, n_points, 3)."""
    p = rng.standard_normal((n_hulls, n_points, 3))
    p /= np.linalg.norm(p, axis=2, keepdims=True)
    return p - hull_centroid_batch(p)[:, None, :]


def hull_centroid_batch(p):
    """Volume centroids of the convex hulls of p (H, n, 3), points in
    general position: a triple of points spans a facet when every other
    point lies on one side of its plane; the facet, oriented outwards, adds
    the signed tetrahedron it forms with the origin."""
    return hull_volume_centroid_batch(p)[1]


def hull_volume_centroid_batch(p):
    """(volumes (H,), centroids (H, 3)) of the convex hulls of p (H, n, 3)."""
    H, n, _ = p.shape
    V = np.zeros(H)
    c = np.zeros((H, 3))
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                a, b, cc = p[:, i], p[:, j], p[:, k]
                nrm = np.cross(b - a, cc - a)
                s = np.einsum("hpk,hk->hp", p - a[:, None, :], nrm)
                s[:, [i, j, k]] = 0.0
                tol = 1e-14 * np.linalg.norm(nrm, axis=1)[:, None]
                pos = (s > tol).any(1)
                neg = (s < -tol).any(1)
                facet = ~(pos & neg)
                # outward orientation: the other points lie on the negative side
                b2 = np.where(pos[:, None], cc, b)
                c2 = np.where(pos[:, None], b, cc)
                d = np.einsum("hk,hk->h", a, np.cross(b2, c2)) / 6.0 * facet
                V += d
                c += d[:, None] * (a + b2 + c2) / 4.0
    return V, c / V[:, None]


def hull_centroid(p):
    n = len(p)
    V = 0.0
    c = np.zeros(3)
    for i in range(n):
        for j in range(i + 1, n):
            for k in range(j + 1, n):
                a, b, cc = p[i], p[j], p[k]
                nrm = np.cross(b - a, cc - a)
                s = (p - a) @ nrm
                s[[i, j, k]] = 0
                tol = 1e-14 * np.linalg.norm(nrm)
                pos, neg = (s > tol).any(), (s < -tol).any()
                if pos and neg:
                    continue
                if pos:
                    b, cc = cc, b
                d = a @ np.cross(b, cc) / 6.0
                V += d
                c += d * (a + b + cc) / 4.0
    return c / V
