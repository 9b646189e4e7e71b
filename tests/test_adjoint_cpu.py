"""The vector-Jacobian products of one step (grains3dgpu_amd/csrc/gr_adjoint.hpp,
run on the host through tests/cpp/adjoint_check.cc) against central finite
differences of the forward functions: the FP64 oracle's contact dispatch and
Hooke force (oracle/grains_oracle.c) and the integration kernel's own
particle step (grains3dgpu_amd/csrc/gr_move.hpp).  Each check is a
dot-product test: <y_bar, (f(t + h d) - f(t - h d)) / 2h> against
<VJP(y_bar), d> for random directions d and cotangents y_bar.

Rotations are perturbed as rotations (R = exp([phi]x) R0): the dynamics keep
every basis orthonormal and the oracle's GJK inverts a basis by transposition.

Finite-difference tolerances: the closed forms (spheres, deep cores, walls,
Hooke, integration) to 1e-6 relative; the implicit GJK geometry to 1e-4,
since the forward's witness points converge only like the square root of
the GJK gap (1e-15 here), a noise of ~5e-9 that a step h = 3e-4 divides."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S

HERE = os.path.dirname(os.path.abspath(__file__))
LIBP = os.path.join(HERE, "cpp", "_build", "libadjoint_check.so")


@pytest.fixture(scope="module")
def adj():
    from grains3dgpu_amd import build as gbuild
    if not os.path.exists(LIBP):
        gbuild.build_adjoint_check()
    L = C.CDLL(LIBP)
    L.adjc_gjk_geo_vjp.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.adjc_deep_geo_vjp.argtypes = [C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 4
    L.adjc_sphere_geo_vjp.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double] + [C.c_void_p] * 4
    L.adjc_wall_geo_vjp.argtypes = [C.c_int] + [C.c_void_p] * 6
    L.adjc_hooke_vjp.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_double] + [C.c_void_p] * 4
    L.adjc_dmuen_den.restype = C.c_double
    L.adjc_dmuen_den.argtypes = [C.c_double]
    L.adjc_move.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 4
    L.adjc_integrate_vjp.argtypes = [C.c_double, C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 6
    L.adjc_rot.argtypes = [C.c_void_p, C.c_void_p]
    L.adjc_rot_vjp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    return L


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dot_test(fwd, x0, vjp_val, ybar, rng, h, n_dirs=4):
    """max relative mismatch of <ybar, J d> (central FD) and <J^T ybar, d>"""
    worst = 0.0
    for _ in range(n_dirs):
        d = rng.standard_normal(x0.shape)
        fd = (fwd(x0 + h * d) - fwd(x0 - h * d)) / (2 * h)
        a, b = float(np.dot(ybar, fd)), float(np.dot(vjp_val, d))
        worst = max(worst, abs(a - b) / max(abs(a), abs(b), 1e-300))
    return worst


def rodrigues(phi):
    th = np.linalg.norm(phi)
    K = np.array([[0, -phi[2], phi[1]], [phi[2], 0, -phi[0]], [-phi[1], phi[0], 0]])
    if th < 1e-300:
        return np.eye(3) + K
    return np.eye(3) + math.sin(th) / th * K + (1 - math.cos(th)) / th ** 2 * K @ K


def rotated(t12, x, phi):
    """transform with origin x and basis exp([phi]x) R0 (a rotation, as the
    dynamics keep it: the oracle's GJK inverts a basis by transposition)"""
    t = t12.copy()
    t[:9] = (rodrigues(phi) @ t12[:9].reshape(3, 3)).ravel()
    t[9:] = x
    return t


def pose_bar(tb12, t12):
    """(x_bar, phi_bar) of a transform adjoint: phi_bar_k = <R_bar R^T, [e_k]x>"""
    G = tb12[:9].reshape(3, 3) @ t12[:9].reshape(3, 3).T
    return np.concatenate([tb12[9:], [G[2, 1] - G[1, 2], G[0, 2] - G[2, 0], G[1, 0] - G[0, 1]]])


def ell(a, b, c, crust=0.0):
    return dict(kind=A.SUPERQUADRIC, dims=(a, b, c, 2.0, 2.0), crust=crust, density=1000.0)


def sph(r, crust=0.0):
    return dict(kind=A.SPHERE, dims=(r,), crust=crust, density=1000.0)


def core_of(s):
    ct = s.get("crust", 0.0)
    d = list(s["dims"])
    if s["kind"] == A.SPHERE:
        return dict(s, dims=(d[0] - ct,), crust=0.0), np.array([d[0] - ct, 0, 0.0])
    return dict(s, dims=(d[0] - ct, d[1] - ct, d[2] - ct, 2.0, 2.0), crust=0.0), np.array(d[:3]) - ct


def tr(pos, q):
    return S.transform12(np.asarray(pos, float)[None], np.asarray(q, float)[None])[0]


def place_touching(P, sa, sb, qa, qb, u, frac):
    """B's centre on the ray u from A's, so that the cores are `frac` of the
    crust sum apart (bisection on the oracle's GJK distance of the cores)"""
    ca, _ = core_of(sa)
    cb, _ = core_of(sb)
    target = frac * (sa["crust"] + sb["crust"])
    lo, hi = 0.0, 5.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        d = P.gjk(ca, cb, tr([0, 0, 0], qa), tr(mid * u, qb), rel_tol=1e-15)[0]
        lo, hi = (mid, hi) if d < target else (lo, mid)
    return tr([0, 0, 0], qa), tr(0.5 * (lo + hi) * u, qb)


@pytest.mark.parametrize("pair", ["ell_ell", "sph_ell", "ell_sph"])
def test_gjk_geometry_vjp_matches_finite_differences(oracle_lib, adj, pair):
    P = oracle_lib
    rng = np.random.default_rng({"ell_ell": 1, "sph_ell": 2, "ell_sph": 3}[pair])
    worst = 0.0
    for trial in range(4):
        ea = ell(*rng.uniform(0.3, 0.6, 3), crust=0.05)
        eb = ell(*rng.uniform(0.3, 0.6, 3), crust=0.04)
        sa, sb = {"ell_ell": (ea, eb), "sph_ell": (sph(0.45, 0.05), eb), "ell_sph": (ea, sph(0.4, 0.04))}[pair]
        qa, qb = S.random_quaternions(rng, 2)
        u = rng.standard_normal(3)
        u /= np.linalg.norm(u)
        tA, tB = place_touching(P, sa, sb, qa, qb, u, 0.5)
        ca, dA = core_of(sa)
        cb, dB = core_of(sb)
        d, pa, pb, _ = P.gjk(ca, cb, tA, tB, rel_tol=1e-15)
        n = (pb - pa) / np.linalg.norm(pb - pa)
        ct = sa["crust"] + sb["crust"]
        assert 0 < d < ct
        cvb, cpb = rng.standard_normal(3), rng.standard_normal(3)
        tAb, tBb = np.zeros(12), np.zeros(12)
        adj.adjc_gjk_geo_vjp(sa["kind"], p(f64(dA)), p(f64(tA)), sb["kind"], p(f64(dB)), p(f64(tB)), p(f64(n)), d, ct,
                             p(f64(cvb)), p(f64(cpb)), p(tAb), p(tBb))

        def fwd(z):  # z = (xA, phiA, xB, phiB)
            hit, cp, cv, ov = P.closest_points(sa, sb, rotated(tA, z[:3], z[3:6]), rotated(tB, z[6:9], z[9:]),
                                               rel_tol=1e-15)
            assert hit
            return np.concatenate([cv, cp])
        z0 = np.concatenate([tA[9:], np.zeros(3), tB[9:], np.zeros(3)])
        worst = max(worst, dot_test(fwd, z0, np.concatenate([pose_bar(tAb, tA), pose_bar(tBb, tB)]),
                                    np.concatenate([cvb, cpb]), rng, 3e-4))
    assert worst < 1e-4, worst


def test_deep_and_sphere_geometry_vjp(oracle_lib, adj):
    P = oracle_lib
    rng = np.random.default_rng(4)
    # overlapping ellipsoid cores: the centre line
    sa, sb = ell(0.5, 0.4, 0.45, crust=0.05), ell(0.35, 0.5, 0.4, crust=0.05)
    qa, qb = S.random_quaternions(rng, 2)
    tA, tB = tr([0, 0, 0], qa), tr([0.3, 0.2, -0.1], qb)
    cvb, cpb = rng.standard_normal(3), rng.standard_normal(3)
    xAb, xBb = np.zeros(3), np.zeros(3)
    adj.adjc_deep_geo_vjp(p(f64(tA[9:])), p(f64(tB[9:])), 0.1, p(f64(cvb)), p(f64(cpb)), p(xAb), p(xBb))

    def fwd_deep(x):
        a, b = tA.copy(), tB.copy()
        a[9:], b[9:] = x[:3], x[3:]
        hit, cp, cv, ov = P.closest_points(sa, sb, a, b, rel_tol=1e-15)
        assert hit and abs(ov + 0.1) < 1e-15
        return np.concatenate([cv, cp])
    assert dot_test(fwd_deep, np.concatenate([tA[9:], tB[9:]]), np.concatenate([xAb, xBb]),
                    np.concatenate([cvb, cpb]), rng, 1e-6) < 1e-6
    # two spheres: closestPointsSpheres
    ra, rb = 0.5, 0.3
    tA, tB = tr([0, 0, 0], [0, 0, 0, 1.0]), tr([0.5, 0.4, 0.2], [0, 0, 0, 1.0])
    adj.adjc_sphere_geo_vjp(p(f64(tA[9:])), p(f64(tB[9:])), ra, rb, p(f64(cvb)), p(f64(cpb)), p(xAb), p(xBb))

    def fwd_sph(x):
        a, b = tA.copy(), tB.copy()
        a[9:], b[9:] = x[:3], x[3:]
        hit, cp, cv, ov = P.closest_points(sph(ra), sph(rb), a, b)
        assert hit
        return np.concatenate([cv, cp])
    assert dot_test(fwd_sph, np.concatenate([tA[9:], tB[9:]]), np.concatenate([xAb, xBb]),
                    np.concatenate([cvb, cpb]), rng, 1e-6) < 1e-6


@pytest.mark.parametrize("kind", ["ell", "sph"])
def test_wall_geometry_vjp(oracle_lib, adj, kind):
    P = oracle_lib
    rng = np.random.default_rng(5)
    sa = ell(0.5, 0.3, 0.4) if kind == "ell" else sph(0.45)
    dims = np.array(sa["dims"][:3] if kind == "ell" else [sa["dims"][0], 0, 0], float)
    wall = dict(kind=A.RECTANGLE, dims=(5.0, 5.0), density=0.0)
    tW = S.transform12(np.zeros((1, 3)), np.array([[0.1, -0.05, 0.02, 1.0]]) / np.linalg.norm([0.1, -0.05, 0.02, 1]))[0]
    rN = tW[[2, 5, 8]]
    q = S.random_quaternions(rng, 1)[0]
    tA = tr(0.35 * rN + np.array([0.2, -0.1, 0.0]), q)
    hit, cp, cv, ov = P.closest_points(sa, wall, tA, tW)
    assert hit
    rNorm = rN * math.copysign(1.0, float(np.dot(rN, tA[9:] - tW[9:])))
    cvb, cpb = rng.standard_normal(3), rng.standard_normal(3)
    tAb = np.zeros(12)
    adj.adjc_wall_geo_vjp(sa["kind"], p(f64(dims)), p(f64(tA)), p(f64(rNorm)), p(f64(cvb)), p(f64(cpb)), p(tAb))

    def fwd(z):
        hit, cp, cv, ov = P.closest_points(sa, wall, rotated(tA, z[:3], z[3:]), tW)
        assert hit
        return np.concatenate([cv, cp])
    z0 = np.concatenate([tA[9:], np.zeros(3)])
    assert dot_test(fwd, z0, pose_bar(tAb, tA), np.concatenate([cvb, cpb]), rng, 1e-6) < 1e-6


@pytest.mark.parametrize("branch", ["viscous", "coulomb", "wall"])
def test_hooke_vjp(oracle_lib, adj, branch):
    P = oracle_lib
    rng = np.random.default_rng({"viscous": 6, "coulomb": 7, "wall": 8}[branch])
    kn, en, etat, muc = 1e5, 0.8, 2e3, 0.3
    m1, m2 = 1.3e-6, (0.0 if branch == "wall" else 0.9e-6)
    n = rng.standard_normal(3)
    n /= np.linalg.norm(n)
    cv = -2e-6 * n
    xA = rng.uniform(-1e-3, 1e-3, 3)
    xB = xA + 1e-3 * n
    cp = xA + 0.5e-3 * n + rng.uniform(-1e-5, 1e-5, 3)
    tang = 1e-4 if branch == "viscous" else 0.3
    vA, vB = rng.standard_normal(3) * tang, rng.standard_normal(3) * tang
    wA, wB = rng.standard_normal(3) * 10, rng.standard_normal(3) * 10
    if branch == "wall":
        vB, wB, xB = np.zeros(3), np.zeros(3), cp.copy()
    x0 = np.concatenate([cv, cp, vA, wA, vB, wB, xA, xB])
    prm = np.array([kn, en, etat, muc])

    def fwd(z):
        x, pr = z[:24], z[24:]
        cv_, cp_, vA_, wA_, vB_, wB_, xA_, xB_ = x.reshape(8, 3)
        rel = (vA_ + np.cross(wA_, cp_ - xA_)) - (vB_ + np.cross(wB_, cp_ - xB_))
        F, M = P.hooke([pr[0], pr[1], pr[2], pr[3], 0.0], cp_, cv_, -np.linalg.norm(cv_), rel, wA_ - wB_, m1, m2, xA_)
        return np.concatenate([F, M])
    Fb, Mb = rng.standard_normal(3), rng.standard_normal(3) * 1e3
    out, F = np.zeros(28), np.zeros(3)
    muen = math.log(en) / math.sqrt(math.pi ** 2 + math.log(en) ** 2)
    adj.adjc_hooke_vjp(p(f64([kn, muen, etat, muc])), p(f64(x0)), m1, m2, p(f64(Fb)), p(f64(Mb)), p(out), p(F))
    F_or, _ = P.hooke([kn, en, etat, muc, 0.0], cp, cv, -np.linalg.norm(cv),
                      (vA + np.cross(wA, cp - xA)) - (vB + np.cross(wB, cp - xB)), wA - wB, m1, m2, xA)
    np.testing.assert_allclose(F, F_or, rtol=1e-12)
    # Coulomb or viscous as intended
    ft = 2 * etat * (m1 * m2 / (m1 + m2) if m2 else 0.5 * m1)
    g = np.concatenate([out[:24], [out[24], out[25] * adj.adjc_dmuen_den(en), out[26], out[27]]])
    # scale the inputs so one relative step fits every component
    vs = max(tang, 1e-3)
    scale = np.concatenate([np.full(3, 2e-6), np.full(3, 1e-3), np.full(3, vs), np.full(3, 10), np.full(3, vs),
                            np.full(3, 10), np.full(6, 1e-3), prm])

    def fwd_s(z):
        return fwd(z * scale)
    z0 = np.concatenate([x0, prm]) / scale
    assert dot_test(fwd_s, z0, g * scale, np.concatenate([Fb, Mb]), rng, 1e-6) < 1e-6
    del ft


def test_integrate_vjp(adj):
    rng = np.random.default_rng(9)
    I = np.array([2.0, 0.1, -0.05, 1.5, 0.02, 1.2]) * 1e-10
    Im = np.array([[I[0], I[1], I[2]], [I[1], I[3], I[4]], [I[2], I[4], I[5]]])
    Iv = np.linalg.inv(Im)
    Ii = np.array([Iv[0, 0], Iv[0, 1], Iv[0, 2], Iv[1, 1], Iv[1, 2], Iv[2, 2]])
    mass, dt = 1.3e-6, 1e-4
    q = S.random_quaternions(rng, 1)[0]
    st = np.concatenate([rng.standard_normal(3) * 1e-3, q, rng.standard_normal(3) * 0.1, rng.standard_normal(3) * 50])
    F, Mt = rng.standard_normal(3) * 1e-5, rng.standard_normal(3) * 1e-9
    x0 = np.concatenate([st, F, Mt])

    def fwd(z):
        out = np.zeros(13)
        adj.adjc_move(mass, p(f64(I)), p(f64(Ii)), dt, p(f64(z[:13])), p(f64(z[13:16])), p(f64(z[16:])), p(out))
        return out
    scale = np.concatenate([np.full(3, 1e-3), np.ones(4), np.full(3, 0.1), np.full(3, 50), np.full(3, 1e-5),
                            np.full(3, 1e-9)])
    sb1 = rng.standard_normal(13)
    sb, Fb, Mb = np.zeros(13), np.zeros(3), np.zeros(3)
    adj.adjc_integrate_vjp(mass, p(f64(I)), p(f64(Ii)), dt, p(f64(st)), p(f64(Mt)), p(f64(sb1)), p(sb), p(Fb), p(Mb))
    g = np.concatenate([sb, Fb, Mb])
    assert dot_test(lambda z: fwd(z * scale), x0 / scale, g * scale, sb1, rng, 1e-6) < 1e-7


def test_rot_vjp(adj):
    rng = np.random.default_rng(10)
    q = rng.standard_normal(4)
    Rb = rng.standard_normal(9)
    qb = np.zeros(4)
    adj.adjc_rot_vjp(p(f64(q)), p(f64(Rb)), p(qb))

    def fwd(z):
        R = np.zeros(9)
        adj.adjc_rot(p(f64(z)), p(R))
        return R
    assert dot_test(fwd, q, qb, Rb, rng, 1e-6) < 1e-9
