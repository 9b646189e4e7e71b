"""Pins of the oracle's narrow phase against closed forms and brute force:
sphere and rectangle contacts (CollisionDetection.cpp:27-90), support maps
(Superquadric.cpp:169-205, Box.cpp:129-134, Sphere.cpp:111-114), OBB
(OBB.cpp:32-94), GJK distances (GJK_JH.cpp:310-391 for the reference's
Johnson GJK; the signed-volume GJK of the physics semantics) and the
reference's early-exit quirk."""
import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S
from grains3dgpu_amd.shapes import superquadric_rc_exact

I12 = np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0.0])


def tr(pos, quat=(0, 0, 0, 1.0)):
    return S.transform12(np.asarray(pos, float), np.asarray(quat, float))[0]


def sq(a, b, c, n1, n2, crust=0.0):
    return dict(kind=A.SUPERQUADRIC, dims=(a, b, c, n1, n2), crust=crust, density=1000.0)


def sq_implicit(p, a, b, c, n1, n2):
    """Superquadric surface function of the support map's shape:
    (|x/a|^n2 + |y/b|^n2)^(n1/n2) + |z/c|^n1."""
    return (np.abs(p[..., 0] / a) ** n2 + np.abs(p[..., 1] / b) ** n2) ** (n1 / n2) + \
        np.abs(p[..., 2] / c) ** n1


def sq_surface_samples(a, b, c, n1, n2, m=200):
    """Points on the surface: x = a cos^e1(eta) cos^e2(w) ... with e = 2/n."""
    e1, e2 = 2.0 / n1, 2.0 / n2
    eta = np.linspace(-np.pi / 2, np.pi / 2, m)
    om = np.linspace(-np.pi, np.pi, 2 * m)
    E, W = np.meshgrid(eta, om, indexing="ij")
    f = lambda t, e: np.sign(t) * np.abs(t) ** e
    x = a * f(np.cos(E), e1) * f(np.cos(W), e2)
    y = b * f(np.cos(E), e1) * f(np.sin(W), e2)
    z = c * f(np.sin(E), e1)
    return np.stack([x.ravel(), y.ravel(), z.ravel()], 1)


def test_sphere_contact_closed_form(oracle_lib):
    P = oracle_lib
    sa = dict(kind=A.SPHERE, dims=(0.5,), density=1.0)
    sb = dict(kind=A.SPHERE, dims=(0.3,), density=1.0)
    d = np.array([0.3, -0.4, 0.2])
    hit, cp, cv, ov = P.closest_points(sa, sb, tr([1, 2, 3]), tr(np.array([1, 2, 3]) + d))
    n = d / np.linalg.norm(d)
    dist = np.linalg.norm(d)
    assert hit
    np.testing.assert_allclose(ov, dist - 0.8, rtol=1e-14)
    np.testing.assert_allclose(cv, (dist - 0.8) * n, rtol=1e-13)
    np.testing.assert_allclose(cp, np.array([1, 2, 3]) + (0.5 + 0.5 * (dist - 0.8)) * n, rtol=1e-13)
    hit, _, _, ov = P.closest_points(sa, sb, tr([0, 0, 0]), tr([0.81, 0, 0]))
    assert not hit and ov == 0.0


@pytest.mark.parametrize("kind", ["sphere", "box", "sq"])
def test_rectangle_contact_plane_projection(oracle_lib, kind):
    P = oracle_lib
    rect = dict(kind=A.RECTANGLE, dims=(5.0, 5.0), density=0.0)
    rng = np.random.default_rng(1)
    q = S.random_quaternions(rng, 1)[0]
    if kind == "sphere":
        s = dict(kind=A.SPHERE, dims=(0.5,), density=1.0)
        deepest_h = lambda R: 0.5
    elif kind == "box":
        s = dict(kind=A.BOX, dims=(0.3, 0.4, 0.5), density=1.0)
        deepest_h = lambda R: np.abs(R[2, :]) @ np.array([0.3, 0.4, 0.5])
    else:
        s = sq(0.3, 0.4, 0.5, 2.0, 2.0)
        # ellipsoid support height along -z: sqrt(sum (a_i R_zi)^2)
        deepest_h = lambda R: np.sqrt(((np.array([0.3, 0.4, 0.5]) * R[2, :]) ** 2).sum())
    R = S.quat_to_matrix(q).reshape(3, 3)
    h = deepest_h(R)
    z = h - 0.07
    hit, cp, cv, ov = P.closest_points(s, rect, tr([0.2, -0.1, z], q), I12)
    assert hit
    np.testing.assert_allclose(ov, -0.07, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(cv, [0, 0, 0.07], atol=1e-13)
    assert abs(cp[2] - (-0.035)) < 1e-12
    # same from above a ceiling: the normal flips toward the particle
    # (symmetric particle: the deepest point toward +z is at height h too)
    ceil = tr([0, 0, z + h - 0.03], [1.0, 0, 0, 0])  # rotated 180 deg about x
    hit, _, cv, ov = P.closest_points(s, rect, tr([0.2, -0.1, z], q), ceil)
    assert hit
    np.testing.assert_allclose(ov, -0.03, rtol=1e-9)
    np.testing.assert_allclose(cv, [0, 0, -0.03], atol=1e-13)
    ceil = tr([0, 0, z + h + 0.01], [1.0, 0, 0, 0])
    assert not P.closest_points(s, rect, tr([0.2, -0.1, z], q), ceil)[0]
    hit, _, _, _ = P.closest_points(s, rect, tr([0, 0, h + 1e-3], q), I12)
    assert not hit


def test_superquadric_support_ellipsoid_closed_form(oracle_lib):
    P = oracle_lib
    a, b, c = 0.7, 0.4, 0.55
    rng = np.random.default_rng(2)
    for v in rng.standard_normal((200, 3)):
        s = P.support(sq(a, b, c, 2.0, 2.0), v)
        ref = np.array([a * a * v[0], b * b * v[1], c * c * v[2]]) / np.linalg.norm([a * v[0], b * v[1], c * v[2]])
        np.testing.assert_allclose(s, ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("n1,n2", [(3.0, 3.0), (2.5, 6.0), (8.0, 2.0), (5.0, 4.0)])
def test_superquadric_support_is_maximal_and_on_surface(oracle_lib, n1, n2):
    P = oracle_lib
    a, b, c = 0.7, 0.4, 0.55
    pts = sq_surface_samples(a, b, c, n1, n2)
    rng = np.random.default_rng(4)
    dirs = np.concatenate([rng.standard_normal((100, 3)),
                           [[0, 0, 1.0], [0, 1.0, 0], [1.0, 0, 0], [0, -0.3, 0.8], [-0.2, 0, 0.9]]])
    for v in dirs:
        s = P.support(sq(a, b, c, n1, n2), v)
        assert abs(sq_implicit(s, a, b, c, n1, n2) - 1.0) < 1e-9
        assert v @ s >= (pts @ v).max() - 1e-9 * np.linalg.norm(v)
    # zero components follow sgn(0) = 0 (Superquadric.cpp:180-183)
    np.testing.assert_array_equal(P.support(sq(a, b, c, n1, n2), [0, 0, -2.0]), [0, 0, -c])


def test_box_and_polyhedron_support_brute_force(oracle_lib):
    P = oracle_lib
    rng = np.random.default_rng(5)
    e = np.array([0.3, 0.2, 0.6])
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * e
    verts = rng.standard_normal((8, 3))
    poly = dict(kind=A.POLYHEDRON, dims=(0,) * 5, density=1.0, n_vertices=8, vertex_offset=0)
    for v in rng.standard_normal((100, 3)):
        s = P.support(dict(kind=A.BOX, dims=tuple(e), density=1.0), v)
        assert np.isclose(v @ s, (corners @ v).max(), rtol=1e-14)
        s = P.support(poly, v, vertices=verts)
        np.testing.assert_array_equal(s, verts[np.argmax(verts @ v)])


def sampled_distance_lower_bound(sa, sb, ta, tb, P, dirs, verts=None):
    """Separating-axis lower bound: dist >= max_u (min_B u.x - max_A u.x)."""
    Ra, Rb = ta[:9].reshape(3, 3), tb[:9].reshape(3, 3)
    best = -np.inf
    for u in dirs:
        u = u / np.linalg.norm(u)
        ha = u @ (Ra @ P.support(sa, Ra.T @ u, verts) + ta[9:])
        hb = u @ (Rb @ P.support(sb, Rb.T @ -u, verts) + tb[9:])
        best = max(best, hb - ha)
    return best


@pytest.mark.parametrize("flags", [A.REF_QUIRKS, 0, A.NESTEROV])
def test_gjk_distance_closed_forms(oracle_lib, flags):
    P = oracle_lib
    # spheres as superquadrics n1 = n2 = 2
    d, pa, pb, it = P.gjk(sq(0.5, 0.5, 0.5, 2, 2), sq(0.3, 0.3, 0.3, 2, 2), tr([0, 0, 0]),
                          tr([1.2, 0.9, -0.4]), flags=flags, rel_tol=1e-12)
    assert abs(d - (np.linalg.norm([1.2, 0.9, -0.4]) - 0.8)) < 1e-9
    # axis-aligned boxes separated along x
    bx = dict(kind=A.BOX, dims=(0.5, 0.4, 0.3), density=1.0)
    d, pa, pb, it = P.gjk(bx, bx, tr([0, 0, 0]), tr([1.25, 0.1, 0.05]), flags=flags, rel_tol=1e-12)
    assert abs(d - 0.25) < 1e-9
    assert abs(pb[0] - pa[0] - 0.25) < 1e-9


@pytest.mark.parametrize("flags", [A.REF_QUIRKS, 0, A.NESTEROV])
def test_gjk_random_separated_pairs_bounds_and_symmetry(oracle_lib, flags):
    P = oracle_lib
    rng = np.random.default_rng(7)
    dirs = rng.standard_normal((400, 3))
    n_ok = 0
    for t in range(25):
        pa_ = rng.uniform(0.3, 0.8, 3)
        pb_ = rng.uniform(0.3, 0.8, 3)
        A_ = sq(*pa_, *rng.uniform(2, 8, 2))
        B_ = sq(*pb_, *rng.uniform(2, 8, 2))
        qa, qb = S.random_quaternions(rng, 2)
        off = rng.standard_normal(3)
        off *= (np.linalg.norm(pa_) + np.linalg.norm(pb_) + rng.uniform(0.05, 0.5)) / np.linalg.norm(off)
        ta, tb = tr([0, 0, 0], qa), tr(off, qb)
        d, wa, wb, it = P.gjk(A_, B_, ta, tb, flags=flags, rel_tol=1e-10)
        d2, _, _, _ = P.gjk(B_, A_, tb, ta, flags=flags, rel_tol=1e-10)
        lb = sampled_distance_lower_bound(A_, B_, ta, tb, P, dirs)
        # witness distance is an upper bound, sampled separating axes a lower one
        assert np.linalg.norm(wb - wa) >= d * (1 - 1e-6) - 1e-12
        assert d >= lb - 1e-9
        assert abs(np.linalg.norm(wb - wa) - d) < 1e-6 * max(d, 1e-3)
        assert abs(d - d2) < 1e-6 * max(d, 1e-3)
        # witness points on the surfaces
        Ra = ta[:9].reshape(3, 3)
        la = Ra.T @ (wa - ta[9:])
        # (accelerated simplices may hold support points of other directions:
        # the barycentric witness then sits slightly inside the surface)
        tol = 1e-2 if flags & A.NESTEROV else 1e-4
        assert abs(sq_implicit(la, *A_["dims"]) - 1.0) < tol
        n_ok += 1
    assert n_ok == 25


@pytest.mark.parametrize("flags", [0, A.NESTEROV])
def test_gjk_gap_bound_at_tight_tolerance_mm_scale(oracle_lib, flags):
    """The gap test bounds |d - d*| by gap * d, so searches stopped at gaps
    1e-10 and 1e-12 agree to 1e-10 relative.  Millimetre superquadrics in SI
    units, random blockiness: near convergence the simplex spans ~1e-6 of
    |v|, where barycentric coordinates taken from cross products of two
    far-away vertices (b x c) lose ~1e-4 and the distance drifts by ~1e-7
    in either direction.  Momentum must not stall the search either."""
    P = oracle_lib
    rng = np.random.default_rng(5)
    prm = S.superquadric_params(rng, 40)
    worst, max_it = 0.0, 0
    for t in range(40):
        A_ = sq(*prm[t])
        B_ = sq(*prm[(t + 7) % 40])
        qa, qb = S.random_quaternions(rng, 2)
        off = rng.standard_normal(3)
        ra = superquadric_rc_exact(*prm[t])
        rb = superquadric_rc_exact(*prm[(t + 7) % 40])
        off *= (ra + rb + rng.uniform(0.05e-3, 0.5e-3)) / np.linalg.norm(off)
        ta, tb = tr([0, 0, 0], qa), tr(off, qb)
        d12, _, _, it = P.gjk(A_, B_, ta, tb, flags=flags, rel_tol=1e-12)
        d10, _, _, _ = P.gjk(A_, B_, ta, tb, flags=flags, rel_tol=1e-10)
        worst = max(worst, abs(d12 - d10) / d12)
        max_it = max(max_it, it)
    assert worst < 1.01e-10
    assert max_it < 100


def test_obb_never_rejects_overlapping_pairs(oracle_lib):
    """Pairs whose shapes share a sampled point must reach GJK: the physics
    semantics then report a contact."""
    P = oracle_lib
    rng = np.random.default_rng(9)
    hits = 0
    for t in range(200):
        A_ = sq(*rng.uniform(0.3, 0.8, 3), *rng.uniform(2, 8, 2), crust=0.02)
        B_ = sq(*rng.uniform(0.3, 0.8, 3), *rng.uniform(2, 8, 2), crust=0.02)
        qa, qb = S.random_quaternions(rng, 2)
        ta, tb = tr([0, 0, 0], qa), tr(rng.uniform(-1.2, 1.2, 3), qb)
        # a sample point inside both shapes (cores, shrunk by the crust)?
        pts = rng.uniform(-1, 1, (4000, 3))
        la = pts @ ta[:9].reshape(3, 3)  # R^T p (rows)
        lb = (pts - tb[9:]) @ tb[:9].reshape(3, 3)
        da = np.array(A_["dims"]); db = np.array(B_["dims"])
        ina = sq_implicit(la, *(da[:3] - 0.02), *da[3:]) < 1
        inb = sq_implicit(lb, *(db[:3] - 0.02), *db[3:]) < 1
        if (ina & inb).any():
            hit, _, _, ov = P.closest_points(A_, B_, ta, tb, flags=0)
            assert hit and ov < -0.02
            hits += 1
    assert hits > 20


def test_reference_early_exit_quirk(oracle_lib):
    """REF_QUIRKS reproduces GJK_JH.cpp:370-372: overlapping superquadrics whose
    centre line shows overlap exit at iteration 0 with no contact; the physics
    semantics (crust cores) find the contact."""
    P = oracle_lib
    A_ = sq(0.5, 0.6, 0.7, 4, 4, crust=0.05)
    hit_q, _, _, ov_q = P.closest_points(A_, A_, tr([0, 0, 0]), tr([0.9, 0, 0]), flags=A.REF_QUIRKS)
    assert not hit_q
    d, pa, pb, it = P.gjk(A_, A_, tr([0, 0, 0]), tr([0.9, 0, 0]), flags=A.REF_QUIRKS)
    assert it == 0 and np.isnan(pa).all()
    hit, cp, cv, ov = P.closest_points(A_, A_, tr([0, 0, 0]), tr([0.9, 0, 0]), flags=0)
    assert hit
    # cores are boxes-like superquadrics of half width 0.45 along x: gap 0,
    # overlap = 0.9 - 2 * 0.45 - 2 * crust = -0.1
    np.testing.assert_allclose(ov, -0.1, atol=1e-6)
    np.testing.assert_allclose(cv, [-0.1 * -1, 0, 0] if False else [-0.1, 0, 0], atol=1e-6)
    # spheres are unaffected by the quirk
    s = dict(kind=A.SPHERE, dims=(0.5,), density=1.0)
    assert P.closest_points(s, s, tr([0, 0, 0]), tr([0.9, 0, 0]), flags=A.REF_QUIRKS)[0]


def test_sphere_vs_nonsphere_quirk_never_contacts(oracle_lib):
    """Sphere::support(0) is 0/0 (Sphere.cpp:111-114), so the reference's GJK
    starts from NaN and a sphere never touches a non-sphere."""
    P = oracle_lib
    s = dict(kind=A.SPHERE, dims=(0.5,), density=1.0, crust=0.15)
    b = dict(kind=A.BOX, dims=(0.5, 0.5, 0.5), density=1.0, crust=0.15)
    assert not P.closest_points(s, b, tr([0, 0, 0]), tr([0.9, 0, 0]), flags=A.REF_QUIRKS)[0]
    hit, _, _, ov = P.closest_points(s, b, tr([0, 0, 0]), tr([0.9, 0, 0]), flags=0)
    assert hit and abs(ov - (-0.1)) < 0.2 * 1e-5  # GJK relative tolerance LOWEPS
