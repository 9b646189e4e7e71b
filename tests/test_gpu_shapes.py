"""Parity of the HIP path with the FP64 oracle for cylinders and cones
(Cylinder.cpp:118-131, Cone.cpp:126-141) mixed with superquadrics, and for
the rolling-resistance moment (HookeContactForceModel.cpp:143-156).

Bars: cell keys, sorted order, candidate and contact lists, contact counts
bit-exact.  Cylinders and cones alone: the support maps use only correctly
rounded sqrt and division, so FP64 forces, torques and trajectories match
the oracle bit for bit.  Mixed with superquadrics, the libm-versus-device pow
ulps enter; the cylinder's rim/cap and the cone's apex/rim choices are
discontinuous in the direction, so an ulp can send a search down another
path to the same gap.  A search stopped at relative gap eps pins the
contact normal to sqrt(2 eps) (|v| <= d*/(1 - eps) and v.n* >= d*), so each
side's contact force is within sqrt(2 eps) |F_c| of the exact one; a
particle sums up to ~10 contacts: bar 10 sqrt(2 eps) = 1.4e-5 at eps = 1e-12
(1e-9 with the reference's Johnson GJK, whose paths coincide).  Spheres
with rolling resistance: bit-exact.
"""
import dataclasses

import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S
from tests.scenes import dense_cylinders_cones, dense_spheres, lattice_cylinders_cones, rel_err
from tests.test_gpu_parity import both, check_broad_phase

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("flags", [0, A.REF_QUIRKS, A.NESTEROV])
def test_cylinders_cones_first_step_fp64(flags):
    sc = dense_cylinders_cones(n=300, seed=41)
    e, o = both(sc, flags=flags, tol=1e-12)
    e.step(1)
    o.step(1)
    cand = check_broad_phase(e, o)
    assert len(cand) > 300
    ce, co = e.get_pairs(A.PAIRS_CONTACT), o.get_pairs(A.PAIRS_CONTACT)
    np.testing.assert_array_equal(ce, co)
    if not flags & A.REF_QUIRKS:
        assert len(ce) > 50
    np.testing.assert_array_equal(e.get_contact_count(), o.get_contact_count())
    fe, te = e.get_torce()
    fo, to = o.get_torce()
    bar = 1e-9 if flags & A.REF_QUIRKS else 10.0 * np.sqrt(2e-12)
    assert rel_err(fe, fo) < bar, rel_err(fe, fo)
    assert rel_err(te, to) < bar, rel_err(te, to)
    se, so = e.get_stats(), o.get_stats()
    assert se["gjk_calls"] == so["gjk_calls"]


def test_cylinders_cones_only_bit_exact():
    """Cylinders and cones alone: no pow anywhere, FP64 forces and torques
    bit-exact on the first step (reference GJK and signed-volume GJK)."""
    sc = dense_cylinders_cones(n=300, seed=43, shares=(0.5, 0.5, 0.0))
    for flags in (A.REF_QUIRKS, 0):
        e, o = both(sc, flags=flags, tol=1e-12)
        e.step(1)
        o.step(1)
        check_broad_phase(e, o)
        np.testing.assert_array_equal(e.get_pairs(A.PAIRS_CONTACT), o.get_pairs(A.PAIRS_CONTACT))
        fe, te = e.get_torce()
        fo, to = o.get_torce()
        np.testing.assert_array_equal(fe, fo)
        np.testing.assert_array_equal(te, to)
        e.close()


def test_cylinders_cones_trajectory_fp64():
    """100 steps of the cylinder-cone lattice (crusts pressed by a quarter
    crust: ~7 m/s rebounds, chaotic).  The oracle keeps rotation matrices and
    the device quaternions, so the two differ by ulps from the first step on;
    the bar is ten times the oracle's own spread under an ulp-level change of
    the initial positions (1e-15 relative), measured here."""
    from oracle import pyoracle
    sc = lattice_cylinders_cones()
    e, o = both(sc, tol=1e-12)
    e.step(1)
    o.step(1)
    fe, te = e.get_torce()
    fo, to = o.get_torce()
    np.testing.assert_array_equal(fe, fo)
    np.testing.assert_array_equal(te, to)
    e.step(99)
    o.step(99)
    sc2 = dataclasses.replace(sc, pos=sc.pos * (1 + 1e-15))
    o2 = S.load(pyoracle.Oracle(dt=sc.dt, gjk_rel_tol=1e-12), sc2)
    o2.step(100)
    pe, _, ve, _ = e.get_state()
    po, _, vo, _ = o.get_state()
    p2, _, v2, _ = o2.get_state()
    assert e.get_stats()["contacts"] > 50 * 99
    L = sc.sims[0]["domain_max"][0]
    spread_p = np.abs(p2 - po).max()
    assert 0 < spread_p < 1e-6 * L
    assert np.abs(pe - po).max() < 10 * spread_p + 1e-14 * L
    assert rel_err(ve, vo) < 10 * rel_err(v2, vo) + 1e-13


def test_cylinders_cones_fp32_mixed():
    """FP32-mixed at gap 1e-6 against the FP64 oracle at 1e-12.  On the
    superquadric lattice the FP32 search follows the FP64 one and lands
    within three times the FP64 oracle's own error at the same gap.  Flat
    caps and straight mantles do not hold the path: nearly parallel faces
    leave the normal free within the gap's bound, sqrt(2 eps) = 1.4e-3 at
    eps = 1e-6 (module docstring), which is the bar here; the contact
    decisions may differ only at the threshold."""
    sc = dense_cylinders_cones(n=300, seed=45)
    e, o = both(sc, precision=A.FP32_MIXED, tol=1e-6, oracle_tol=1e-12)
    e.step(1)
    o.step(1)
    check_broad_phase(e, o)
    se = set(map(tuple, e.get_pairs(A.PAIRS_CONTACT)))
    so = set(map(tuple, o.get_pairs(A.PAIRS_CONTACT)))
    assert len(se ^ so) <= max(2, len(so) // 100)
    ok = np.array([i not in {a for a, _ in se ^ so} for i in range(sc.n)])
    fe, _ = e.get_torce()
    fo, _ = o.get_torce()
    assert rel_err(fe[ok], fo[ok]) < np.sqrt(2e-6)


@pytest.mark.parametrize("scene", ["spheres", "mixed"])
def test_rolling_resistance(scene):
    """kr > 0: the moment -kr |F_N| (1e-3 w_n + w_t) joins every contact's
    torque.  kr = 1e-11 m s keeps the explicit step stable (kr |F_N| dt / I
    ~ 1e-2 for 0.35 mm spheres under ~30 N).  Spheres: bit-exact forces and
    torques on the first step, then 49 steps against the oracle's own
    spread; the mixed scene to the bars of the first-step test."""
    if scene == "spheres":
        sc = dense_spheres(n=400, seed=22, walls=True)
        sc.tables = np.array([S.hooke_row(1e5, 0.8, 2e4, 0.3, 1e-11)])
    else:
        sc = dense_cylinders_cones(n=300, seed=46, kr=1e-11)
    e, o = both(sc, tol=1e-12)
    e.step(1)
    o.step(1)
    fe, te = e.get_torce()
    fo, to = o.get_torce()
    # the moment is there: torques differ from those without rolling resistance
    from oracle import pyoracle
    sc0 = dataclasses.replace(sc, tables=sc.tables.copy())
    sc0.tables[:, 4] = 0.0
    o0 = S.load(pyoracle.Oracle(dt=sc.dt, gjk_rel_tol=1e-12), sc0)
    o0.step(1)
    assert rel_err(o0.get_torce()[1], to) > 1e-8
    if scene == "spheres":
        np.testing.assert_array_equal(fe, fo)
        np.testing.assert_array_equal(te, to)
        # 49 more steps: the angular velocities differ by ulps after the
        # first integration (matrix versus quaternion), so the bar is ten
        # times the oracle's spread under a 1e-15 change of the positions
        e.step(49)
        o.step(49)
        o2 = S.load(pyoracle.Oracle(dt=sc.dt, gjk_rel_tol=1e-12), dataclasses.replace(sc, pos=sc.pos * (1 + 1e-15)))
        o2.step(50)
        pe, _, _, we = e.get_state()
        po, _, _, wo = o.get_state()
        p2, _, _, w2 = o2.get_state()
        assert rel_err(pe, po) <= 10 * rel_err(p2, po) + 1e-15
        assert rel_err(we, wo) <= 10 * rel_err(w2, wo) + 1e-13
    else:
        bar = 10.0 * np.sqrt(2e-12)
        assert rel_err(fe, fo) < bar and rel_err(te, to) < bar


def test_gjk_batch_cylinders_cones():
    from grains3dgpu_amd import engine
    from oracle import pyoracle
    rng = np.random.default_rng(6)
    n = 256
    shapes = [dict(kind=A.CYLINDER if k % 2 else A.CONE, dims=(rng.uniform(0.3e-3, 0.6e-3),
                                                              rng.uniform(0.6e-3, 1.6e-3)), density=1.0)
              for k in range(8)]
    shapes.append(dict(kind=A.SUPERQUADRIC, dims=tuple(S.superquadric_params(rng, 1)[0]), density=1.0))
    sc = S.Scene(shapes=shapes, tables=np.array([S.hooke_row(1e5, 0.8, 0, 0.3)]), n_materials=1,
                 sims=[dict(domain_min=(0, 0, 0), domain_max=(1, 1, 1))],
                 particle_shape=np.zeros(1, np.uint32), pos=np.zeros((1, 3)),
                 quat=np.array([[0, 0, 0, 1.0]]), vel=np.zeros((1, 3)), omega=np.zeros((1, 3)))
    sa = rng.integers(0, 9, n).astype(np.uint32)
    sb = rng.integers(0, 9, n).astype(np.uint32)
    qa, qb = S.random_quaternions(rng, n), S.random_quaternions(rng, n)
    rc = np.array([pyoracle.shape_props(s)["rc_exact"] for s in shapes])
    off = rng.standard_normal((n, 3))
    off *= ((rc[sa] + rc[sb] + rng.uniform(0.05e-3, 0.5e-3, n)) / np.linalg.norm(off, axis=1))[:, None]
    ta = S.transform12(rng.uniform(0, 1e-2, (n, 3)), qa)
    tb = S.transform12(ta[:, 9:] + off, qb)
    for prec, flags, gt, tol in [(A.FP64, A.REF_QUIRKS, 0.0, 1e-12), (A.FP64, 0, 1e-12, 1e-9),
                                 (A.FP32_MIXED, 0, 1e-6, 2e-4), (A.FP64, A.NESTEROV, 1e-12, 1e-9)]:
        e = engine.Engine(precision=prec, flags=flags, gjk_rel_tol=gt)
        e.set_shapes(sc)
        d, pa, pb, it = e.gjk_batch(sa, sb, ta, tb)
        ref = np.array([pyoracle.gjk(shapes[a], shapes[b], x, y, flags=flags,
                                     rel_tol=gt if prec == A.FP64 else 1e-12)[0]
                        for a, b, x, y in zip(sa, sb, ta, tb)])
        assert np.isfinite(d).all()
        assert rel_err(d, ref) < tol, (prec, flags, rel_err(d, ref))
        e.close()


def with_obstacles(sc):
    """The scene plus three obstacles at rest between its walls: a tilted Box
    at the centre, a Cylinder lying along x and a Superquadric, all of
    density 0 (mass 0: the avmass branch of HookeContactForceModel.cpp:124-129)."""
    L = sc.sims[0]["domain_max"][0]
    shapes = list(sc.shapes)
    first = len(shapes)
    shapes += [dict(kind=A.BOX, material=0, dims=(L / 8, L / 10, L / 12), crust=1e-5, density=0.0),
               dict(kind=A.CYLINDER, material=0, dims=(L / 12, 0.9 * L), crust=1e-5, density=0.0),
               dict(kind=A.SUPERQUADRIC, material=0, dims=(L / 10, L / 9, L / 8, 4.0, 3.0), crust=1e-5, density=0.0)]
    q = np.array([[0.1, 0.2, 0.05, 1.0], [0.0, 0.0, np.sin(np.pi / 4), np.cos(np.pi / 4)], [0.3, -0.1, 0.2, 0.9]])
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    pos = np.array([[L / 2, L / 2, L / 2], [L / 2, L / 4, 0.3 * L], [0.25 * L, 0.75 * L, 0.7 * L]])
    otr = np.concatenate([sc.obstacle_tr, S.transform12(pos, q)])
    osh = np.concatenate([sc.obstacle_shape, np.arange(first, first + 3, dtype=np.uint32)])
    return dataclasses.replace(sc, shapes=shapes, obstacle_shape=osh, obstacle_tr=otr)


@pytest.mark.parametrize("flags", [0, A.REF_QUIRKS])
def test_general_obstacles_fp64(flags):
    """Obstacles of any convex kind go through the bounding-sphere, OBB and
    GJK tests against every particle of their simulation
    (ComponentManagerCPU.cpp:301-351).  Wall counts and GJK calls equal the
    oracle's; forces to the bars of the first-step test."""
    sc = with_obstacles(dense_cylinders_cones(n=300, seed=48))
    e, o = both(sc, flags=flags, tol=1e-12)
    e.step(1)
    o.step(1)
    se, so = e.get_stats(), o.get_stats()
    # (the reference's Johnson GJK stops at once on overlapping shapes, so
    # under REF_QUIRKS no obstacle contact of this scene is found)
    assert se["wall_contacts"] == so["wall_contacts"] and so["wall_contacts"] >= (0 if flags else 20)
    assert se["gjk_calls"] == so["gjk_calls"]
    np.testing.assert_array_equal(e.get_contact_count(), o.get_contact_count())
    fe, te = e.get_torce()
    fo, to = o.get_torce()
    bar = 1e-9 if flags & A.REF_QUIRKS else 10.0 * np.sqrt(2e-12)
    assert rel_err(fe, fo) < bar and rel_err(te, to) < bar
    # the obstacles' share: forces without them differ
    from oracle import pyoracle
    o0 = S.load(pyoracle.Oracle(dt=sc.dt, flags=flags, gjk_rel_tol=1e-12),
                dataclasses.replace(sc, obstacle_shape=sc.obstacle_shape[:2], obstacle_tr=sc.obstacle_tr[:2]))
    o0.step(1)
    if not flags:
        assert rel_err(o0.get_torce()[0], fo) > 1e-3


def test_general_obstacles_fp32_mixed():
    sc = with_obstacles(dense_cylinders_cones(n=300, seed=49))
    e, o = both(sc, precision=A.FP32_MIXED, tol=1e-6, oracle_tol=1e-12)
    e.step(1)
    o.step(1)
    se, so = e.get_stats(), o.get_stats()
    assert abs(int(se["wall_contacts"]) - int(so["wall_contacts"])) <= 2
    fe, _ = e.get_torce()
    fo, _ = o.get_torce()
    assert rel_err(fe, fo) < np.sqrt(2e-6)
