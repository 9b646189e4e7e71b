"""Edge cases of the HIP path against the FP64 oracle and of the C-ABI's
error behaviour: a lone particle, a sim without particles inside an
ensemble, particles outside the domain and on cell boundaries, the
candidate capacity overflow, and configurations the ABI rejects."""

import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S
from tests.scenes import dense_spheres, rel_err
from tests.test_gpu_parity import both, check_broad_phase

pytestmark = pytest.mark.gpu


def sphere_scene(pos, vel=None, r=0.5e-3, lo=(0, 0, 0), hi=(5e-3,) * 3, g=(0, 0, -9.81), sims=None,
                 particle_sim=None):
    n = len(pos)
    return S.Scene(shapes=[dict(kind=A.SPHERE, material=0, dims=(r,), density=2500.0, crust=2.5e-5)],
                   tables=np.array([S.hooke_row(1e5, 0.8, 2e4, 0.3)] * (len(sims) if sims else 1)),
                   n_materials=1, sims=sims or [dict(domain_min=lo, domain_max=hi, gravity=g)],
                   particle_shape=np.zeros(n, np.uint32), particle_sim=particle_sim,
                   pos=np.asarray(pos, float), quat=np.tile([0, 0, 0, 1.0], (n, 1)),
                   vel=np.zeros((n, 3)) if vel is None else np.asarray(vel, float),
                   omega=np.zeros((n, 3)), dt=1e-6)


def test_lone_particle():
    sc = sphere_scene([[2.5e-3, 2.5e-3, 2.5e-3]], vel=[[0.1, -0.2, 0.3]])
    e, o = both(sc)
    e.step(25)
    o.step(25)
    check_broad_phase(e, o)
    assert len(e.get_pairs(A.PAIRS_CANDIDATE)) == 0
    st = e.get_stats()
    assert st["candidate_pairs"] == 0 and st["contacts"] == 0 and st["gjk_calls"] == 0
    for a, b in zip(e.get_state()[::2], o.get_state()[::2]):   # positions and velocities
        np.testing.assert_array_equal(a, b)


def test_ensemble_with_an_empty_sim():
    """Three sims, the middle one without particles: its cells exist, nothing
    visits them, and the others step as on their own.  (A sim's cell size
    derives from its particles' radii, Grains.cpp:226-245; an empty sim
    takes an explicit one.)"""
    a = dense_spheres(n=120, seed=61)
    c = dense_spheres(n=90, seed=62)
    L = a.sims[0]["domain_max"][0]
    empty = dict(domain_min=(0, 0, 0), domain_max=(L, L, L), gravity=(0, 0, -9.81), cell_size=L / 4)
    sc = S.Scene(shapes=a.shapes, tables=np.concatenate([a.tables] * 3), n_materials=1,
                 sims=[a.sims[0], dict(empty, contact_table=1), dict(c.sims[0], contact_table=2)],
                 particle_shape=np.concatenate([a.particle_shape, c.particle_shape]),
                 particle_sim=np.concatenate([np.zeros(a.n, np.uint32), np.full(c.n, 2, np.uint32)]),
                 pos=np.concatenate([a.pos, c.pos]), quat=np.concatenate([a.quat, c.quat]),
                 vel=np.concatenate([a.vel, c.vel]), omega=np.concatenate([a.omega, c.omega]))
    e, o = both(sc)
    e.step(5)
    o.step(5)
    check_broad_phase(e, o)
    pe, _, _, _ = e.get_state()
    po, _, _, _ = o.get_state()
    assert rel_err(pe, po) < 1e-13


def test_outside_the_domain_and_on_cell_boundaries():
    """Centres beyond the domain fold into the border cells and are counted
    (out_of_domain: three here; the maximum corner itself is inside);
    centres exactly on cell faces and on the maximum corner hash like the
    oracle's; the pairs and the state follow the oracle."""
    h = 1e-3
    pts = [[-0.3e-3, 2.5e-3, 2.5e-3], [5.4e-3, 2.5e-3, 2.5e-3], [2.5e-3, 2.5e-3, 6.0e-3],
           [2.0e-3, 3.0e-3, 1.0e-3], [3.0e-3, 3.0e-3, 1.0e-3], [5e-3, 5e-3, 5e-3],
           [1.0e-3, 1.0e-3, 4.0e-3], [1.0e-3 + 0.85e-3, 1.0e-3, 4.0e-3]]
    sc = sphere_scene(pts, r=0.45e-3, g=(0, 0, 0))
    sc.sims[0]["cell_size"] = h
    e, o = both(sc)
    e.step(1)
    o.step(1)
    check_broad_phase(e, o)
    assert e.get_stats()["out_of_domain"] == o.get_stats()["out_of_domain"] == 3
    assert len(e.get_pairs(A.PAIRS_CONTACT)) == 2   # the overlapping pair, both ways
    for a, b in zip(e.get_state()[::2], o.get_state()[::2]):
        np.testing.assert_array_equal(a, b)


def test_candidate_capacity_overflow():
    """More candidates than max_neighbors: the step fails with
    GRAINS_EOVERFLOW and names the count needed; with that capacity the
    same scene steps like the oracle."""
    from grains3dgpu_amd import engine
    sc = dense_spheres(n=300, seed=63)
    e = S.load(engine.Engine(dt=sc.dt, max_neighbors=4), sc)
    with pytest.raises(engine.GrainsError) as err:
        e.step(1)
    assert "max_neighbors" in str(err.value)
    need = e.get_stats()["max_candidates"]
    assert need > 4
    e.close()
    e, o = both(sc, max_neighbors=int(need))
    e.step(1)
    o.step(1)
    check_broad_phase(e, o)


def test_rejected_configurations():
    from grains3dgpu_amd import engine
    for kw in (dict(integrator=7), dict(narrow_passes=3), dict(candidate_staging=5), dict(block_size=100)):
        with pytest.raises(engine.GrainsError):
            engine.Engine(**kw)
    e = engine.Engine()
    bad = sphere_scene([[1e-3, 1e-3, 1e-3]])
    bad.shapes = [dict(kind=A.SPHERE, material=0, dims=(0.5e-3,), density=0.0)]
    with pytest.raises(engine.GrainsError):
        S.load(e, bad)   # a particle needs mass > 0
    e.close()
    e = engine.Engine()
    wall = sphere_scene([[1e-3, 1e-3, 1e-3]])
    wall.shapes = [dict(kind=A.RECTANGLE, material=0, dims=(1e-3, 1e-3), density=1.0)]
    with pytest.raises(engine.GrainsError):
        S.load(e, wall)  # a particle cannot be a Rectangle
    e.close()
