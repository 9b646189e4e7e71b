"""Parity of the device's time integration schemes (grains_config_t.integrator)
with the FP64 oracle, and the schemes' history across snapshots and graph
replays.

Bars: the first step of a sphere scene is bit-exact in positions and
velocities for every scheme (forces are bit-exact there and both sides use
the same update arithmetic); later steps inherit the ulp-level difference of
the orientations (the oracle keeps rotation matrices, the device
quaternions), so after 20 steps the bar is ten times the oracle's own spread
under a 1e-15 change of the initial positions."""
import dataclasses

import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S
from tests.scenes import dense_spheres, lattice_superquadrics, rel_err

pytestmark = pytest.mark.gpu

SCHEMES = [A.FIRST_ORDER_EXPLICIT, A.SECOND_ORDER_EXPLICIT, A.SECOND_ORDER_LEAPFROG,
           A.SECOND_ORDER_ADAMS_BASHFORTH]


def pair(sc, scheme, tol=0.0):
    from grains3dgpu_amd import engine
    from oracle import pyoracle
    e = S.load(engine.Engine(dt=sc.dt, integrator=scheme, gjk_rel_tol=tol), sc)
    o = S.load(pyoracle.Oracle(dt=sc.dt, integrator=scheme, gjk_rel_tol=tol), sc)
    return e, o


@pytest.mark.parametrize("scheme", SCHEMES)
def test_spheres(scheme):
    from oracle import pyoracle
    sc = dense_spheres(n=300, seed=23, walls=True)
    e, o = pair(sc, scheme)
    e.step(1)
    o.step(1)
    pe, _, ve, _ = e.get_state()
    po, _, vo, _ = o.get_state()
    np.testing.assert_array_equal(pe, po)
    np.testing.assert_array_equal(ve, vo)
    e.step(19)
    o.step(19)
    o2 = S.load(pyoracle.Oracle(dt=sc.dt, integrator=scheme), dataclasses.replace(sc, pos=sc.pos * (1 + 1e-15)))
    o2.step(20)
    pe, _, ve, we = e.get_state()
    po, _, vo, wo = o.get_state()
    p2, _, v2, w2 = o2.get_state()
    assert rel_err(pe, po) <= 10 * rel_err(p2, po) + 1e-15
    assert rel_err(ve, vo) <= 10 * rel_err(v2, vo) + 1e-13
    assert rel_err(we, wo) <= 10 * rel_err(w2, wo) + 1e-12


@pytest.mark.parametrize("scheme", [A.SECOND_ORDER_LEAPFROG, A.SECOND_ORDER_ADAMS_BASHFORTH])
def test_superquadric_lattice(scheme):
    """The superquadric lattice of the first-order trajectory test, 100
    steps: positions to 1e-9 of the box, as there."""
    sc = lattice_superquadrics()
    e, o = pair(sc, scheme, tol=1e-12)
    e.step(100)
    o.step(100)
    pe, _, ve, _ = e.get_state()
    po, _, vo, _ = o.get_state()
    L = sc.sims[0]["domain_max"][0]
    assert np.abs(pe - po).max() < 1e-9 * L
    assert rel_err(ve, vo) < 1e-6


@pytest.mark.parametrize("scheme", [A.SECOND_ORDER_LEAPFROG, A.SECOND_ORDER_ADAMS_BASHFORTH])
def test_history_in_snapshots_and_graphs(scheme):
    """The schemes' history travels with grains_snapshot / grains_restore and
    is carried through graph replays: restored and replayed runs equal the
    uninterrupted eager one bit for bit."""
    from grains3dgpu_amd import engine
    sc = dense_spheres(n=200, seed=72)
    e = S.load(engine.Engine(dt=sc.dt, integrator=scheme), sc)
    e.step(5)
    blob = e.snapshot()
    e.step(7)
    eager = e.get_state()
    e.restore(blob)
    e.step(7)
    for x, y in zip(eager, e.get_state()):
        np.testing.assert_array_equal(x, y)
    e.restore(blob)
    e.step(3, graph=True)
    e.step(4, graph=True)
    for x, y in zip(eager, e.get_state()):
        np.testing.assert_array_equal(x, y)
    # and the history matters: a first-order run from the snapshot differs
    e1 = S.load(engine.Engine(dt=sc.dt), sc)
    e1.step(5)
    e1.step(7)
    assert rel_err(e1.get_state()[0], eager[0]) > 0
