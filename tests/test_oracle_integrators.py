"""Pins of the oracle's time integration schemes (grains_config_t.integrator):
FirstOrderExplicit (FirstOrderExplicit.cpp:46-60) and the three second-order
schemes TimeIntegratorBuilderFactory.cpp:45-49 names (not built by the
reference; their updates are stated in grains_hip.h).

Closed forms under constant acceleration (free fall), convergence orders on a
harmonic contact oscillator (a sphere held on a wall by gravity: the force
depends on the position only) and on torque-free rotation of a box (the
angular acceleration depends on the angular velocity), and the history reset
of the two-step and staggered schemes."""
import math

import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S

SCHEMES = [A.FIRST_ORDER_EXPLICIT, A.SECOND_ORDER_EXPLICIT, A.SECOND_ORDER_LEAPFROG,
           A.SECOND_ORDER_ADAMS_BASHFORTH]


def oracle(sc, scheme):
    from oracle import pyoracle
    return S.load(pyoracle.Oracle(dt=sc.dt, integrator=scheme), sc)


def single(shape, pos, vel, omega, g, dt, quat=(0, 0, 0, 1.0), walls=False, kn=1e5):
    shapes = [shape]
    osh, otr = np.zeros(0, np.uint32), np.zeros((0, 12))
    if walls:
        shapes.append(dict(kind=A.RECTANGLE, dims=(4e-3, 4e-3), density=0.0))
        osh, otr = np.array([1], np.uint32), np.array([[1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0.0]])
    return S.Scene(shapes=shapes, tables=np.array([S.hooke_row(kn, 1.0, 0.0, 0.0)]), n_materials=1,
                   sims=[dict(domain_min=(-5e-3,) * 3, domain_max=(5e-3,) * 3, gravity=g)],
                   particle_shape=np.zeros(1, np.uint32), pos=np.array([pos], float),
                   quat=np.array([quat], float) / np.linalg.norm(quat), vel=np.array([vel], float),
                   omega=np.array([omega], float), obstacle_shape=osh, obstacle_tr=otr, dt=dt)


@pytest.mark.parametrize("scheme", SCHEMES)
def test_free_fall_closed_forms(oracle_lib, scheme):
    g = np.array([0.0, 0.0, -9.81])
    x0, v0 = np.array([1e-3, -2e-3, 0.0]), np.array([0.5, -0.2, 3.0])
    dt, N = 1e-5, 40
    sc = single(dict(kind=A.SPHERE, dims=(0.5e-3,), density=2500.0), x0, v0, (0, 0, 0), tuple(g), dt)
    o = oracle(sc, scheme)
    o.step(N)
    x, _, v, _ = o.get_state()
    t = N * dt
    if scheme == A.FIRST_ORDER_EXPLICIT:       # x_N = x0 + v0 t + g dt^2 N (N - 1) / 2
        xe, ve = x0 + v0 * t + g * dt * dt * N * (N - 1) / 2, v0 + g * t
    elif scheme == A.SECOND_ORDER_EXPLICIT:    # exact
        xe, ve = x0 + v0 * t + g * t * t / 2, v0 + g * t
    elif scheme == A.SECOND_ORDER_LEAPFROG:    # exact; the velocity lives at t - dt/2
        xe, ve = x0 + v0 * t + g * t * t / 2, v0 + g * (t - dt / 2)
    else:                                      # exact after the first-order first step
        xe, ve = x0 + v0 * t + g * t * t / 2 - g * dt * dt / 2, v0 + g * t
    np.testing.assert_allclose(x[0], xe, rtol=1e-12, atol=1e-17)
    np.testing.assert_allclose(v[0], ve, rtol=1e-12, atol=1e-15)


def oscillator_error(scheme, per_period, periods=2.25):
    """A 0.5 mm sphere held on a wall by g = -1e5 m/s^2, released one
    amplitude below equilibrium: z(t) = z_eq - A cos(w0 t) while in contact
    (the overlap never reaches 0).  Error at a quarter phase, where the
    phase error shows to first order."""
    r, rho, kn, g = 0.5e-3, 2500.0, 1e5, -1e5
    m = rho * 4 / 3 * math.pi * r ** 3
    w0 = math.sqrt(kn / m)
    amp = -m * g / kn
    z_eq = r - amp
    dt = 2 * math.pi / w0 / per_period
    sc = single(dict(kind=A.SPHERE, dims=(r,), density=rho), (0, 0, z_eq - amp), (0, 0, 0), (0, 0, 0),
                (0, 0, g), dt, walls=True, kn=kn)
    o = oracle(sc, scheme)
    n = int(round(periods * per_period))
    o.step(n)
    return abs(o.get_state()[0][0, 2] - (z_eq - amp * math.cos(w0 * n * dt)))


@pytest.mark.parametrize("scheme", [A.SECOND_ORDER_LEAPFROG, A.SECOND_ORDER_ADAMS_BASHFORTH])
def test_second_order_on_a_harmonic_contact(oracle_lib, scheme):
    e = [oscillator_error(scheme, 200 * 2 ** k) for k in range(3)]
    assert 3.7 < e[0] / e[1] < 4.3 and 3.7 < e[1] / e[2] < 4.3, e


def rotation_error(scheme, dt, ref):
    T = 2e-3
    sc = single(dict(kind=A.BOX, dims=(0.3e-3, 0.5e-3, 0.8e-3), density=2500.0), (0, 0, 0), (0.1, 0, 0),
                (10.0, 200.0, 5.0), (0, 0, 0), dt, quat=(0.1, 0.2, 0.3, 0.9))
    o = oracle(sc, scheme)
    o.step(int(round(T / dt)))
    R = o.get_transforms()[0][:9]
    return R if ref is None else np.abs(R - ref).max()


@pytest.mark.parametrize("scheme,order", [(A.FIRST_ORDER_EXPLICIT, 1), (A.SECOND_ORDER_EXPLICIT, 1),
                                          (A.SECOND_ORDER_LEAPFROG, 1), (A.SECOND_ORDER_ADAMS_BASHFORTH, 2)])
def test_orders_on_torque_free_rotation(oracle_lib, scheme, order):
    """Euler's equations make the angular acceleration a function of the
    angular velocity: only the two-step scheme stays second order (the
    Taylor step and the staggered step integrate w by Euler steps)."""
    ref = rotation_error(scheme, 1e-5 / 64, None)
    e = [rotation_error(scheme, 1e-5 / 2 ** k, ref) for k in range(3)]
    lo, hi = (1.8, 2.3) if order == 1 else (3.7, 4.3)
    assert lo < e[0] / e[1] < hi and lo < e[1] / e[2] < hi, e


@pytest.mark.parametrize("scheme", [A.SECOND_ORDER_LEAPFROG, A.SECOND_ORDER_ADAMS_BASHFORTH])
def test_set_state_restarts_the_history(oracle_lib, scheme):
    """After grains_set_state the next step starts the scheme afresh: the
    two-step scheme takes a first-order step, the staggered one a half kick."""
    sc = single(dict(kind=A.BOX, dims=(0.3e-3, 0.5e-3, 0.8e-3), density=2500.0), (0, 0, 0), (0.1, 0, 0.2),
                (10.0, 200.0, 5.0), (0, 0, -9.81), 1e-5, quat=(0.1, 0.2, 0.3, 0.9))
    o = oracle(sc, scheme)
    o.step(10)
    st = o.get_state()
    o.set_state(*st)
    o.step(1)
    # the same first step from that state in a fresh world
    o2 = oracle(sc, scheme)
    o2.set_state(*st)
    o2.step(1)
    for a, b in zip(o.get_state(), o2.get_state()):
        np.testing.assert_array_equal(a, b)
    if scheme == A.SECOND_ORDER_ADAMS_BASHFORTH:
        o1 = oracle(sc, A.FIRST_ORDER_EXPLICIT)
        o1.set_state(*st)
        o1.step(1)
        for a, b in zip(o.get_state(), o1.get_state()):
            np.testing.assert_array_equal(a, b)
