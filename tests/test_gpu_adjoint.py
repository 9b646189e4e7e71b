"""The adjoint sweep (grains_adjoint) against central finite differences of
the FP64 forward, through the C-ABI and the torch.autograd.Function.

Scene: a small config-4 bed (spheres and ellipsoids, two materials, three
contact-table entries, walls, gravity, random velocities) for 150 steps,
checkpoints every 40 (several blocks, the last one partial).  The loss is
linear in the final state (random weights), so dL/dstate is exact and the
check isolates the sweep.

Finite differences of a DEM trajectory are trustworthy only where they are
stable in the step: the adjoint differentiates each step at its fixed
contact set, while a parameter change of h can make a contact begin, end or
switch between sticking and sliding inside the trajectory (a kink in the
loss, felt by large h), and the GJK witness points of curved contacts
converge like the square root of the gap (here 1e-15: a noise that small h
amplifies).  Measured on this bed, no single step serves every component
(at 1e-3 the restitution of one entry is 5x off, at 1e-4 another entry's
friction 1 % off).  So each component's difference is taken at relative
steps 1e-2, 1e-3, 1e-4 and 1e-5; where two adjacent steps agree to 1e-3 of
the entry's largest sensitivity (dL / dlog p, so that kn and the
dimensionless en, mu compare), their mean is the reference, and the adjoint
must match every such reference to 2e-3 (measured: <= 1e-4); at least 7 of
the 9 components must have one."""
import numpy as np
import pytest

from grains3dgpu_amd import _abi as A
from grains3dgpu_amd import scene as S

pytestmark = pytest.mark.gpu

STEPS, EVERY = 150, 40


@pytest.fixture(scope="module")
def bed():
    """The floor (the first wall) is raised to 5 um above the lowest point of
    any particle, so that wall contacts are in the differentiated steps."""
    # 0.4-0.6 mm grains, so that all three table entries see contacts in 150 steps
    sc = S.config4(n=64, seed=5, vel=0.5, a_rng=(0.4e-3, 0.6e-3), jitter=0.05, layout="lattice")
    R = S.quat_to_matrix(sc.quat).reshape(-1, 3, 3)
    low = np.zeros(sc.n)
    for i, k in enumerate(sc.particle_shape):
        sh = sc.shapes[int(k)]
        d = np.asarray(sh["dims"], float)
        # extent along z: r for a sphere, |diag(a, b, c) R^T e_z| for an ellipsoid
        ext = d[0] if int(sh["kind"]) == A.SPHERE else float(np.linalg.norm(d[:3] * R[i, 2, :]))
        low[i] = sc.pos[i, 2] - ext
    tr = np.array(sc.obstacle_tr, copy=True)
    assert tr[0, 11] == sc.sims[0]["domain_min"][2]  # the floor: normal +z, origin on z = min
    tr[0, 11] = low.min() + 5e-6
    sc.obstacle_tr = tr
    return sc


def run(sc, params, precision=A.FP64):
    from grains3dgpu_amd import adjoint, engine
    e = S.load(engine.Engine(dt=sc.dt, precision=precision, gjk_rel_tol=1e-15), sc)
    s2 = adjoint.with_params(sc, params)
    e.set_contact_tables(s2)
    e.set_state(s2.pos, s2.quat, s2.vel, s2.omega)
    e.adjoint_record(EVERY)
    e.step(STEPS)
    return e


def weights(n, seed=0):
    r = np.random.default_rng(seed)
    return r.standard_normal((n, 3)), r.standard_normal((n, 4)) * 1e-3, r.standard_normal((n, 3)) * 1e-3, \
        r.standard_normal((n, 3)) * 1e-6


def loss(state, w):
    return sum(float((a * b).sum()) for a, b in zip(state, w))


def base_params(sc):
    return np.array(sc.tables)[:, [0, 1, 3]].copy()


def test_parameter_gradient_matches_finite_differences(bed):
    p0 = base_params(bed)
    w = weights(bed.n)
    e = run(bed, p0)
    st = e.get_stats()
    assert st["contacts"] > 50 and st["wall_contacts"] > 0
    g = e.adjoint(*w)[:, [0, 1, 3]]
    e.close()
    steps = (1e-2, 1e-3, 1e-4, 1e-5)
    fd = np.zeros((len(steps),) + p0.shape)
    for k, rel in enumerate(steps):
        for r in range(p0.shape[0]):
            for c in range(3):
                h = rel * p0[r, c]
                pp, pm = p0.copy(), p0.copy()
                pp[r, c] += h
                pm[r, c] -= h
                ep, em = run(bed, pp), run(bed, pm)
                fd[k, r, c] = (loss(ep.get_state(), w) - loss(em.get_state(), w)) / (2 * h)
                ep.close()
                em.close()
    # compared as sensitivities to relative changes, dL / dlog p, so that kn
    # (N/m) and en, mu (dimensionless) share one scale per entry
    gs, fs = g * p0, fd * p0
    scale = np.maximum(np.abs(gs), np.abs(fs).max(axis=0)).max(axis=1)
    assert (scale > 0).all()
    trusted, bad = 0, []
    for r in range(p0.shape[0]):
        for c in range(3):
            for k in range(len(steps) - 1):
                if abs(fs[k, r, c] - fs[k + 1, r, c]) <= 1e-3 * scale[r]:
                    ref = 0.5 * (fs[k, r, c] + fs[k + 1, r, c])
                    trusted += 1
                    if abs(gs[r, c] - ref) > 2e-3 * scale[r]:
                        bad.append((r, c, steps[k], gs[r, c], ref))
                    break
    assert trusted >= 7, (g, fd)
    assert not bad, (bad, g, fd)


def test_initial_state_gradient(bed):
    p0 = base_params(bed)
    w = weights(bed.n, 1)
    e = run(bed, p0)
    _, (dpos0, dq0, dv0, dw0) = e.adjoint(*w, want_state0=True)
    e.close()
    from grains3dgpu_amd import engine
    rng = np.random.default_rng(2)
    for i in rng.choice(bed.n, 4, replace=False):
        for k in range(3):
            h = 1e-8
            vals = []
            for sgn in (1, -1):
                sc = S.Scene(**{f: getattr(bed, f) for f in bed.__dataclass_fields__})
                sc.pos = bed.pos.copy()
                sc.pos[i, k] += sgn * h
                ee = run(sc, p0)
                vals.append(loss(ee.get_state(), w))
                ee.close()
            fdv = (vals[0] - vals[1]) / (2 * h)
            assert abs(fdv - dpos0[i, k]) <= 5e-3 * max(abs(fdv), abs(dpos0).max() * 1e-2), (i, k, fdv, dpos0[i, k])
    del engine


def test_autograd_function_and_reproducibility(bed):
    import torch
    from grains3dgpu_amd import adjoint, engine
    p0 = base_params(bed)
    w = [torch.from_numpy(a) for a in weights(bed.n, 3)]
    e = S.load(engine.Engine(dt=bed.dt, precision=A.FP64, gjk_rel_tol=1e-15), bed)
    grads = []
    for _ in range(2):
        p = torch.tensor(p0, requires_grad=True)
        out = adjoint.simulate(e, bed, p, STEPS, EVERY)
        L = sum((a * b).sum() for a, b in zip(out, w))
        L.backward()
        grads.append(p.grad.numpy().copy())
    np.testing.assert_array_equal(grads[0], grads[1])
    # the same as the C-ABI sweep on a fresh engine
    e2 = run(bed, p0)
    g = e2.adjoint(*[a.numpy() for a in w])[:, [0, 1, 3]]
    np.testing.assert_array_equal(grads[0], g)
    e.close()
    e2.close()
