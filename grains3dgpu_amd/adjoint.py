"""Differentiable trajectories: the library's hand-written adjoint
(grains_adjoint_record / grains_adjoint, over the C-ABI) behind a
torch.autograd.Function.

    final = simulate(engine, scene, params, n_steps, checkpoint_every=100)

`params` is a torch tensor (contact-table entries x 3) of (kn, en, muc) --
stiffness, restitution (which sets the normal damping) and friction -- in
the row order of the scene's contact tables; the tangential damping and
kr stay as the scene holds them.  `final` is (pos, quat, vel, omega) after
n_steps from the scene's state, as float64 tensors; any loss of them
backpropagates to `params` through one reverse sweep on the device, which
re-runs every block of `checkpoint_every` steps from its checkpoint.

Tensors are host memory (the engine's boundary takes host buffers); torch is
only the autograd plumbing here.
"""
import numpy as np
import torch

from . import _abi as A
from . import scene as S


def with_params(scene, params):
    """A copy of `scene` whose contact tables carry params (entries x 3:
    kn, en, muc)."""
    p = np.asarray(params, np.float64)
    tables = np.array(scene.tables, np.float64, copy=True)
    assert p.shape == (tables.shape[0], 3), (p.shape, tables.shape)
    tables[:, 0] = p[:, 0]
    tables[:, 1] = p[:, 1]
    tables[:, 3] = p[:, 2]
    sc = S.Scene(**{k: getattr(scene, k) for k in scene.__dataclass_fields__})
    sc.tables = tables
    return sc


class Trajectory(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, engine, scene, n_steps, checkpoint_every, graph):
        sc = with_params(scene, params.detach().cpu().numpy())
        engine.set_contact_tables(sc)
        engine.set_state(sc.pos, sc.quat, sc.vel, sc.omega)
        engine.adjoint_record(checkpoint_every)
        engine.step(n_steps, graph=graph)
        ctx.engine = engine
        ctx.n = sc.n
        return tuple(torch.from_numpy(np.ascontiguousarray(a)) for a in engine.get_state())

    @staticmethod
    def backward(ctx, gpos, gquat, gvel, gomega):
        n = ctx.n

        def arr(g, k):
            return np.zeros((n, k)) if g is None else g.detach().cpu().numpy().astype(np.float64)
        d = ctx.engine.adjoint(arr(gpos, 3), arr(gquat, 4), arr(gvel, 3), arr(gomega, 3))
        # columns kn, en, etat, muc -> the parameters kn, en, muc
        g = torch.from_numpy(np.ascontiguousarray(d[:, [0, 1, 3]]))
        return g, None, None, None, None, None


def simulate(engine, scene, params, n_steps, checkpoint_every=100, graph=True):
    """(pos, quat, vel, omega) after n_steps from scene's state with contact
    parameters `params` (entries x 3: kn, en, muc), differentiable in params."""
    return Trajectory.apply(params, engine, scene, n_steps, checkpoint_every, graph)


def masses(scene):
    """Particle masses of a scene of spheres and ellipsoids (superquadrics
    with n1 = n2 = 2), the shapes the adjoint covers: rho * 4/3 pi r^3 or
    rho * 4/3 pi a b c."""
    out = np.zeros(scene.n)
    for i, s in enumerate(scene.particle_shape):
        sh = scene.shapes[int(s)]
        d = sh["dims"]
        v = 4.0 / 3.0 * np.pi * (d[0] ** 3 if int(sh["kind"]) == A.SPHERE else d[0] * d[1] * d[2])
        out[i] = sh["density"] * v
    return out


def bed_loss(pos, vel, mass, target):
    """Calibration loss of config 4: relative squared errors of the bed
    height (mean centre height) and the mean translational kinetic energy
    against a target (h*, KE*)."""
    h = pos[:, 2].mean()
    ke = (0.5 * mass * (vel * vel).sum(1)).mean()
    return ((h - target[0]) / target[0]) ** 2 + ((ke - target[1]) / target[1]) ** 2, h, ke
