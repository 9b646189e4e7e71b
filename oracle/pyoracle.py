"""ctypes wrapper of the FP64 CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, the smoke check and bench.py's CPU-baseline leg may import this
module; the product never does.  It exposes the same methods as
grains3dgpu_amd.engine.Engine so a test drives both with one scene.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from grains3dgpu_amd import _abi as A

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "libgrains_oracle.so")
_lib = None

u32p = C.POINTER(C.c_uint32)


def build(force=False):
    if force or not os.path.exists(_LIB):
        subprocess.check_call(["make", "-s", "-C", _HERE])
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        vp = C.c_void_p
        L.go_create.restype = vp
        L.go_create.argtypes = [C.POINTER(A.Config)]
        L.go_destroy.argtypes = [vp]
        L.go_last_error.restype = C.c_char_p
        L.go_last_error.argtypes = [vp]
        L.go_set_shapes.argtypes = [vp, vp, C.c_uint32, vp, C.c_uint32]
        L.go_set_contact_tables.argtypes = [vp, vp, C.c_uint32, C.c_uint32]
        L.go_set_sims.argtypes = [vp, vp, C.c_uint32]
        L.go_set_obstacles.argtypes = [vp, vp, vp, vp, C.c_uint32]
        L.go_set_particles.argtypes = [vp, vp, vp, C.c_uint64]
        L.go_set_state.argtypes = [vp, vp, vp, vp, vp, C.c_uint64]
        L.go_get_state.argtypes = [vp, vp, vp, vp, vp, C.c_uint64]
        L.go_set_transforms.argtypes = [vp, vp, C.c_uint64]
        L.go_get_transforms.argtypes = [vp, vp, C.c_uint64]
        L.go_step.argtypes = [vp, C.c_uint32, C.c_int]
        L.go_get_cell_keys.argtypes = [vp, vp, vp, C.c_uint64]
        L.go_get_cell_layout.argtypes = [vp, C.c_uint32, vp, C.POINTER(C.c_double), u32p]
        L.go_get_pairs.argtypes = [vp, C.c_int, vp, C.c_uint64, C.POINTER(C.c_uint64)]
        L.go_get_contact_count.argtypes = [vp, vp, C.c_uint64]
        L.go_get_torce.argtypes = [vp, vp, vp, C.c_uint64]
        L.go_get_stats.argtypes = [vp, C.POINTER(A.Stats)]
        L.go_shape_props.argtypes = [C.POINTER(A.Shape), vp, vp]
        L.go_support.argtypes = [C.POINTER(A.Shape), vp, vp, vp]
        L.go_cell_counts.argtypes = [vp, vp, C.c_double, vp]
        L.go_neighbor_table.argtypes = [vp, vp]
        L.go_cell_hash.restype = C.c_uint64
        L.go_cell_hash.argtypes = [vp, C.c_double, vp, vp, C.c_uint64, vp]
        L.go_gjk.restype = C.c_double
        L.go_gjk.argtypes = [C.POINTER(A.Shape), C.POINTER(A.Shape), vp, vp, vp, C.c_uint32,
                             C.c_double, vp, vp, C.POINTER(C.c_int)]
        L.go_closest_points.argtypes = [C.POINTER(A.Shape), C.POINTER(A.Shape), vp, vp, vp,
                                        C.c_uint32, C.c_double, vp]
        L.go_hooke.argtypes = [C.POINTER(A.Hooke), vp, vp, vp, C.c_double, C.c_double, vp, vp]
        L.go_quat_from_matrix.argtypes = [vp, vp]
        L.go_matrix_from_quat.argtypes = [vp, vp]
        L.go_basis_from_angles.argtypes = [C.c_double, C.c_double, C.c_double, vp]
        _lib = L
    return _lib


def _c(a, dtype=np.float64):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


class Oracle:
    """FP64 CPU oracle world; method names follow grains3dgpu_amd.engine.Engine."""

    def __init__(self, dt=1e-6, precision=A.FP64, flags=0, gjk_rel_tol=0.0, threads=0,
                 integrator=A.FIRST_ORDER_EXPLICIT):
        self.L = lib()
        cfg = A.Config()
        cfg.dt, cfg.precision, cfg.flags, cfg.gjk_rel_tol = dt, precision, flags, gjk_rel_tol
        cfg.integrator = integrator
        self.w = self.L.go_create(C.byref(cfg))
        self.threads = threads
        self.n = 0
        self._keep = []

    def close(self):
        if self.w:
            self.L.go_destroy(self.w)
            self.w = None

    __del__ = close

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {self.L.go_last_error(self.w).decode()}")

    # scene -------------------------------------------------------------
    def set_shapes(self, sc):
        arr = sc.shape_array()
        v = _c(sc.vertices.reshape(-1, 3)) if len(sc.vertices) else None
        self._keep += [arr, v]
        self._chk(self.L.go_set_shapes(self.w, A.aptr(arr), len(sc.shapes), A.dptr(v),
                                       0 if v is None else v.shape[0]))

    def set_contact_tables(self, sc):
        arr = sc.table_array()
        self._chk(self.L.go_set_contact_tables(self.w, C.cast(arr, C.c_void_p), sc.n_materials,
                                               sc.n_tables))

    def set_sims(self, sc):
        arr = sc.sim_array()
        self._chk(self.L.go_set_sims(self.w, C.cast(arr, C.c_void_p), len(sc.sims)))

    def set_obstacles(self, sc):
        sh = _c(sc.obstacle_shape, np.uint32)
        sm = _c(sc.obstacle_sim, np.uint32)
        tr = _c(sc.obstacle_tr)
        self._chk(self.L.go_set_obstacles(self.w, A.dptr(sh), A.dptr(sm), A.dptr(tr), len(sh)))

    def set_particles(self, sc):
        sh = _c(sc.particle_shape, np.uint32)
        sm = _c(sc.particle_sim, np.uint32)
        self.n = len(sh)
        self._chk(self.L.go_set_particles(self.w, A.dptr(sh), A.dptr(sm), self.n))

    def set_state(self, pos=None, quat=None, vel=None, omega=None):
        p, q, v, o = _c(pos), _c(quat), _c(vel), _c(omega)
        self._chk(self.L.go_set_state(self.w, A.dptr(p), A.dptr(q), A.dptr(v), A.dptr(o), self.n))

    def set_transforms(self, tr):
        t = _c(tr)
        self._chk(self.L.go_set_transforms(self.w, A.dptr(t), self.n))

    # run / results -----------------------------------------------------------
    def step(self, n_steps=1):
        self._chk(self.L.go_step(self.w, n_steps, self.threads))

    def get_state(self):
        n = self.n
        p, q, v, o = (np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3)))
        self._chk(self.L.go_get_state(self.w, A.dptr(p), A.dptr(q), A.dptr(v), A.dptr(o), n))
        return p, q, v, o

    def get_transforms(self):
        t = np.zeros((self.n, 12))
        self._chk(self.L.go_get_transforms(self.w, A.dptr(t), self.n))
        return t

    def get_cell_keys(self):
        k = np.zeros(self.n, np.uint32)
        s = np.zeros(self.n, np.uint32)
        self._chk(self.L.go_get_cell_keys(self.w, A.dptr(k), A.dptr(s), self.n))
        return k, s

    def get_cell_layout(self, sim=0):
        n3 = np.zeros(3, np.uint32)
        h = C.c_double()
        base = C.c_uint32()
        self._chk(self.L.go_get_cell_layout(self.w, sim, A.dptr(n3), C.byref(h), C.byref(base)))
        return n3, h.value, base.value

    def get_pairs(self, which=A.PAIRS_CANDIDATE):
        m = C.c_uint64()
        self._chk(self.L.go_get_pairs(self.w, which, None, 0, C.byref(m)))
        out = np.zeros((m.value, 2), np.uint32)
        self._chk(self.L.go_get_pairs(self.w, which, A.dptr(out), m.value, C.byref(m)))
        return out

    def get_contact_count(self):
        c = np.zeros(self.n, np.int32)
        self._chk(self.L.go_get_contact_count(self.w, A.dptr(c), self.n))
        return c

    def get_torce(self):
        f = np.zeros((self.n, 3))
        t = np.zeros((self.n, 3))
        self._chk(self.L.go_get_torce(self.w, A.dptr(f), A.dptr(t), self.n))
        return f, t

    def get_stats(self):
        st = A.Stats()
        self._chk(self.L.go_get_stats(self.w, C.byref(st)))
        return st.as_dict()


# ---- primitives ---------------------------------------------------------------
def _shape(s):
    sh = A.Shape()
    sh.kind = s["kind"]
    sh.material = s.get("material", 0)
    d = list(s["dims"]) + [0.0] * (5 - len(s["dims"]))
    for k in range(5):
        sh.dims[k] = d[k]
    sh.crust = s.get("crust", 0.0)
    sh.density = s.get("density", 0.0)
    sh.n_vertices = s.get("n_vertices", 0)
    sh.vertex_offset = s.get("vertex_offset", 0)
    return sh


def shape_props(s, vertices=None):
    out = np.zeros(19)
    v = _c(vertices)
    rc = lib().go_shape_props(C.byref(_shape(s)), A.dptr(v), A.dptr(out))
    if rc:
        raise RuntimeError(rc)
    return dict(volume=out[0], mass=out[1], inertia=out[2:8].copy(), inertia_inv=out[8:14].copy(),
                rc=out[14], bbox=out[15:18].copy(), rc_exact=out[18])


def support(s, d, vertices=None):
    out = np.zeros(3)
    d = _c(d)
    v = _c(vertices)
    rc = lib().go_support(C.byref(_shape(s)), A.dptr(v), A.dptr(d), A.dptr(out))
    if rc:
        raise RuntimeError(rc)
    return out


def cell_counts(lo, hi, h):
    n3 = np.zeros(3, np.uint32)
    lib().go_cell_counts(A.dptr(_c(lo)), A.dptr(_c(hi)), h, A.dptr(n3))
    return n3


def neighbor_table(n3):
    n3 = _c(n3, np.uint32)
    t = np.zeros(int(np.prod(n3)) * 27, np.uint32)
    lib().go_neighbor_table(A.dptr(n3), A.dptr(t))
    return t.reshape(-1, 27)


def cell_hash(lo, h, n3, pos):
    pos = _c(pos)
    keys = np.zeros(len(pos), np.uint32)
    bad = lib().go_cell_hash(A.dptr(_c(lo)), h, A.dptr(_c(n3, np.uint32)), A.dptr(pos), len(pos),
                             A.dptr(keys))
    return keys, int(bad)


def gjk(sa, sb, ta, tb, flags=0, rel_tol=0.0, vertices=None):
    pa, pb = np.zeros(3), np.zeros(3)
    it = C.c_int()
    v = _c(vertices)
    d = lib().go_gjk(C.byref(_shape(sa)), C.byref(_shape(sb)), A.dptr(v), A.dptr(_c(ta)),
                     A.dptr(_c(tb)), flags, rel_tol, A.dptr(pa), A.dptr(pb), C.byref(it))
    return d, pa, pb, it.value


def closest_points(sa, sb, ta, tb, flags=0, vertices=None, rel_tol=0.0):
    out = np.zeros(7)
    v = _c(vertices)
    hit = lib().go_closest_points(C.byref(_shape(sa)), C.byref(_shape(sb)), A.dptr(v),
                                  A.dptr(_c(ta)), A.dptr(_c(tb)), flags, C.c_double(rel_tol), A.dptr(out))
    return bool(hit), out[:3].copy(), out[3:6].copy(), out[6]


def hooke(params, cp, cv, overlap, rel_vel, rel_ang_vel, m1, m2, origin_a):
    h = A.Hooke(*params)
    ci = _c(np.concatenate([cp, cv, [overlap]]))
    out = np.zeros(6)
    lib().go_hooke(C.byref(h), A.dptr(ci), A.dptr(_c(rel_vel)), A.dptr(_c(rel_ang_vel)), m1, m2,
                   A.dptr(_c(origin_a)), A.dptr(out))
    return out[:3].copy(), out[3:].copy()


def quat_from_matrix(m9):
    q = np.zeros(4)
    lib().go_quat_from_matrix(A.dptr(_c(m9)), A.dptr(q))
    return q


def matrix_from_quat(q):
    m = np.zeros(9)
    lib().go_matrix_from_quat(A.dptr(_c(q)), A.dptr(m))
    return m


def basis_from_angles(ax, ay, az):
    m = np.zeros(9)
    lib().go_basis_from_angles(ax, ay, az, A.dptr(m))
    return m
