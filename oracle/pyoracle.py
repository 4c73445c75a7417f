"""ctypes access to the CPU restatement (oracle/build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker or as the timed CPU baseline.
The product path never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liboracle.so")

_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_UP = C.POINTER(C.c_uint32)
_P = C.c_void_p

SIGNATURES = {
    "oracle_create": (C.c_int, [C.c_int, _P, C.c_int, _FP, C.c_int, _IP, C.POINTER(_P)]),
    "oracle_destroy": (None, [_P]),
    "oracle_configure": (C.c_int, [_P, C.c_float]),
    "oracle_set_young": (C.c_int, [_P, C.c_float, C.c_float]),
    "oracle_set_kstretch": (C.c_int, [_P, C.c_float]),
    "oracle_set_kbend": (C.c_int, [_P, C.c_float]),
    "oracle_pin_vertices": (C.c_int, [_P, _IP, C.c_int]),
    "oracle_fix_orientation": (C.c_int, [_P, _IP, C.c_int]),
    "oracle_sag_free": (C.c_int, [_P, _IP, C.c_int, C.c_int]),
    "oracle_set_params": (C.c_int, [_P, _P]),
    "oracle_upload": (C.c_int, [_P]),
    "oracle_set_pos_vel": (C.c_int, [_P, _FP, _FP]),
    "oracle_get_pos_vel": (C.c_int, [_P, _FP, _FP]),
    "oracle_set_last_vel": (C.c_int, [_P, _FP]),
    "oracle_set_connections": (C.c_int, [_P, _IP]),
    "oracle_get_pos_vel64": (C.c_int, [_P, _DP, _DP]),
    "oracle_get_q": (C.c_int, [_P, _FP]),
    "oracle_get_q64": (C.c_int, [_P, _DP]),
    "oracle_set_q": (C.c_int, [_P, _FP]),
    "oracle_get_rest": (C.c_int, [_P, _FP, _FP, _FP, _UP, _FP]),
    "oracle_set_sdf": (C.c_int, [_P, _IP, _FP, _FP, _FP, _FP, C.c_float, _FP, _FP]),
    "oracle_enable_sdf": (C.c_int, [_P, C.c_int]),
    "oracle_set_strain_limit": (C.c_int, [_P, C.c_int, C.c_float, C.POINTER(C.c_float)]),
    "oracle_advance": (C.c_int, [_P, C.c_float]),
    "oracle_solve_next_step": (C.c_int, [_P, C.c_float]),
    "oracle_set_threads": (C.c_int, [_P, C.c_int]),
    "oracle_num_verts": (C.c_int, [_P]),
    "oracle_rotor_mul": (None, [_FP, _FP, _FP]),
    "oracle_wind_cos": (None, [_DP, _DP, C.c_long]),
    "oracle_contact_pair": (None, [C.c_void_p, _DP, C.c_double, _DP]),
    "oracle_mesh_sdf": (C.c_long, [_FP, C.c_int, _IP, C.c_int, _IP, _FP, _FP, _FP, _FP, _DP, C.c_long]),
    "oracle_mesh_distance": (C.c_int, [_FP, C.c_int, _IP, C.c_int, _DP, C.c_int, _DP, _DP, _IP]),
    "oracle_cubic_shape": (None, [_DP, _DP, _DP]),
    "oracle_cubic_nodes": (C.c_long, [_FP, C.c_int, _IP, _DP, C.c_long]),
    "oracle_cubic_interpolate": (C.c_int, [_FP, C.c_int, _IP, _DP, _DP, C.c_int, _DP, _DP]),
    "oracle_det_logf": (None, [_FP, _FP, C.c_long]),
    "oracle_detect_only": (None, [C.c_void_p, C.c_float]),
    "oracle_max_step": (None, [C.c_void_p, _FP]),
    "oracle_dx": (None, [C.c_void_p, _FP]),
    "oracle_wind_powf": (None, [_FP, _FP, _FP, C.c_long]),
    "oracle_rotor_rotate": (None, [_FP, _FP, _FP]),
    "oracle_rotor_from_to": (None, [_FP, _FP, _FP]),
    "oracle_segment_closest_points": (None, [_FP, _FP, _FP, _FP, _FP]),
    "oracle_morton3": (C.c_uint32, [C.c_int, C.c_int, C.c_int]),
    "oracle_sym_solve": (None, [_FP, _FP, _FP]),
    "oracle_lbvh_build": (_P, [_FP, C.c_int]),
    "oracle_lbvh_refit": (None, [_P, _FP]),
    "oracle_lbvh_destroy": (None, [_P]),
    "oracle_lbvh_query": (C.c_long, [_P, _IP, C.c_long]),
    "oracle_lbvh_get": (None, [_P, _UP, _UP, _UP, _FP, _UP, _FP, _IP]),
    "oracle_num_pairs": (C.c_long, [_P]),
    "oracle_error": (C.c_int, [_P]),
    "oracle_contacts": (None, [_P, _IP, _IP]),
    "oracle_set_material": (C.c_int, [_P, _FP, _FP]),
    "oracle_set_material64": (C.c_int, [_P, _DP, _DP]),
    "oracle_set_pos_vel64": (C.c_int, [_P, _DP, _DP]),
    "oracle_gradient": (C.c_double, [_P, C.c_float, C.c_int, C.c_int, _IP, C.c_int, _DP, _DP, _DP, C.c_int, _DP, _DP]),
    "oracle_clamp_log": (C.c_long, [_P, C.c_int, C.POINTER(C.c_uint8), C.c_long]),
}

_lib = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _f(a):
    return a.ctypes.data_as(_FP)


def _i(a):
    return a.ctypes.data_as(_IP)


class OracleSim:
    """Same call sequence as kamisimulator_amd._lib.Sim, on the CPU restatement."""

    def __init__(self, params, positions, strand_offsets, precision=32):
        self.lib = load()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        off = np.ascontiguousarray(strand_offsets, dtype=np.int32)
        self.n = pos.shape[0]
        self.precision = precision
        h = C.c_void_p()
        self._params = params  # keep alive
        rc = self.lib.oracle_create(precision, C.cast(C.byref(params), C.c_void_p), self.n, _f(pos), off.size - 1,
                                    _i(off), C.byref(h))
        if rc:
            raise RuntimeError(f"oracle_create failed with status {rc}")
        self.h = h

    def _check(self, rc):
        if rc:
            raise RuntimeError(f"oracle call failed with status {rc}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, density):
        self._check(self.lib.oracle_configure(self.h, density))

    def set_young(self, y, dev=1.0):
        self._check(self.lib.oracle_set_young(self.h, y, dev))

    def set_kstretch(self, k):
        self._check(self.lib.oracle_set_kstretch(self.h, k))

    def set_kbend(self, k):
        self._check(self.lib.oracle_set_kbend(self.h, k))

    def pin(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._check(self.lib.oracle_pin_vertices(self.h, _i(ids), ids.size))

    def fix_orientation(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._check(self.lib.oracle_fix_orientation(self.h, _i(ids), ids.size))

    def sag_free(self, roots, iterations=1):
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        self._check(self.lib.oracle_sag_free(self.h, _i(roots), roots.size, int(iterations)))

    def set_params(self, params):
        self._params = params
        self._check(self.lib.oracle_set_params(self.h, C.cast(C.byref(params), C.c_void_p)))

    def upload(self):
        self._check(self.lib.oracle_upload(self.h))

    def set_threads(self, t):
        self.lib.oracle_set_threads(self.h, int(t))

    def set_pos_vel(self, pos=None, vel=None):
        p = None if pos is None else np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
        self.lib.oracle_set_pos_vel(self.h, None if p is None else _f(p), None if v is None else _f(v))

    def set_pos_vel64(self, pos=None, vel=None):
        """Positions and velocities without rounding to float (for the FP64 state)."""
        p = None if pos is None else np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        self.lib.oracle_set_pos_vel64(self.h, None if p is None else p.ctypes.data_as(_DP),
                                      None if v is None else v.ctypes.data_as(_DP))

    def set_connections(self, conn=None):
        c = None if conn is None else np.ascontiguousarray(conn, dtype=np.int32)
        self.lib.oracle_set_connections(self.h, None if c is None else _i(c))

    def set_last_vel(self, u):
        u = np.ascontiguousarray(u, dtype=np.float32).reshape(-1)
        self.lib.oracle_set_last_vel(self.h, _f(u))

    def get_pos_vel(self, f64=False):
        if f64:
            p = np.empty((self.n, 3), np.float64)
            v = np.empty((self.n, 3), np.float64)
            self.lib.oracle_get_pos_vel64(self.h, p.ctypes.data_as(_DP), v.ctypes.data_as(_DP))
        else:
            p = np.empty((self.n, 3), np.float32)
            v = np.empty((self.n, 3), np.float32)
            self.lib.oracle_get_pos_vel(self.h, _f(p), _f(v))
        return p, v

    def get_q(self, f64=False):
        if f64:
            q = np.empty((self.n, 4), np.float64)
            self.lib.oracle_get_q64(self.h, q.ctypes.data_as(_DP))
        else:
            q = np.empty((self.n, 4), np.float32)
            self.lib.oracle_get_q(self.h, _f(q))
        return q

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(-1)
        self.lib.oracle_set_q(self.h, _f(q))

    def get_rest(self):
        im = np.empty(self.n, np.float32)
        lr = np.empty(self.n, np.float32)
        ks = np.empty(self.n, np.float32)
        fl = np.empty(self.n, np.uint32)
        qr = np.empty((self.n, 4), np.float32)
        self.lib.oracle_get_rest(self.h, _f(im), _f(lr), _f(ks), fl.ctypes.data_as(_UP), _f(qr))
        return dict(inv_mass=im, l_rest=lr, k_stretch=ks, flags=fl, q_rest=qr)

    def set_sdf(self, grid, threshold=0.0, rotation=(0, 0, 0, 1), translation=(0, 0, 0)):
        dim = np.ascontiguousarray(grid.dim, dtype=np.int32)
        gmin = np.ascontiguousarray(grid.gmin, dtype=np.float32)
        gmax = np.ascontiguousarray(grid.gmax, dtype=np.float32)
        phi = np.ascontiguousarray(grid.phi, dtype=np.float32).reshape(-1)
        grad = np.ascontiguousarray(grid.grad, dtype=np.float32).reshape(-1)
        rot = np.ascontiguousarray(rotation, dtype=np.float32)
        tr = np.ascontiguousarray(translation, dtype=np.float32)
        self._check(self.lib.oracle_set_sdf(self.h, _i(dim), _f(gmin), _f(gmax), _f(phi), _f(grad), threshold,
                                            _f(rot), _f(tr)))

    def enable_sdf(self, on=True):
        self.lib.oracle_enable_sdf(self.h, int(bool(on)))

    def set_strain_limit(self, enable=True, strain_error=0.1, rest_lengths=None):
        rl = None if rest_lengths is None else np.ascontiguousarray(rest_lengths, np.float32).reshape(-1)
        if rl is not None and rl.size != self.n:
            raise ValueError("rest_lengths must hold one value per vertex")
        self.lib.oracle_set_strain_limit(self.h, int(bool(enable)), float(strain_error),
                                         None if rl is None else _f(rl))

    def advance(self, dt):
        self.lib.oracle_advance(self.h, dt)

    def solve_next_step(self, dt):
        self.lib.oracle_solve_next_step(self.h, dt)


    def set_material(self, k_stretch=None, q_rest=None):
        ks = None if k_stretch is None else np.ascontiguousarray(k_stretch, np.float32)
        qr = None if q_rest is None else np.ascontiguousarray(q_rest, np.float32).reshape(-1)
        self.lib.oracle_set_material(self.h, None if ks is None else _f(ks), None if qr is None else _f(qr))

    def set_material64(self, k_stretch, q_rest):
        """Per-vertex materials in the state's precision (call after upload)."""
        ks = np.ascontiguousarray(k_stretch, np.float64)
        qr = np.ascontiguousarray(q_rest, np.float64).reshape(-1)
        self.lib.oracle_set_material64(self.h, ks.ctypes.data_as(_DP), qr.ctypes.data_as(_DP))

    def gradient(self, dt, nsteps, ckpt_every, tips, targets, with_sdf=False, initial=False):
        """Loss and (dL/dkStretch[n], dL/dqRest[n,4]) in float64; with_sdf:
        each step followed by the SDF pass; initial: also dL/dx and dL/dv of
        the initial state ([n,3] each)."""
        tips = np.ascontiguousarray(tips, np.int32)
        tg = np.ascontiguousarray(targets, np.float64).reshape(-1)
        ksb = np.empty(self.n, np.float64)
        qrb = np.empty((self.n, 4), np.float64)
        x0b = np.empty((self.n, 3), np.float64)
        v0b = np.empty((self.n, 3), np.float64)
        loss = self.lib.oracle_gradient(self.h, dt, nsteps, ckpt_every, _i(tips), tips.size, tg.ctypes.data_as(_DP),
                                        ksb.ctypes.data_as(_DP), qrb.ctypes.data_as(_DP), 1 if with_sdf else 0,
                                        x0b.ctypes.data_as(_DP), v0b.ctypes.data_as(_DP))
        if initial:
            return loss, ksb, qrb, x0b, v0b
        return loss, ksb, qrb

    def log_clamp(self, enable=True):
        """Starts (enable) or stops logging the branch of the adaptive initial
        guess's clamp per vertex and substep; returns the log so far as
        uint8[substeps, n] (0: sigma clamped to 0, 1: interior, 2: clamped
        to 1, 3: no guess)."""
        m = self.lib.oracle_clamp_log(self.h, -1, None, 0)
        out = np.empty(m, np.uint8)
        self.lib.oracle_clamp_log(self.h, int(bool(enable)), out.ctypes.data_as(C.POINTER(C.c_uint8)), m)
        return out.reshape(-1, self.n)

    def num_pairs(self):
        return self.lib.oracle_num_pairs(self.h)

    def error(self):
        return self.lib.oracle_error(self.h)

    def contacts(self):
        num = np.empty(self.n, np.int32)
        cols = np.empty((self.n, 16), np.int32)
        self.lib.oracle_contacts(self.h, _i(num), _i(cols))
        return num, cols

    def detect_only(self, dt):
        """initItr + a forced detection + step limits of one substep (no iterations)."""
        self.lib.oracle_detect_only(self.h, dt)

    def max_step(self):
        out = np.empty(self.n, np.float32)
        self.lib.oracle_max_step(self.h, _f(out))
        return out

    def dx(self):
        out = np.empty((self.n, 3), np.float32)
        self.lib.oracle_dx(self.h, _f(out))
        return out


class OracleLbvh:
    """CPU restatement of Kitten::LBVH (lbvh_oracle.hpp)."""

    def __init__(self, boxes):
        self.lib = load()
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        self.n = b.shape[0]
        self.boxes = b
        self.h = self.lib.oracle_lbvh_build(_f(b), self.n)

    def refit(self, boxes):
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        self.boxes = b
        self.lib.oracle_lbvh_refit(self.h, _f(b))

    def query(self, cap=None):
        cap = cap if cap is not None else 64 * self.n
        out = np.empty((cap, 2), np.int32)
        m = self.lib.oracle_lbvh_query(self.h, _i(out), cap)
        return out[:min(m, cap)].copy()

    def get(self):
        n = self.n
        codes = np.empty(n, np.uint32)
        ids = np.empty(n, np.uint32)
        links = np.empty((max(n - 1, 1), 4), np.uint32)
        bounds = np.empty((max(n - 1, 1), 12), np.float32)
        lp = np.empty(n, np.uint32)
        root = np.empty(6, np.float32)
        ms = C.c_int(0)
        u = lambda a: a.ctypes.data_as(_UP)
        self.lib.oracle_lbvh_get(self.h, u(codes), u(ids), u(links), _f(bounds), u(lp), _f(root), C.byref(ms))
        return dict(codes=codes, ids=ids, links=links[:n - 1], bounds=bounds[:n - 1], leaf_parent=lp, root=root,
                    max_stack=ms.value)

    def __del__(self):
        try:
            self.lib.oracle_lbvh_destroy(self.h)
        except Exception:
            pass


# ------------------------------------------------------------------ primitives
def rotor_mul(a, b):
    lib = load()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.empty(4, np.float32)
    lib.oracle_rotor_mul(_f(a), _f(b), _f(out))
    return out


def rotor_rotate(q, v):
    lib = load()
    q = np.ascontiguousarray(q, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    out = np.empty(3, np.float32)
    lib.oracle_rotor_rotate(_f(q), _f(v), _f(out))
    return out


def rotor_from_to(a, b):
    lib = load()
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    out = np.empty(4, np.float32)
    lib.oracle_rotor_from_to(_f(a), _f(b), _f(out))
    return out


def segment_closest_points(a0, a1, b0, b1):
    lib = load()
    arrs = [np.ascontiguousarray(x, np.float32) for x in (a0, a1, b0, b1)]
    out = np.empty(2, np.float32)
    lib.oracle_segment_closest_points(*[_f(x) for x in arrs], _f(out))
    return out


def morton3(x, y, z):
    return load().oracle_morton3(int(x), int(y), int(z))


def mesh_sdf(verts, tris, dim, with_coef=False):
    """The body SDF grid of a triangle mesh (meshsdf_oracle.hpp): phi[dz, dy, dx],
    grad[dz, dy, dx, 3], the grid box (gmin, gmax) and, with with_coef, the
    cubic grid's node values."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.int32)
    d = np.ascontiguousarray(dim, np.int32)
    n = int(np.prod(d))
    phi = np.empty(n, np.float32)
    grad = np.empty(3 * n, np.float32)
    gmin = np.empty(3, np.float32)
    gmax = np.empty(3, np.float32)
    nn = lib.oracle_cubic_nodes(_f(v), len(v), _i(d), None, 0)
    coef = np.empty(nn if with_coef else 0, np.float64)
    lib.oracle_mesh_sdf(_f(v), len(v), _i(t), len(t), _i(d), _f(phi), _f(grad), _f(gmin), _f(gmax),
                        coef.ctypes.data_as(_DP) if with_coef else None, nn if with_coef else 0)
    out = (phi.reshape(d[2], d[1], d[0]), grad.reshape(d[2], d[1], d[0], 3), gmin, gmax)
    return out + (coef,) if with_coef else out


def mesh_distance(verts, tris, points):
    """Signed distance (double), nearest point and nearest triangle of each point."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32)
    t = np.ascontiguousarray(tris, np.int32)
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    dist = np.empty(len(p), np.float64)
    near = np.empty((len(p), 3), np.float64)
    tri = np.empty(len(p), np.int32)
    lib.oracle_mesh_distance(_f(v), len(v), _i(t), len(t), p.ctypes.data_as(_DP), len(p), dist.ctypes.data_as(_DP),
                             near.ctypes.data_as(_DP), _i(tri))
    return dist, near, tri


def cubic_shape(xi):
    """Shape functions N[32] and derivatives dN[32, 3] of the cubic grid cell."""
    lib = load()
    x = np.ascontiguousarray(xi, np.float64)
    N = np.empty(32, np.float64)
    dN = np.empty((32, 3), np.float64)
    lib.oracle_cubic_shape(x.ctypes.data_as(_DP), N.ctypes.data_as(_DP), dN.ctypes.data_as(_DP))
    return N, dN


def cubic_interpolate(verts, dim, coef, points):
    """The cubic grid with node values coef, interpolated: value[n], grad[n, 3]."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32)
    d = np.ascontiguousarray(dim, np.int32)
    c = np.ascontiguousarray(coef, np.float64)
    assert len(c) == lib.oracle_cubic_nodes(_f(v), len(v), _i(d), None, 0)
    p = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    val = np.empty(len(p), np.float64)
    grad = np.empty((len(p), 3), np.float64)
    lib.oracle_cubic_interpolate(_f(v), len(v), _i(d), c.ctypes.data_as(_DP), p.ctypes.data_as(_DP), len(p),
                                 val.ctypes.data_as(_DP), grad.ctypes.data_as(_DP))
    return val, grad


def cubic_nodes(verts, dim):
    """Node positions [n, 3] of the cubic grid over the padded vertex box."""
    lib = load()
    v = np.ascontiguousarray(verts, np.float32)
    d = np.ascontiguousarray(dim, np.int32)
    n = lib.oracle_cubic_nodes(_f(v), len(v), _i(d), None, 0)
    pos = np.empty((n, 3), np.float64)
    lib.oracle_cubic_nodes(_f(v), len(v), _i(d), pos.ctypes.data_as(_DP), n)
    return pos


def contact_pair(params, xi, xj, dxi, dxj, b0, b1, db0, db1, damp=0.0):
    """orc::contact_pair in double: (f, H6, f2, H2_6) from zero accumulators."""
    lib = load()
    inp = np.ascontiguousarray(np.concatenate([np.asarray(a, np.float64).reshape(3)
                                               for a in (xi, xj, dxi, dxj, b0, b1, db0, db1)]))
    out = np.zeros(18, np.float64)
    lib.oracle_contact_pair(C.byref(params), inp.ctypes.data_as(_DP), float(damp), out.ctypes.data_as(_DP))
    return out[0:3], out[3:9], out[9:12], out[12:18]


def det_logf(d):
    lib = load()
    d = np.ascontiguousarray(d, np.float32)
    out = np.empty_like(d)
    lib.oracle_det_logf(_f(d), _f(out), d.size)
    return out


def wind_cos(x):
    """cos of doubles by the restatement's fixed operation sequence."""
    lib = load()
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty_like(x)
    lib.oracle_wind_cos(x.ctypes.data_as(_DP), out.ctypes.data_as(_DP), x.size)
    return out


def wind_powf(b, e):
    """powf of float pairs by the restatement's fixed operation sequence."""
    lib = load()
    b, e = np.broadcast_arrays(np.asarray(b, np.float32), np.asarray(e, np.float32))
    b = np.ascontiguousarray(b)
    e = np.ascontiguousarray(e)
    out = np.empty_like(b)
    lib.oracle_wind_powf(_f(b), _f(e), _f(out), b.size)
    return out


def sym_solve(H6, f):
    lib = load()
    H6 = np.ascontiguousarray(H6, np.float32)
    f = np.ascontiguousarray(f, np.float32)
    out = np.empty(3, np.float32)
    lib.oracle_sym_solve(_f(H6), _f(f), _f(out))
    return out
