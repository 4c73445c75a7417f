"""ctypes binding of libkscr.so (the C ABI of include/kscr.h).

The library is the product path.  If it is missing this module raises: there
is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libkscr.so")

SCHEDULE_JACOBI = 0
SCHEDULE_COLORED = 1

BATCH_GRID = 0
BATCH_PERSISTENT = 1

FLAG_HAS_PREV = 1
FLAG_HAS_NEXT = 2
FLAG_HAS_NEXT_ORIENTATION = 4
FLAG_FIX_ORIENTATION = 8

STATUS = {
    0: "KSCR_OK",
    1: "KSCR_ERR_INVALID_ARGUMENT",
    2: "KSCR_ERR_TOO_FEW_VERTICES",
    3: "KSCR_ERR_DANGLING_SEGMENT",
    4: "KSCR_ERR_ZERO_LENGTH_SEGMENT",
    5: "KSCR_ERR_DETECTION_SCALER",
    6: "KSCR_ERR_SEGMENT_TOO_SHORT",
    7: "KSCR_ERR_NOT_CONFIGURED",
    8: "KSCR_ERR_MAX_COLLISIONS",
    9: "KSCR_ERR_STRAND_TOO_LONG",
    10: "KSCR_ERR_HIP",
    11: "KSCR_ERR_OUT_OF_MEMORY",
    12: "KSCR_ERR_UNSUPPORTED",
}


class KscrParams(C.Structure):
    """Mirror of ``kscr_params`` (include/kscr.h)."""

    _fields_ = [
        ("gravity", C.c_float * 3),
        ("num_itr", C.c_int32),
        ("max_h", C.c_float),
        ("damping", C.c_float),
        ("velocity_decay", C.c_float),
        ("radius", C.c_float),
        ("barrier_thickness", C.c_float),
        ("acceleration_ratio", C.c_float),
        ("k_collision", C.c_float),
        ("k_friction", C.c_float),
        ("friction_coeff", C.c_float),
        ("detection_scaler", C.c_float),
        ("bvh_rebuild_period", C.c_float),
        ("detection_period", C.c_int32),
        ("use_step_size_limit", C.c_int32),
        ("use_velocity_radius", C.c_int32),
        ("enable_hair_hair_coll", C.c_int32),
        ("external_force_multiplier", C.c_float),
        ("wind_peak", C.c_float),
        ("wind_y_freq", C.c_float),
        ("wind_z_freq", C.c_float),
        ("wind_time_freq", C.c_float),
        ("wind_sharpness", C.c_float),
        ("schedule", C.c_int32),
        ("block_size", C.c_int32),
    ]


# Every entry point declared in include/kscr.h, with its ctypes signature.
_P = C.c_void_p
_FP = C.POINTER(C.c_float)
_IP = C.POINTER(C.c_int)
_UP = C.POINTER(C.c_uint32)
SIGNATURES = {
    "kscr_params_default": (C.c_int, [C.POINTER(KscrParams)]),
    "kscr_create": (C.c_int, [C.POINTER(KscrParams), C.c_int, _FP, C.c_int, _IP, C.c_int, C.POINTER(_P)]),
    "kscr_destroy": (None, [_P]),
    "kscr_configure": (C.c_int, [_P, C.c_float]),
    "kscr_set_young": (C.c_int, [_P, C.c_float, C.c_float]),
    "kscr_set_kstretch": (C.c_int, [_P, C.c_float]),
    "kscr_set_kbend": (C.c_int, [_P, C.c_float]),
    "kscr_pin_vertices": (C.c_int, [_P, _IP, C.c_int]),
    "kscr_fix_orientation": (C.c_int, [_P, _IP, C.c_int]),
    "kscr_sag_free": (C.c_int, [_P, _IP, C.c_int, C.c_int]),
    "kscr_set_params": (C.c_int, [_P, C.POINTER(KscrParams)]),
    "kscr_get_params": (C.c_int, [_P, C.POINTER(KscrParams)]),
    "kscr_upload": (C.c_int, [_P]),
    "kscr_set_pos_vel": (C.c_int, [_P, _FP, _FP]),
    "kscr_get_pos_vel": (C.c_int, [_P, _FP, _FP]),
    "kscr_download_pos_vel": (C.c_int, [_P, C.POINTER(_FP), C.POINTER(_FP)]),
    "kscr_get_last_vel": (C.c_int, [_P, _FP]),
    "kscr_set_connections": (C.c_int, [_P, _IP]),
    "kscr_get_q": (C.c_int, [_P, _FP]),
    "kscr_set_q": (C.c_int, [_P, _FP]),
    "kscr_get_rest": (C.c_int, [_P, _FP, _FP, _FP, _UP, _FP]),
    "kscr_set_sdf": (C.c_int, [_P, _IP, _FP, _FP, _FP, _FP, C.c_float, _FP, _FP]),
    "kscr_enable_sdf": (C.c_int, [_P, C.c_int]),
    "kscr_set_strain_limit": (C.c_int, [_P, C.c_int, C.c_float, _FP]),
    "kscr_selftest_rounding": (C.c_int, [C.c_int, _FP, C.c_int, C.POINTER(C.c_ulonglong)]),
    "kscr_mesh_sdf": (C.c_int, [C.c_int, _FP, C.c_int, _IP, C.c_int, _IP, _FP, _FP, _FP, _FP, C.c_char_p,
                                C.c_size_t]),
    "kscr_set_sdf_from_mesh": (C.c_int, [_P, _FP, C.c_int, _IP, C.c_int, _IP, C.c_float, _FP, _FP]),
    "kscr_advance": (C.c_int, [_P, C.c_float]),
    "kscr_solve_next_step": (C.c_int, [_P, C.c_float]),
    "kscr_set_stream": (C.c_int, [_P, _P]),
    "kscr_synchronize": (C.c_int, [_P]),
    "kscr_num_verts": (C.c_int, [_P]),
    "kscr_num_strands": (C.c_int, [_P]),
    "kscr_block_size": (C.c_int, [_P]),
    "kscr_last_error": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "kscr_create_error": (C.c_char_p, []),
    "kscr_autotune": (C.c_int, [_P, C.c_float, C.c_int, _IP]),
    "kscr_batch_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.POINTER(_P)]),
    "kscr_batch_destroy": (None, [_P]),
    "kscr_batch_step": (C.c_int, [_P, C.c_float, C.c_int, C.c_int]),
    "kscr_batch_set_graph": (C.c_int, [_P, C.c_int]),
    "kscr_batch_synchronize": (C.c_int, [_P]),
    "kscr_batch_autotune": (C.c_int, [_P, C.c_float, C.c_int, _IP]),
    "kscr_batch_block_size": (C.c_int, [_P]),
    "kscr_batch_set_mapping": (C.c_int, [_P, C.c_int]),
    "kscr_batch_mapping": (C.c_int, [_P]),
    "kscr_batch_last_error": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "kscr_get_contacts": (C.c_int, [_P, _IP, _IP]),
    "kscr_set_material": (C.c_int, [_P, _FP, _FP]),
    "kscr_adjoint_forward": (C.c_int, [_P, C.c_float, C.c_int, C.c_int]),
    "kscr_adjoint_forward_sdf": (C.c_int, [_P, C.c_float, C.c_int, C.c_int]),
    "kscr_adjoint_backward": (C.c_int, [_P, _FP, _FP, _FP, _FP]),
    "kscr_adjoint_last_time": (C.c_int, [_P, _FP, _FP, _FP]),
    "kscr_lbvh_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "kscr_lbvh_destroy": (None, [_P]),
    "kscr_lbvh_compute": (C.c_int, [_P, _FP, C.c_int]),
    "kscr_lbvh_refit": (C.c_int, [_P, _FP]),
    "kscr_lbvh_query": (C.c_int, [_P, _IP, C.c_size_t, C.POINTER(C.c_size_t)]),
    "kscr_lbvh_compute_device": (C.c_int, [_P, _P, C.c_int, _P]),
    "kscr_lbvh_refit_device": (C.c_int, [_P, _P, _P]),
    "kscr_lbvh_query_device": (C.c_int, [_P, _P, C.c_size_t, C.POINTER(C.c_size_t), _P]),
    "kscr_lbvh_self_query_static": (C.c_int, [_P, C.c_int, _P, C.c_size_t, C.POINTER(C.c_size_t), _P, C.c_char_p,
                                              C.c_size_t]),
    "kscr_lbvh_get": (C.c_int, [_P, _UP, _UP, _UP, _FP, _UP, _FP, _IP]),
    "kscr_lbvh_set_fit_slots": (C.c_int, [_P, C.c_int]),
}

_lib = None


def load(path=LIB_PATH):
    """Loads libkscr.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build it with `python -m kamisimulator_amd.build` "
            "(the Stable Cosserat Rods step has no CPU fallback)")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class KscrError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


def fptr(a):
    return a.ctypes.data_as(_FP)


def iptr(a):
    return a.ctypes.data_as(_IP)


def default_params(**kw):
    lib = load()
    p = KscrParams()
    lib.kscr_params_default(C.byref(p))
    for k, v in kw.items():
        if k == "gravity":
            p.gravity = (C.c_float * 3)(*v)
        else:
            setattr(p, k, v)
    return p


def params_to_dict(p):
    d = {}
    for name, _ in KscrParams._fields_:
        val = getattr(p, name)
        d[name] = list(val) if name == "gravity" else val
    return d


def selftest_rounding(divisors=(1e-3,), device=-1):
    """Differing-result counts of the kernels' fast correctly rounded
    arithmetic against IEEE on the GPU (kscr_selftest_rounding): sqrt,
    1/sqrt, the reverse sweep's clamped sqrt and reciprocal, division by each
    divisor."""
    lib = load()
    d = np.ascontiguousarray(divisors, dtype=np.float32)
    counts = (C.c_ulonglong * 4)()
    rc = lib.kscr_selftest_rounding(int(device), fptr(d), len(d), counts)
    if rc:
        raise KscrError(rc, "kscr_selftest_rounding failed")
    return [int(c) for c in counts]


def mesh_sdf(vertices, triangles, dim, device=-1):
    """Body SDF grid of a closed triangle mesh, built on the GPU
    (kscr_mesh_sdf): a scenes.SdfGrid over the vertex box."""
    from .scenes import SdfGrid
    lib = load()
    v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
    d = np.ascontiguousarray(dim, dtype=np.int32)
    n = int(np.prod(d.astype(np.int64)))
    phi = np.empty(n, np.float32)
    grad = np.empty(3 * n, np.float32)
    gmin = np.empty(3, np.float32)
    gmax = np.empty(3, np.float32)
    err = C.create_string_buffer(512)
    rc = lib.kscr_mesh_sdf(int(device), fptr(v), len(v), iptr(t), len(t), iptr(d), fptr(phi), fptr(grad), fptr(gmin),
                           fptr(gmax), err, 512)
    if rc:
        raise KscrError(rc, err.value.decode())
    return SdfGrid(d, gmin, gmax, phi.reshape(d[2], d[1], d[0]), grad.reshape(d[2], d[1], d[0], 3))


class Sim:
    """Python view of one ``kscr_handle`` (mirrors YarnBall::Sim's lifecycle)."""

    def __init__(self, params, positions, strand_offsets, device=-1):
        self.lib = load()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        off = np.ascontiguousarray(strand_offsets, dtype=np.int32)
        self.n = pos.shape[0]
        self.num_strands = off.shape[0] - 1
        self.strand_offsets = off
        h = C.c_void_p()
        rc = self.lib.kscr_create(C.byref(params), self.n, fptr(pos), self.num_strands, iptr(off), device, C.byref(h))
        if rc:
            raise KscrError(rc, self.lib.kscr_create_error().decode())
        self.h = h

    def _check(self, rc):
        if rc:
            buf = C.create_string_buffer(512)
            self.lib.kscr_last_error(self.h, buf, 512)
            raise KscrError(rc, buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.kscr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def configure(self, density):
        self._check(self.lib.kscr_configure(self.h, density))

    def set_young(self, young_gpa, deviation=1.0):
        self._check(self.lib.kscr_set_young(self.h, young_gpa, deviation))

    def set_kstretch(self, k):
        self._check(self.lib.kscr_set_kstretch(self.h, k))

    def set_kbend(self, k):
        self._check(self.lib.kscr_set_kbend(self.h, k))

    def pin(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._check(self.lib.kscr_pin_vertices(self.h, iptr(ids), ids.size))

    def fix_orientation(self, ids):
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        self._check(self.lib.kscr_fix_orientation(self.h, iptr(ids), ids.size))

    def sag_free(self, roots=None, iterations=1):
        """Sim::SagFree: after configure/set_young/pin, before upload."""
        if roots is None:
            self._check(self.lib.kscr_sag_free(self.h, None, 0, int(iterations)))
            return
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        self._check(self.lib.kscr_sag_free(self.h, iptr(roots), roots.size, int(iterations)))

    def set_params(self, params):
        self._check(self.lib.kscr_set_params(self.h, C.byref(params)))

    def get_params(self):
        p = KscrParams()
        self._check(self.lib.kscr_get_params(self.h, C.byref(p)))
        return p

    def upload(self):
        self._check(self.lib.kscr_upload(self.h))

    def set_pos_vel(self, pos=None, vel=None):
        p = None if pos is None else np.ascontiguousarray(pos, dtype=np.float32).reshape(-1)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(-1)
        self._check(self.lib.kscr_set_pos_vel(self.h, None if p is None else fptr(p), None if v is None else fptr(v)))

    def get_pos_vel(self):
        p = np.empty((self.n, 3), np.float32)
        v = np.empty((self.n, 3), np.float32)
        self._check(self.lib.kscr_get_pos_vel(self.h, fptr(p), fptr(v)))
        return p, v

    def download_pos_vel(self):
        """Positions and velocities as read-only views of the handle's
        page-locked host mirror (valid until the next call on the handle)."""
        p, v = _FP(), _FP()
        self._check(self.lib.kscr_download_pos_vel(self.h, C.byref(p), C.byref(v)))
        pa = np.ctypeslib.as_array(p, shape=(self.n, 3))
        va = np.ctypeslib.as_array(v, shape=(self.n, 3))
        pa.flags.writeable = False
        va.flags.writeable = False
        return pa, va

    def set_connections(self, conn=None):
        c = None if conn is None else np.ascontiguousarray(conn, dtype=np.int32)
        self._check(self.lib.kscr_set_connections(self.h, None if c is None else iptr(c)))

    def get_last_vel(self):
        u = np.empty((self.n, 3), np.float32)
        self._check(self.lib.kscr_get_last_vel(self.h, fptr(u)))
        return u

    def get_q(self):
        q = np.empty((self.n, 4), np.float32)
        self._check(self.lib.kscr_get_q(self.h, fptr(q)))
        return q

    def set_q(self, q):
        q = np.ascontiguousarray(q, dtype=np.float32).reshape(-1)
        self._check(self.lib.kscr_set_q(self.h, fptr(q)))

    def get_rest(self):
        im = np.empty(self.n, np.float32)
        lr = np.empty(self.n, np.float32)
        ks = np.empty(self.n, np.float32)
        fl = np.empty(self.n, np.uint32)
        qr = np.empty((self.n, 4), np.float32)
        self._check(self.lib.kscr_get_rest(self.h, fptr(im), fptr(lr), fptr(ks), fl.ctypes.data_as(_UP), fptr(qr)))
        return dict(inv_mass=im, l_rest=lr, k_stretch=ks, flags=fl, q_rest=qr)

    def set_sdf(self, grid, threshold=0.0, rotation=(0, 0, 0, 1), translation=(0, 0, 0)):
        dim = np.ascontiguousarray(grid.dim, dtype=np.int32)
        gmin = np.ascontiguousarray(grid.gmin, dtype=np.float32)
        gmax = np.ascontiguousarray(grid.gmax, dtype=np.float32)
        phi = np.ascontiguousarray(grid.phi, dtype=np.float32).reshape(-1)
        grad = np.ascontiguousarray(grid.grad, dtype=np.float32).reshape(-1)
        rot = np.ascontiguousarray(rotation, dtype=np.float32)
        tr = np.ascontiguousarray(translation, dtype=np.float32)
        self._check(self.lib.kscr_set_sdf(self.h, iptr(dim), fptr(gmin), fptr(gmax), fptr(phi), fptr(grad),
                                          threshold, fptr(rot), fptr(tr)))

    def set_sdf_from_mesh(self, vertices, triangles, dim, threshold=0.0, rotation=(0, 0, 0, 1),
                          translation=(0, 0, 0)):
        """The body SDF grid built on the device from a closed triangle mesh
        (kscr_set_sdf_from_mesh)."""
        v = np.ascontiguousarray(vertices, dtype=np.float32).reshape(-1, 3)
        t = np.ascontiguousarray(triangles, dtype=np.int32).reshape(-1, 3)
        d = np.ascontiguousarray(dim, dtype=np.int32)
        rot = np.ascontiguousarray(rotation, dtype=np.float32)
        tr = np.ascontiguousarray(translation, dtype=np.float32)
        self._check(self.lib.kscr_set_sdf_from_mesh(self.h, fptr(v), len(v), iptr(t), len(t), iptr(d), threshold,
                                                    fptr(rot), fptr(tr)))

    def enable_sdf(self, on=True):
        self._check(self.lib.kscr_enable_sdf(self.h, int(bool(on))))

    def set_strain_limit(self, enable=True, strain_error=0.1, rest_lengths=None):
        """Strain limiting after the SDF pass (kscr_set_strain_limit);
        rest_lengths (n,) or None for the current segment lengths."""
        rl = None
        if rest_lengths is not None:
            rl = np.ascontiguousarray(rest_lengths, dtype=np.float32).reshape(-1)
            if rl.size != self.n:
                raise ValueError("rest_lengths must hold one value per vertex")
        self._check(self.lib.kscr_set_strain_limit(self.h, int(bool(enable)), float(strain_error),
                                                   None if rl is None else fptr(rl)))

    def advance(self, dt):
        self._check(self.lib.kscr_advance(self.h, dt))

    def solve_next_step(self, dt):
        self._check(self.lib.kscr_solve_next_step(self.h, dt))

    def set_stream(self, stream_ptr):
        self._check(self.lib.kscr_set_stream(self.h, C.c_void_p(stream_ptr)))

    def synchronize(self):
        self._check(self.lib.kscr_synchronize(self.h))

    @property
    def block_size(self):
        return self.lib.kscr_block_size(self.h)

    def get_contacts(self):
        num = np.empty(self.n, np.int32)
        cols = np.empty((self.n, 16), np.int32)
        self._check(self.lib.kscr_get_contacts(self.h, iptr(num), iptr(cols)))
        return num, cols

    def set_material(self, k_stretch=None, q_rest=None):
        """Per-vertex kStretch (n) and qRest (n, 4)."""
        ks = None if k_stretch is None else np.ascontiguousarray(k_stretch, dtype=np.float32).reshape(-1)
        qr = None if q_rest is None else np.ascontiguousarray(q_rest, dtype=np.float32).reshape(-1)
        if ks is not None and ks.size != self.n:
            raise ValueError("k_stretch must hold one value per vertex")
        if qr is not None and qr.size != 4 * self.n:
            raise ValueError("q_rest must hold one rotor per vertex")
        self._check(self.lib.kscr_set_material(self.h, None if ks is None else fptr(ks),
                                               None if qr is None else fptr(qr)))

    def adjoint_forward(self, dt, nsteps, checkpoint_every, with_sdf=False):
        """nsteps steps of dt keeping a checkpoint every checkpoint_every steps
        (with_sdf: each followed by the SDF pass, as solve_next_step)."""
        fn = self.lib.kscr_adjoint_forward_sdf if with_sdf else self.lib.kscr_adjoint_forward
        self._check(fn(self.h, dt, nsteps, checkpoint_every))

    def adjoint_backward(self, x_bar=None, v_bar=None):
        """Adjoints of the final (x, v) -> (d/dkStretch (n,), d/dqRest (n, 4))."""
        xb = None if x_bar is None else np.ascontiguousarray(x_bar, dtype=np.float32).reshape(-1)
        vb = None if v_bar is None else np.ascontiguousarray(v_bar, dtype=np.float32).reshape(-1)
        for a in (xb, vb):
            if a is not None and a.size != 3 * self.n:
                raise ValueError("adjoint seeds must hold 3 floats per vertex")
        ksb = np.empty(self.n, np.float32)
        qrb = np.empty((self.n, 4), np.float32)
        self._check(self.lib.kscr_adjoint_backward(self.h, None if xb is None else fptr(xb),
                                                   None if vb is None else fptr(vb), fptr(ksb), fptr(qrb)))
        return ksb, qrb

    def adjoint_time(self):
        """(re-record ms, reverse sweep ms, whole backward ms) of the last
        adjoint_backward; re-recording overlaps the sweeps."""
        r = C.c_float(0)
        w = C.c_float(0)
        b = C.c_float(0)
        self._check(self.lib.kscr_adjoint_last_time(self.h, C.byref(r), C.byref(w), C.byref(b)))
        return r.value, w.value, b.value

    def autotune(self, dt, steps=3):
        best = C.c_int(0)
        self._check(self.lib.kscr_autotune(self.h, dt, steps, C.byref(best)))
        return best.value


class Batch:
    """Many handles (grooms) advanced by one launch per step (kscr_batch_*)."""

    def __init__(self, sims, block_size=0):
        self.lib = load()
        self.sims = list(sims)  # keep the handles alive
        arr = (C.c_void_p * len(self.sims))(*[s.h for s in self.sims])
        b = C.c_void_p()
        rc = self.lib.kscr_batch_create(arr, len(self.sims), block_size, C.byref(b))
        if rc:
            raise KscrError(rc, "kscr_batch_create failed")
        self.b = b

    def _check(self, rc):
        if rc:
            buf = C.create_string_buffer(512)
            self.lib.kscr_batch_last_error(self.b, buf, 512)
            raise KscrError(rc, buf.value.decode())

    def step(self, dt, nsteps=1, with_sdf=True):
        self._check(self.lib.kscr_batch_step(self.b, dt, nsteps, int(bool(with_sdf))))

    def autotune(self, dt, steps=3):
        best = C.c_int(0)
        self._check(self.lib.kscr_batch_autotune(self.b, dt, steps, C.byref(best)))
        return best.value

    @property
    def mapping(self):
        """BATCH_GRID or BATCH_PERSISTENT (kscr_batch_set_mapping)."""
        return self.lib.kscr_batch_mapping(self.b)

    def set_mapping(self, mapping):
        self._check(self.lib.kscr_batch_set_mapping(self.b, int(mapping)))

    def set_graph(self, on):
        self._check(self.lib.kscr_batch_set_graph(self.b, int(bool(on))))

    def synchronize(self):
        self._check(self.lib.kscr_batch_synchronize(self.b))

    @property
    def block_size(self):
        return self.lib.kscr_batch_block_size(self.b)

    def close(self):
        if getattr(self, "b", None):
            self.lib.kscr_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Lbvh:
    """Python view of a ``kscr_lbvh`` (the reference's Kitten::LBVH interface)."""

    def __init__(self, device=-1):
        self.lib = load()
        b = C.c_void_p()
        rc = self.lib.kscr_lbvh_create(device, C.byref(b))
        if rc:
            raise KscrError(rc, "kscr_lbvh_create failed")
        self.b = b
        self.n = 0

    def compute(self, boxes):
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        self.n = b.shape[0]
        rc = self.lib.kscr_lbvh_compute(self.b, fptr(b), self.n)
        if rc:
            raise KscrError(rc, "kscr_lbvh_compute failed")

    def refit(self, boxes):
        b = np.ascontiguousarray(boxes, np.float32).reshape(-1, 6)
        rc = self.lib.kscr_lbvh_refit(self.b, fptr(b))
        if rc:
            raise KscrError(rc, "kscr_lbvh_refit failed")

    def query(self, cap=None):
        cap = cap if cap is not None else 64 * self.n
        out = np.empty((cap, 2), np.int32)
        cnt = C.c_size_t(0)
        rc = self.lib.kscr_lbvh_query(self.b, iptr(out), cap, C.byref(cnt))
        if rc:
            raise KscrError(rc, "kscr_lbvh_query failed")
        return out[:min(cnt.value, cap)].copy(), cnt.value

    def compute_device(self, d_aabbs, n, stream=None):
        """Build from boxes already on the device (d_aabbs: device address of
        float[6n]); stream: hipStream_t address or None."""
        self.n = int(n)
        rc = self.lib.kscr_lbvh_compute_device(self.b, C.c_void_p(d_aabbs), self.n, C.c_void_p(stream))
        if rc:
            raise KscrError(rc, "kscr_lbvh_compute_device failed")

    def refit_device(self, d_aabbs, stream=None):
        rc = self.lib.kscr_lbvh_refit_device(self.b, C.c_void_p(d_aabbs), C.c_void_p(stream))
        if rc:
            raise KscrError(rc, "kscr_lbvh_refit_device failed")

    def query_device(self, d_pairs, cap, stream=None):
        """Self query into a device buffer int[2*cap]; returns the pair count."""
        cnt = C.c_size_t(0)
        rc = self.lib.kscr_lbvh_query_device(self.b, C.c_void_p(d_pairs), int(cap), C.byref(cnt), C.c_void_p(stream))
        if rc:
            raise KscrError(rc, "kscr_lbvh_query_device failed")
        return cnt.value

    def set_fit_slots(self, slots):
        """Crossing nodes the fit of a small tree keeps in LDS (0 .. 4096)."""
        rc = self.lib.kscr_lbvh_set_fit_slots(self.b, int(slots))
        if rc:
            raise KscrError(rc, "kscr_lbvh_set_fit_slots failed")

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
        rc = self.lib.kscr_lbvh_get(self.b, u(codes), u(ids), u(links), fptr(bounds), u(lp), fptr(root), C.byref(ms))
        if rc:
            raise KscrError(rc, "kscr_lbvh_get failed")
        return dict(codes=codes, ids=ids, links=links[:n - 1], bounds=bounds[:n - 1], leaf_parent=lp, root=root,
                    max_stack=ms.value)

    def close(self):
        if getattr(self, "b", None):
            self.lib.kscr_lbvh_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
