"""Synthetic grooms and analytic SDF grids for the benchmark workloads.

Everything is generated in float64 with seeded numpy Generators and cast to
float32 at the end.  Parameters follow the reference's Stable Cosserat Rods
adapter (src/solvers/StableCosseratRods/StableCosseratRodsSolver.cpp:177-207):
meta.radius = 0.05 * CurveRadius, barrierThickness = 2 * 0.95 * CurveRadius,
Young's modulus in GPa, roots pinned.

Workloads (name: shape):
    cantilever : 1 strand x 64 vertices, 1 mm segments along +x, root clamped
    groom      : roots on the upper hemisphere of a 0.1 m sphere, strands along
                 the outward normal (1 mm axial spacing) with a seeded helical
                 curl of radius U[2, 8] mm and phase U[0, 2pi), pitch 32 vertices
    lognormal  : groom with per-strand vertex counts round(96 exp(0.35 Z))
                 clipped to [32, 256]
"""
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


@dataclass
class Material:
    """Settings a Kami JSON gives the SCR solver, in the reference's units."""
    curve_radius: float = 4e-5       # CurveRadius [m]
    density: float = 1300.0          # Density [kg/m^3]
    young_gpa: float = 3.0           # YoungsModule [GPa]
    deviation: float = 1.0           # StiffnessDeviationCoeff
    num_itr: int = 10                # NumItr
    velocity_decay: float = 0.05     # VelocityDecay
    gravity: tuple = (0.0, -9.81, 0.0)
    wind_peak: float = 0.0
    # the wind shape of the SCR solver's JSON defaults (used when wind_peak != 0)
    wind_y_freq: float = 1.0
    wind_z_freq: float = 0.7
    wind_time_freq: float = 0.4
    wind_sharpness: float = 3.0
    external_force_multiplier: float = 1.0
    sag_free_itr: int = 0            # SagFreeItr


@dataclass
class SdfGrid:
    dim: np.ndarray      # int32[3]
    gmin: np.ndarray     # float32[3]
    gmax: np.ndarray     # float32[3]
    phi: np.ndarray      # float32[dz, dy, dx]  (x fastest in memory)
    grad: np.ndarray     # float32[dz, dy, dx, 3]


@dataclass
class Scene:
    name: str
    positions: np.ndarray          # float32[N, 3]
    strand_offsets: np.ndarray     # int32[S+1]
    material: Material = field(default_factory=Material)
    sdf: SdfGrid = None
    sdf_threshold: float = 0.0
    clamp_root_orientation: bool = False
    gravity_dirs: np.ndarray = None  # per-groom gravity (batched form)

    @property
    def num_verts(self):
        return int(self.positions.shape[0])

    @property
    def num_strands(self):
        return int(self.strand_offsets.shape[0] - 1)

    @property
    def roots(self):
        return self.strand_offsets[:-1].astype(np.int32)


def make_params(lib_default_params, material: Material, block_size=0, schedule=0):
    """kscr_params for a material, as StableCosseratRodsSolver::BuildYarnBallHair maps them."""
    r = F32(material.curve_radius)
    ratio = F32(0.05)
    return lib_default_params(
        radius=float(ratio * r),
        barrier_thickness=float(F32(2) * (F32(1) - ratio) * r),
        num_itr=int(material.num_itr),
        velocity_decay=float(material.velocity_decay),
        use_step_size_limit=0,
        use_velocity_radius=1,
        enable_hair_hair_coll=0,
        external_force_multiplier=float(material.external_force_multiplier),
        wind_peak=float(material.wind_peak),
        wind_y_freq=float(material.wind_y_freq),
        wind_z_freq=float(material.wind_z_freq),
        wind_time_freq=float(material.wind_time_freq),
        wind_sharpness=float(material.wind_sharpness),
        gravity=tuple(float(g) for g in material.gravity),
        block_size=int(block_size),
        schedule=int(schedule),
    )


def setup_sim(sim, scene: Scene):
    """The reference's setup sequence on any object with the Sim interface:
    configure(density) -> setYoung -> pin roots (-> clamp) -> SagFree ->
    upload -> SDF (StableCosseratRodsSolver.cpp:194-209)."""
    m = scene.material
    sim.configure(m.density)
    sim.set_young(m.young_gpa, m.deviation)
    sim.pin(scene.roots)
    if scene.clamp_root_orientation:
        sim.fix_orientation(scene.roots)
    if m.sag_free_itr:
        sim.sag_free(scene.roots, m.sag_free_itr)
    sim.upload()
    if scene.sdf is not None:
        sim.set_sdf(scene.sdf, threshold=scene.sdf_threshold)
    return sim


# ------------------------------------------------------------------ generators
def cantilever(num_verts=64, seg_len=1e-3, clamp=True):
    x = np.arange(num_verts, dtype=np.float64) * seg_len
    pos = np.stack([x, np.zeros_like(x), np.zeros_like(x)], 1).astype(F32)
    off = np.array([0, num_verts], np.int32)
    return Scene("cantilever", pos, off, clamp_root_orientation=clamp)


def _frames(n):
    a = np.where(np.abs(n[:, 2:3]) < 0.9, np.array([[0.0, 0.0, 1.0]]), np.array([[1.0, 0.0, 0.0]]))
    t1 = np.cross(n, a)
    t1 /= np.linalg.norm(t1, axis=1, keepdims=True)
    t2 = np.cross(n, t1)
    return t1, t2


def groom(num_strands, verts_per_strand, seed=1, sphere_radius=0.1, seg_len=1e-3, curl=(2e-3, 8e-3), pitch=32,
          center=(0.0, 0.0, 0.0), roots="random"):
    """Strands on the upper hemisphere of a sphere; verts_per_strand is an int or an int array.

    roots="random": heights and azimuths uniform (area-uniform, Poisson-like:
    some roots nearly coincide); roots="spiral": a golden-angle spiral with
    uniform height steps, so neighbouring roots are evenly spaced, as on a
    scalp."""
    rng = np.random.default_rng(seed)
    counts = np.broadcast_to(np.asarray(verts_per_strand, np.int64), (num_strands,)).copy()
    u = rng.random(num_strands)
    phi = rng.random(num_strands) * 2 * np.pi
    if roots == "spiral":
        k = np.arange(num_strands)
        u = (k + 0.5) / num_strands
        phi = np.mod(k * np.pi * (3.0 - np.sqrt(5.0)) + phi[0], 2 * np.pi)
    elif roots != "random":
        raise ValueError(f"unknown root layout {roots!r}")
    rho = curl[0] + (curl[1] - curl[0]) * rng.random(num_strands)
    th0 = rng.random(num_strands) * 2 * np.pi
    R = sphere_radius
    y = R * u
    rr = np.sqrt(np.maximum(R * R - y * y, 0.0))
    n = np.stack([rr * np.cos(phi), y, rr * np.sin(phi)], 1) / R
    root = n * R + np.asarray(center, np.float64)
    t1, t2 = _frames(n)
    off = np.zeros(num_strands + 1, np.int64)
    np.cumsum(counts, out=off[1:])
    N = int(off[-1])
    sid = np.repeat(np.arange(num_strands), counts)
    j = (np.arange(N) - off[sid]).astype(np.float64)
    w = 2 * np.pi / pitch
    ang = th0[sid] + w * j
    cj = (np.cos(ang) - np.cos(th0[sid]))[:, None]
    sj = (np.sin(ang) - np.sin(th0[sid]))[:, None]
    pos = root[sid] + n[sid] * (j * seg_len)[:, None] + rho[sid][:, None] * (cj * t1[sid] + sj * t2[sid])
    return Scene(f"groom{num_strands}x{verts_per_strand if np.isscalar(verts_per_strand) else 'var'}",
                 pos.astype(F32), off.astype(np.int32))


def lognormal_counts(num_strands, seed=2, median=96, sigma=0.35, lo=32, hi=256):
    rng = np.random.default_rng(seed)
    z = rng.standard_normal(num_strands)
    return np.clip(np.round(median * np.exp(sigma * z)), lo, hi).astype(np.int64)


# ------------------------------------------------------------------ SDF grids
def sdf_grid(fn, gmin, gmax, dim):
    """Samples phi and grad(phi) at nodes gmin + i*ds, ds = (gmax - gmin)/dim
    (the node convention of src/PNMesh.cpp:126-147: the last node is one cell
    short of gmax)."""
    gmin = np.asarray(gmin, F32)
    gmax = np.asarray(gmax, F32)
    dim = np.asarray(dim, np.int32)
    ds = ((gmax - gmin) / dim.astype(F32)).astype(np.float64)
    zs = gmin[2] + ds[2] * np.arange(dim[2])
    ys = gmin[1] + ds[1] * np.arange(dim[1])
    xs = gmin[0] + ds[0] * np.arange(dim[0])
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    P = np.stack([X, Y, Z], -1)
    phi, grad = fn(P)
    return SdfGrid(dim, gmin, gmax, phi.astype(F32), grad.astype(F32))


def sphere_sdf(center=(0.0, 0.0, 0.0), radius=0.1):
    c = np.asarray(center, np.float64)

    def fn(P):
        d = P - c
        r = np.linalg.norm(d, axis=-1)
        safe = np.where(r > 0, r, 1.0)[..., None]
        g = np.where(r[..., None] > 0, d / safe, 0.0)
        return r - radius, g

    return fn


def capsule_sdf(a, b, radius):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)

    def fn(P):
        ab = b - a
        t = np.clip(((P - a) @ ab) / (ab @ ab), 0.0, 1.0)
        cp = a + t[..., None] * ab
        d = P - cp
        r = np.linalg.norm(d, axis=-1)
        safe = np.where(r > 0, r, 1.0)[..., None]
        g = np.where(r[..., None] > 0, d / safe, 0.0)
        return r - radius, g

    return fn


def union_sdf(*fns):
    def fn(P):
        best_phi, best_g = fns[0](P)
        for f in fns[1:]:
            phi, g = f(P)
            m = phi < best_phi
            best_phi = np.where(m, phi, best_phi)
            best_g = np.where(m[..., None], g, best_g)
        return best_phi, best_g

    return fn


# ------------------------------------------------------------------ workloads
def sphere_mesh(level=6, radius=0.1, center=(0.0, 0.0, 0.0)):
    """Closed icosphere body mesh (20 * 4^level triangles, outward winding):
    vertices float32[n, 3], triangles int32[m, 3]."""
    p = (1 + 5 ** 0.5) / 2
    v = [(-1, p, 0), (1, p, 0), (-1, -p, 0), (1, -p, 0), (0, -1, p), (0, 1, p), (0, -1, -p), (0, 1, -p),
         (p, 0, -1), (p, 0, 1), (-p, 0, -1), (-p, 0, 1)]
    v = [np.array(x, np.float64) / np.linalg.norm(x) for x in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def midpoint(a, b):
            k = (min(a, b), max(a, b))
            if k not in mid:
                x = v[a] + v[b]
                v.append(x / np.linalg.norm(x))
                mid[k] = len(v) - 1
            return mid[k]

        g = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            g += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = g
    verts = np.array(v) * radius + np.asarray(center, np.float64)
    return verts.astype(F32), np.array(f, np.int32)


def config_cantilever():
    """1 strand x 64 vertices, 1 mm segments, clamped root, g = 9.81, E = 3 GPa,
    r = 40 um, rho = 1300, 10 iterations (dt = 1/600 s, 100 steps)."""
    return cantilever(64, 1e-3, clamp=True)


def config_groom(num_strands=10_000, verts=100, seed=1, sdf_dim=128, with_sdf=True):
    """Synthetic groom on a 0.1 m sphere scalp with the sphere's SDF on a
    sdf_dim^3 grid over [-1.5R, 1.5R]^3 (SDFThreshold 0)."""
    sc = groom(num_strands, verts, seed=seed)
    if with_sdf:
        R = 0.1
        sc.sdf = sdf_grid(sphere_sdf((0, 0, 0), R), (-1.5 * R,) * 3, (1.5 * R,) * 3, (sdf_dim,) * 3)
    return sc


def config_large(num_strands=100_000, seed=2, sdf_dim=256, with_sdf=True, curl=(2e-3, 8e-3), roots="random"):
    """100k strands with lognormal vertex counts (median 96, clipped 32-256) and
    a sphere + capsule (neck) SDF on a sdf_dim^3 grid; `curl` is the range of
    helix radii [m], `roots` the root layout of groom()."""
    counts = lognormal_counts(num_strands, seed=seed)
    sc = groom(num_strands, counts, seed=seed, curl=curl, roots=roots)
    sc.name = f"large{num_strands}"
    if with_sdf:
        R = 0.1
        fn = union_sdf(sphere_sdf((0, 0, 0), R), capsule_sdf((0, -0.05, 0), (0, -0.3, 0), 0.045))
        sc.sdf = sdf_grid(fn, (-0.4, -0.45, -0.4), (0.4, 0.35, 0.4), (sdf_dim,) * 3)
    return sc


def config_batched(num_grooms=64, strands=(1_000, 50_000), verts=(32, 256), seed=3, sdf_dim=128):
    """Independent grooms for the batched form: strand counts log-uniform in
    `strands`, vertex counts uniform in `verts` (inclusive), each groom on its
    own 0.1 m sphere with that sphere's SDF on a sdf_dim^3 grid and its own
    gravity direction (9.81 m/s^2, tilted up to ~35 degrees from -y)."""
    rng = np.random.default_rng(seed)
    lo, hi = np.log(strands[0]), np.log(strands[1])
    out = []
    R = 0.1
    sdf = sdf_grid(sphere_sdf((0, 0, 0), R), (-1.5 * R,) * 3, (1.5 * R,) * 3, (sdf_dim,) * 3) if sdf_dim else None
    for g in range(num_grooms):
        ns = int(round(np.exp(lo + (hi - lo) * rng.random())))
        counts = rng.integers(verts[0], verts[1] + 1, size=ns)
        sc = groom(ns, counts, seed=seed * 1000 + g)
        sc.name = f"batched{g}"
        # every groom owns its grid (same sphere, separate device copy per handle)
        sc.sdf = sdf
        d = np.array([rng.normal() * 0.35, -1.0, rng.normal() * 0.35])
        d /= np.linalg.norm(d)
        sc.material = Material(gravity=tuple(float(c) for c in 9.81 * d))
        out.append(sc)
    return out


def config_differentiation(num_strands=2_000, verts=64, seed=4):
    """Differentiation workload: groom of num_strands x verts (10 iterations
    per substep) and the per-strand parameters its targets come from: Young's
    modulus 3 GPa and rest-curvature scale 1, each perturbed by a seeded
    uniform +-20 %.  Returns (scene, E_target[S], kappa_target[S])."""
    sc = groom(num_strands, verts, seed=seed)
    sc.name = f"diff{num_strands}x{verts}"
    sc.material = Material(num_itr=10)
    rng = np.random.default_rng(seed)
    e_t = 3.0 * (1 + 0.2 * rng.uniform(-1, 1, num_strands))
    k_t = 1 + 0.2 * rng.uniform(-1, 1, num_strands)
    return sc, e_t, k_t


def lbvh_recipe_boxes(n=100_000, pad=0.001):
    """The reference's LBVH test set (Kitten::testLBVH,
    KE/KittenGpuLBVH/lbvh.cu:554-565): srand(1), point i = (rand(), rand(),
    rand()) / (float)RAND_MAX, each box the point padded by R = 0.001.  The C
    library's rand() is called through ctypes, so on glibc these are the
    reference's own boxes.  float32[n, 6] (min xyz, max xyz)."""
    import ctypes
    import ctypes.util

    libc = ctypes.CDLL(ctypes.util.find_library("c"))
    libc.rand.restype = ctypes.c_int
    libc.srand(1)
    rand_max = 2147483647
    raw = np.array([libc.rand() for _ in range(3 * n)], np.float64).reshape(n, 3)
    p = (raw.astype(np.float32) / np.float32(rand_max)).astype(np.float32)
    return np.concatenate([p - np.float32(pad), p + np.float32(pad)], axis=1).astype(np.float32)


def bundle(nx=6, ny=6, nv=32, spacing=6e-5, seg_len=1e-3, jitter=1e-5, seed=7):
    """A grid of nx x ny horizontal strands (along +x) with lateral spacing
    `spacing`; neighbours sit inside the contact barrier (hair-hair contacts)."""
    rng = np.random.default_rng(seed)
    ys, zs = np.meshgrid(np.arange(ny) * spacing, np.arange(nx) * spacing, indexing="ij")
    roots = np.stack([np.zeros(nx * ny), ys.ravel(), zs.ravel()], 1)
    x = np.arange(nv) * seg_len
    pos = roots[:, None, :] + np.stack([x, np.zeros(nv), np.zeros(nv)], 1)[None]
    pos[:, 1:, 1:] += rng.uniform(-jitter, jitter, size=(nx * ny, nv - 1, 2))
    off = np.arange(0, nx * ny * nv + 1, nv).astype(np.int32)
    return Scene(f"bundle{nx}x{ny}x{nv}", pos.reshape(-1, 3).astype(F32), off)


def ribbon(num_strands=12, nv=32, spacing=6e-5, seg_len=1e-3, jitter=1e-5, seed=7):
    """One row of horizontal strands `spacing` apart: each segment touches only
    the two neighbouring strands (inside the contact barrier, well below the
    16-contact list capacity)."""
    return bundle(1, num_strands, nv, spacing=spacing, seg_len=seg_len, jitter=jitter, seed=seed)


def contact_params(lib_default_params, material, **kw):
    """kscr_params with hair-hair contacts on (EnableHairHairColl: contacts and
    step limits, StableCosseratRodsSolver.cpp:184-186)."""
    p = make_params(lib_default_params, material, **kw)
    p.enable_hair_hair_coll = 1
    p.use_step_size_limit = 1
    return p
