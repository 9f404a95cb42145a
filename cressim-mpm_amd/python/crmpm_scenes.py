"""Synthetic MPM scenes with seeded inputs.

A scene is a plain dict of numpy arrays laid out exactly as the reference
engine binds them to a solver (include/solver/particle_data.h:44-58,
include/solver/material_data.h:50-64, include/engine/shape.h:58-89,
include/engine/geometry.h:65-118):

  particles: xm float32[N,4] (position, mass), v float32[N,4], params float32[N,4],
             type int32[N], mask uint8[N]
  shapes:    ids int32[S], type int32[cap], geom int32[cap], and float32[cap,4]
             tables pos, rot (quaternion x,y,z,w), invScale, linVel, angVel,
             comInvMass, inertia0, inertia1, params0 (smooth, drag, friction, fatten)
  geometry:  gtype int32[G], gparams float32[G,4]
  sdf:       sdf_dims int32[K,4], sdf_lbcs float32[K,4], sdf_data float32[D,4]

Random numbers come from SplitMix64; a float is (u >> 40) * 2^-24.
"""
import math

import numpy as np

M64 = (1 << 64) - 1

# enums (include/solver/mpm_solver.h:51-61, include/engine/geometry.h:42-55,
# include/engine/shape.h:43-48, include/solver/material_data.h:41-46)
MLS, PB, STD = 0, 1, 2
BOX, SPHERE, PLANE, CAPSULE, TRIMESH, QUAD_SLICER, TRIMESH_SLICER, LINES, ARC = range(9)
STATIC, KINEMATIC, DYNAMIC = range(3)
COROTATIONAL, NEOHOOKEAN, LINEAR = range(3)


class SplitMix64:
    """SplitMix64 stream (Steele, Lea, Flood 2014)."""

    def __init__(self, seed):
        self.state = np.uint64(seed & M64)

    def next_u64(self, n):
        with np.errstate(over="ignore"):
            k = np.arange(1, n + 1, dtype=np.uint64)
            z = self.state + k * np.uint64(0x9E3779B97F4A7C15)
            self.state = self.state + np.uint64(n) * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.next_u64(n) >> np.uint64(40)).astype(np.float64) * (2.0 ** -24)
        return lo + (hi - lo) * u


def lame(E, nu):
    """Lame parameters as the reference's tests compute them (tests/test_benchmark/main.cu:111-114)."""
    lam = (E * nu) / ((1 + nu) * (1 - 2 * nu))
    mu = E / (2 * (1 + nu))
    return lam, mu


def new_scene(kind, dx, lo, hi, n_particles, dt, gravity=(0.0, -9.81, 0.0), iterations=10):
    return dict(kind=kind, dx=float(dx), lo=np.asarray(lo, np.float32), hi=np.asarray(hi, np.float32),
                gravity=tuple(float(g) for g in gravity), dt=float(dt), iterations=int(iterations),
                contact_model=0,
                xm=np.zeros((n_particles, 4), np.float32), v=np.zeros((n_particles, 4), np.float32),
                params=np.zeros((n_particles, 4), np.float32), type=np.zeros(n_particles, np.int32),
                mask=np.ones(n_particles, np.uint8), shape_list=[], geoms=[],
                sdf_dims=np.zeros((1, 4), np.int32), sdf_lbcs=np.zeros((1, 4), np.float32),
                sdf_data=np.zeros((1, 4), np.float32), frames=[])


def grid_dims(scene):
    """Node count per axis, int(extent/dx + 1) in FP32 (src/solver/gpu_mls_mpm_solver.cu:48-50)."""
    ext = scene["hi"] - scene["lo"]
    return tuple(int(np.float32(e) / np.float32(scene["dx"]) + np.float32(1.0)) for e in ext)


def add_shape(scene, gtype, gparams, pos=(0, 0, 0), rot=(0, 0, 0, 1), stype=STATIC, params0=(0.0, 1.0, 0.0, 0.0),
              inv_scale=(1, 1, 1), lin_vel=(0, 0, 0), ang_vel=(0, 0, 0), com_inv_mass=(0, 0, 0, 0),
              inertia0=(0, 0, 0, 0), inertia1=(0, 0, 0, 0)):
    """Append a geometry and a shape that uses it; returns the shape id."""
    scene["geoms"].append((int(gtype), tuple(float(x) for x in gparams)))
    sid = len(scene["shape_list"])
    scene["shape_list"].append(dict(type=int(stype), geom=len(scene["geoms"]) - 1, pos=tuple(pos) + (0.0,),
                                    rot=tuple(rot), invScale=tuple(inv_scale) + (0.0,),
                                    linVel=tuple(lin_vel) + (0.0,), angVel=tuple(ang_vel) + (0.0,),
                                    comInvMass=tuple(com_inv_mass), inertia0=tuple(inertia0),
                                    inertia1=tuple(inertia1), params0=tuple(params0)))
    return sid


def int_bits(i):
    """An int32 carried in a float32 geometry parameter (CR_FLOAT_AS_INT,
    include/math/type_punning.h:52-58)."""
    return float(np.array([i], np.int32).view(np.float32)[0])


def add_sdf_grid(scene, dims, lower, cell, data):
    """Append one SDF table (GeometrySdfData, include/engine/geometry.h:113-118):
    dims (x, y, z), lower bound and cell size, and `data` float4 rows (gradient,
    value).  Returns (offset of the first row in sdf_data, table id)."""
    off, sid = len(scene["sdf_data"]), len(scene["sdf_dims"])
    scene["sdf_dims"] = np.vstack([scene["sdf_dims"], np.array([[dims[0], dims[1], dims[2], 0]], np.int32)])
    scene["sdf_lbcs"] = np.vstack([scene["sdf_lbcs"], np.array([[lower[0], lower[1], lower[2], cell]], np.float32)])
    scene["sdf_data"] = np.vstack([scene["sdf_data"], np.asarray(data, np.float32).reshape(-1, 4)])
    return off, sid


def _grid_points(dims, lower, cell):
    # node (i, j, k) at row i + j nx + k nx ny (getGridIndex, include/geometry_utils/sdf.h)
    k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    return np.stack([i.ravel(), j.ravel(), k.ravel()], 1) * np.float32(cell) + np.asarray(lower, np.float32)


def sphere_sdf_grid(radius, cell, n):
    """Sampled SDF of a sphere as a triangle-mesh SDF table: n^3 nodes centred
    on the origin, rows (unit gradient, |p| - radius)."""
    lower = (-(n - 1) / 2 * cell,) * 3
    pts = _grid_points((n, n, n), lower, cell).astype(np.float64)
    r = np.linalg.norm(pts, axis=1)
    g = np.where(r[:, None] > 0, pts / np.maximum(r, 1e-12)[:, None], np.array([0.0, 1.0, 0.0]))
    return (n, n, n), lower, cell, np.concatenate([g, (r - radius)[:, None]], 1)


def planar_slicer_grid(half_x, half_z, cell, spine):
    """Mesh-slicer table of a planar sheet y = 0 (|x| <= half_x): a spine row
    (normal, offset) then the grid of signed distances y with gradient
    (0, sign y, 0), NaN beyond the sheet's x extent (no surface there)."""
    nx, ny, nz = int(round(2 * (half_x + 2 * cell) / cell)) + 1, 7, int(round(2 * (half_z + 2 * cell) / cell)) + 1
    lower = (-(nx - 1) / 2 * cell, -(ny - 1) / 2 * cell, -(nz - 1) / 2 * cell)
    pts = _grid_points((nx, ny, nz), lower, cell).astype(np.float64)
    val = pts[:, 1].copy()
    grad = np.zeros((len(pts), 3))
    grad[:, 1] = np.where(val >= 0, 1.0, -1.0)
    val[np.abs(pts[:, 0]) > half_x] = np.nan
    rows = np.concatenate([grad, val[:, None]], 1)
    return (nx, ny, nz), lower, cell, np.vstack([np.asarray(spine, np.float64)[None, :], rows])


def icosphere(radius, subdivisions=2, center=(0.0, 0.0, 0.0)):
    """Closed triangle mesh of a sphere: an icosahedron with each face split
    into 4 `subdivisions` times, vertices projected to the sphere.
    Returns float32 points (n, 3) and uint32 triangles (m, 3), outward winding."""
    t = (1.0 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    pts = [np.asarray(x, np.float64) / np.linalg.norm(x) for x in v]
    for _ in range(subdivisions):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                q = pts[a] + pts[b]
                pts.append(q / np.linalg.norm(q))
                mid[key] = len(pts) - 1
            return mid[key]

        f = [t2 for a, b, c in f for t2 in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                            (m(a, b), m(b, c), m(c, a)))]
    P = (np.asarray(pts) * radius + np.asarray(center)).astype(np.float32)
    return P, np.asarray(f, np.uint32)


def torus(major, minor, nu=24, nv=12):
    """Closed triangle mesh of a torus around the y axis, outward winding."""
    u = np.arange(nu) * 2 * np.pi / nu
    v = np.arange(nv) * 2 * np.pi / nv
    U, V = np.meshgrid(u, v, indexing="ij")
    x = (major + minor * np.cos(V)) * np.cos(U)
    z = (major + minor * np.cos(V)) * np.sin(U)
    y = minor * np.sin(V)
    P = np.stack([x.ravel(), y.ravel(), z.ravel()], 1).astype(np.float32)
    idx = lambda i, j: (i % nu) * nv + (j % nv)  # noqa: E731
    tris = []
    for i in range(nu):
        for j in range(nv):
            a, b, c, d = idx(i, j), idx(i + 1, j), idx(i + 1, j + 1), idx(i, j + 1)
            tris += [(a, c, b), (a, d, c)]
    return P, np.asarray(tris, np.uint32)


def geometry_bounds(gtype, gparams, sdf=None):
    """Geometry::getBounds as SimulationFactoryImpl::createGeometry sets it
    (src/engine/simulation_factory_impl.cpp:521-577): box, sphere, capsule,
    quad slicer; a triangle-mesh table (sdf = (dims, lower, cell, rows)) spans
    its nodes (lower bound + (dims - 1) cell)."""
    p = [float(x) for x in gparams]
    if gtype == BOX:
        return (-p[0], -p[1], -p[2]), (p[0], p[1], p[2])
    if gtype == SPHERE:
        return (-p[0],) * 3, (p[0],) * 3
    if gtype == CAPSULE:
        return (-p[0], -p[0] - p[1], -p[0]), (p[0], p[0] + p[1], p[0])
    if gtype == QUAD_SLICER:
        return (-p[0], 0.0, -p[1]), (p[0], 0.0, p[1])
    if gtype == TRIMESH and sdf is not None:
        dims, lower, cell = sdf[0], sdf[1], sdf[2]
        return tuple(lower), tuple(lower[k] + (dims[k] - 1) * cell for k in range(3))
    raise ValueError("unbounded geometry")


def lattice_object(gtype, gparams, spacing, mass, position=(0, 0, 0), rotation=(0, 0, 0, 1), inv_scale=(1, 1, 1),
                   sdf=None):
    """A particle object's inputs (ParticleObjectDesc, include/engine/particle_object.h):
    geometry, spacing, mass and pose.  sdf: a table from sphere_sdf_grid for
    a triangle-mesh geometry; gparams then carries its offset and id."""
    o = dict(type=int(gtype), spacing=float(spacing), mass=float(mass), position=tuple(position),
             rotation=tuple(rotation), inv_scale=tuple(inv_scale), sdf_dims=np.zeros((1, 4), np.int32),
             sdf_lbcs=np.zeros((1, 4), np.float32), sdf_data=np.zeros((1, 4), np.float32))
    if sdf is not None:
        off, tid = add_sdf_grid(o, *sdf)
        gparams = (int_bits(off), int_bits(tid), 0.0, 0.0)
    o["params0"] = tuple(float(x) for x in gparams)
    o["bounds_min"], o["bounds_max"] = geometry_bounds(gtype, gparams, sdf)
    return o


SHAPE_F4 = ("pos", "rot", "invScale", "linVel", "angVel", "comInvMass", "inertia0", "inertia1", "params0")


def shape_tables(scene):
    """ShapeData / GeometryData tables (at least one row each, so no pointer is null)."""
    sl = scene["shape_list"]
    cap = max(1, len(sl))
    t = dict(ids=np.arange(len(sl), dtype=np.int32), type=np.zeros(cap, np.int32), geom=np.zeros(cap, np.int32))
    for k in SHAPE_F4:
        t[k] = np.zeros((cap, 4), np.float32)
    t["rot"][:, 3] = 1.0
    for i, s in enumerate(sl):
        t["type"][i] = s["type"]
        t["geom"][i] = s["geom"]
        for k in SHAPE_F4:
            t[k][i] = s[k]
    g = scene["geoms"] or [(BOX, (0.0, 0.0, 0.0, 0.0))]
    t["gtype"] = np.array([x[0] for x in g], np.int32)
    t["gparams"] = np.array([x[1] for x in g], np.float32)
    return t


def apply_frame(scene, frame):
    """Kinematic target for one frame: teleport to the frame's pose, set its
    velocity (src/engine/shape_impl.cpp:72-107).  A frame may move several
    shapes ("moves": [(shape, pos, lin_vel)]) and change geometry parameters
    ("geoms": [(geometry, params)], Geometry::setParams)."""
    moves = frame.get("moves") or [(frame["shape"], frame["pos"], frame["lin_vel"])]
    for sid, pos, vel in moves:
        s = scene["shape_list"][sid]
        s["pos"] = tuple(pos) + (0.0,)
        s["linVel"] = tuple(vel) + (0.0,)
    for gi, params in frame.get("geoms", []):
        scene["geoms"][gi] = (scene["geoms"][gi][0], tuple(float(x) for x in params))


def lattice_block(rng, n, origin, spacing, jitter):
    """n = (nx, ny, nz) lattice, x fastest; jitter is the half-width of a uniform offset per axis."""
    nx, ny, nz = n
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ii.transpose(2, 1, 0).ravel(), jj.transpose(2, 1, 0).ravel(), kk.transpose(2, 1, 0).ravel()], 1)
    pts = np.asarray(origin, np.float64) + (idx + 0.5) * np.asarray(spacing, np.float64)
    if jitter:
        pts = pts + rng.uniform(pts.size, -1.0, 1.0).reshape(pts.shape) * np.asarray(jitter, np.float64)
    return pts


def config1(n_side=32, steps=200, seed=20261203 + 1):
    """Plain MLS elastic cube on an SDF ground plane: 32^3 = 32,768 particles,
    64^3-cell unit box, StVK (co-rotational falls through to it), dt = 1e-4."""
    rng = SplitMix64(seed)
    dx = 1.0 / 64
    spacing = 1.0 / 128
    side = n_side * spacing
    N = n_side ** 3
    sc = new_scene(MLS, dx, (0, 0, 0), (1, 1, 1), N, 1e-4)
    origin = (0.5 - side / 2, 0.1 + dx / 2, 0.5 - side / 2)
    pts = lattice_block(rng, (n_side,) * 3, origin, (spacing,) * 3, 0.25 * dx)
    sc["xm"][:, :3] = pts
    sc["xm"][:, 3] = 1000.0 * spacing ** 3
    sc["v"][:, 1] = -0.5
    lam, mu = lame(1e4, 0.3)
    sc["params"][:, 0] = lam
    sc["params"][:, 1] = mu
    sc["type"][:] = COROTATIONAL
    add_shape(sc, PLANE, (0, 0, 0, 0), pos=(0, 0.1, 0), params0=(0.0, 1.0, 0.0, 0.0))
    sc["steps"] = steps
    return sc


def config2(nx=128, ny=16, nz=128, steps=500, iterations=20, seed=20261203 + 2, frame_substeps=10, embed=0.01):
    """Plain PB-MPM tissue slab: stratified nx x ny x nz particles, the full
    128 x 16 x 128 = 262,144-particle slab 0.961 x 0.2 x 0.961 m, smaller
    counts give a smaller slab at the same spacing, centred in x/z.
    Grid [0,1] x [0,0.5] x [0,1], dx = 1/128 (129 x 65 x 129 nodes).
    The slab keeps 2.5 dx from every lower grid face: a particle closer than
    dx/2 to a lower face gets base cell trunc(s - 0.5) = 0 with fx < 0.5 and
    weights that do not sum to one (src/solver/mls_mpm_algorithm.h:172-174),
    and a slab filling the whole 1 m box blows up in the reference solver
    itself within 3 substeps.  Neo-Hookean with PB parameters (alpha, r) =
    (0.5/dt, 0.9); a kinematic sphere indenter of radius 0.1 m (scaled with
    the slab) starts with its lowest point `embed` (10 mm) below the top and
    moves down at 0.05 m/s, friction 0.3; dt = 2e-3, 20 iterations per
    substep.  The slab's top falls at nearly free fall for its first ~40
    substeps (the PB constraint and the F update act on the dx-scaled affine
    velocity C, src/solver/pb_mpm_algorithm.h:329,360,407, so the slab
    offers little elastic support at dx = 1/128), faster than the indenter
    descends: a sphere starting above the top would never touch it.  Starting
    embedded, it is in contact with the slab for the first ~35 substeps."""
    rng = SplitMix64(seed)
    dx = 1.0 / 128
    dt = 2e-3
    N = nx * ny * nz
    sc = new_scene(PB, dx, (0, 0, 0), (1, 0.5, 1), N, dt, iterations=iterations)
    span = 1.0 - 5 * dx
    cell = (span / 128, 0.2 / 16, span / 128)
    slab = (nx * cell[0], ny * cell[1], nz * cell[2])
    y0 = 2.5 * dx
    origin = (0.5 - slab[0] / 2, y0, 0.5 - slab[2] / 2)
    ii, jj, kk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    idx = np.stack([ii.transpose(2, 1, 0).ravel(), jj.transpose(2, 1, 0).ravel(), kk.transpose(2, 1, 0).ravel()], 1)
    u = rng.uniform(3 * N).reshape(N, 3)
    pts = np.asarray(origin) + (idx + u) * np.asarray(cell)
    sc["xm"][:, :3] = pts
    sc["xm"][:, 3] = 1000.0 * (cell[0] * cell[1] * cell[2])
    sc["params"][:, 0] = 0.5 / dt
    sc["params"][:, 1] = 0.9
    sc["type"][:] = NEOHOOKEAN
    top = y0 + slab[1]
    r = 0.1 * min(nx, nz) / 128
    c0 = (0.5, top - embed + r, 0.5)
    add_shape(sc, SPHERE, (r, 0, 0, 0), pos=c0, stype=KINEMATIC, params0=(0.0, 1.0, 0.3, 0.0))
    sc["steps"] = steps
    # Kinematic target per frame of `frame_substeps` substeps: the shape
    # teleports to the frame's end pose and moves at (0, -0.05, 0) during it
    # (src/engine/shape_impl.cpp:72-107).
    speed = 0.05
    frames = []
    for f in range((steps + frame_substeps - 1) // frame_substeps):
        y = c0[1] - speed * dt * frame_substeps * (f + 1)
        frames.append(dict(substeps=min(frame_substeps, steps - f * frame_substeps),
                           shape=0, pos=(c0[0], y, c0[2]), lin_vel=(0.0, -speed, 0.0)))
    sc["frames"] = frames
    return sc


def small_scene(kind, seed, n_side=8, shapes="plane", dx=1.0 / 32, jitter=0.25, v0=(0.0, -0.5, 0.0),
                material=COROTATIONAL, E=1e4, nu=0.3):
    """Small randomized block used by the parity tests (8^3 particles by default)."""
    rng = SplitMix64(seed)
    spacing = dx / 2
    N = n_side ** 3
    dt = 1e-4 if kind != PB else 1e-3
    sc = new_scene(kind, dx, (0, 0, 0), (1, 1, 1), N, dt, iterations=4)
    side = n_side * spacing
    origin = (0.5 - side / 2, 0.2, 0.5 - side / 2)
    pts = lattice_block(rng, (n_side,) * 3, origin, (spacing,) * 3, jitter * spacing)
    sc["xm"][:, :3] = pts
    sc["xm"][:, 3] = 1000.0 * spacing ** 3
    vr = rng.uniform(3 * N, -0.1, 0.1).reshape(N, 3)
    sc["v"][:, :3] = np.asarray(v0) + vr
    if kind == PB:
        sc["params"][:, 0] = 0.5 / dt
        sc["params"][:, 1] = 0.9
        sc["type"][:] = NEOHOOKEAN if material == COROTATIONAL else material
    else:
        lam, mu = lame(E, nu)
        sc["params"][:, 0] = lam
        sc["params"][:, 1] = mu
        sc["type"][:] = material
    if "plane" in shapes:
        add_shape(sc, PLANE, (0, 0, 0, 0), pos=(0, 0.19, 0), params0=(0.02, 1.0, 0.3, 0.0))
    if "sphere" in shapes:
        add_shape(sc, SPHERE, (0.05, 0, 0, 0), pos=(0.5, 0.2 + side + 0.04, 0.5), stype=KINEMATIC,
                  params0=(0.01, 0.9, 0.5, 0.0), lin_vel=(0.0, -0.3, 0.0))
    if "box" in shapes:
        a = math.radians(20.0)
        add_shape(sc, BOX, (0.03, 0.02, 0.04, 0), pos=(0.5 + side / 2, 0.22, 0.5),
                  rot=(0.0, 0.0, math.sin(a / 2), math.cos(a / 2)), stype=KINEMATIC,
                  params0=(0.01, 1.0, 0.2, 0.002), lin_vel=(-0.2, 0.0, 0.0), ang_vel=(0, 0.5, 0))
    if "capsule" in shapes:
        add_shape(sc, CAPSULE, (0.01, 0.03, 0, 0), pos=(0.5, 0.2 + side * 0.8, 0.5 - side / 2), stype=DYNAMIC,
                  params0=(0.01, 1.0, 0.1, 0.0), lin_vel=(0, 0, 0.2), inv_scale=(1.0, 0.8, 1.25))
    parts = shapes.split("+")
    if "slicer" in parts:
        a = math.radians(90.0)
        # vertical quad slicer through the block centre (local y is the cut normal)
        add_shape(sc, QUAD_SLICER, (side / 3, side, 0, 0), pos=(0.5 + 0.3 * spacing, 0.2 + side, 0.5),
                  rot=(0.0, 0.0, math.sin(a / 2), math.cos(a / 2)), stype=KINEMATIC,
                  params0=(0.0, 1.0, 0.0, 0.001))
    if "mesh" in parts:
        # triangle-mesh SDF table (a sampled sphere) pressing on the block, a
        # polyline under it and an arc around one side at a node row (the
        # last two act on grid nodes only)
        off, tid = add_sdf_grid(sc, *sphere_sdf_grid(0.05, 0.01, 17))
        add_shape(sc, TRIMESH, (int_bits(off), int_bits(tid), 0, 0), pos=(0.5, 0.2 + side + 0.04, 0.5),
                  stype=KINEMATIC, params0=(0.01, 0.9, 0.5, 0.0), lin_vel=(0.0, -0.3, 0.0))
        y = 0.197
        pts = np.array([[0.5 - side, y, 0.5 - side / 4], [0.5, y - 0.002, 0.5], [0.5 + side, y, 0.5 + side / 4]])
        lo, hi = pts.min(0) - 0.03, pts.max(0) + 0.03
        off, tid = add_sdf_grid(sc, (0, 0, 0), lo, 0.0, np.vstack([np.append(hi, 0.0), np.c_[pts, np.zeros(3)]]))
        add_shape(sc, LINES, (int_bits(off), int_bits(tid), int_bits(4), 0), params0=(0.03, 1.0, 0.4, 0.003))
        r = side / 2 + 0.01
        add_shape(sc, ARC, (-math.pi / 2, math.pi / 2, 0, 0), pos=(0.5, 0.25, 0.5),
                  inv_scale=(1 / r, 1 / r, 1 / r), params0=(0.02, 1.0, 0.2, 0.002))
    if "meshslicer" in parts:
        # a mesh slicer (planar sheet table with a spine) through the block centre
        a = math.radians(90.0)
        off, tid = add_sdf_grid(sc, *planar_slicer_grid(side / 3, side, spacing, (0.0, 0.0, 1.5, -side / 2)))
        add_shape(sc, TRIMESH_SLICER, (int_bits(off), int_bits(tid), 0, 0), pos=(0.5 + 0.3 * spacing, 0.2 + side, 0.5),
                  rot=(0.0, 0.0, math.sin(a / 2), math.cos(a / 2)), stype=KINEMATIC,
                  params0=(0.0, 1.0, 0.0, 0.001))
    sc["steps"] = 10
    return sc


def config3(n_scenes=64, kind=MLS, steps=100, seed=20261203 + 3, n_min=2000, n_max=200000, iterations=20, gap=2.0):
    """Batched scenes: n_scenes independent soft bodies with particle counts
    log-uniform in [n_min, n_max] and grids of G = 4 round(U(8, 32)) cells per
    axis over the unit box.  Each body is a box or a sphere (equal odds) at
    8 particles per cell (lattice spacing dx / 2, enlarged when the body would
    exceed half the box), E ~ U(5e3, 5e4), nu ~ U(0.3, 0.45), falling at
    0.5 m/s onto one static SDF box (half extents U(0.05, 0.2)) or capsule
    (radius U(0.02, 0.08), half height U(0.05, 0.2)) below it, the body's
    lowest lattice row `gap` cells above the obstacle's top.  MLS uses the
    co-rotational model (StVK, as the reference evaluates it); PB uses
    Neo-Hookean with (alpha, r) = (0.5 / dt, 0.9) and `iterations` iterations.
    dt = 1e-4."""
    rng = SplitMix64(seed)
    dt = 1e-4
    scenes = []
    for s in range(n_scenes):
        u = rng.uniform(9)
        n_target = int(round(math.exp(math.log(n_min) + u[0] * (math.log(n_max) - math.log(n_min)))))
        G = 4 * int(round(8 + u[1] * 24))
        dx = 1.0 / G
        is_sphere = u[2] < 0.5
        E = 5e3 + u[3] * 4.5e4
        nu = 0.3 + u[4] * 0.15
        h = dx / 2
        if is_sphere:
            k = (3.0 * n_target / (4.0 * math.pi)) ** (1.0 / 3.0)
            radius = min(k * h, 0.25)
            h = radius / k
            m = int(math.ceil(2 * radius / h)) + 1
            ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
            off = (np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1) + 0.5) * h - m * h / 2
            pts = off[np.linalg.norm(off, axis=1) <= radius]
            half = (radius, radius, radius)
        else:
            n_side = max(2, int(round(n_target ** (1.0 / 3.0))))
            h = min(h, 0.5 / n_side)
            ii, jj, kk = np.meshgrid(np.arange(n_side), np.arange(n_side), np.arange(n_side), indexing="ij")
            pts = (np.stack([ii.ravel(), jj.ravel(), kk.ravel()], 1) + 0.5) * h - n_side * h / 2
            half = (n_side * h / 2,) * 3
        N = len(pts)
        jit = rng.uniform(3 * N, -0.25, 0.25).reshape(N, 3) * h
        sc = new_scene(kind, dx, (0, 0, 0), (1, 1, 1), N, dt, iterations=iterations)
        # obstacle centred under the body, its top at y = 0.3
        if u[5] < 0.5:
            hx, hy, hz = 0.05 + u[6] * 0.15, 0.05 + u[7] * 0.15, 0.05 + u[8] * 0.15
            obs_top = 0.3
            add_shape(sc, BOX, (hx, hy, hz, 0), pos=(0.5, obs_top - hy, 0.5), params0=(0.0, 1.0, 0.3, 0.0))
        else:
            r = 0.02 + u[6] * 0.06
            hh = 0.05 + u[7] * 0.15
            obs_top = 0.3
            # capsule axis along local y rotated to lie along x
            a = math.radians(90.0)
            add_shape(sc, CAPSULE, (r, hh, 0, 0), pos=(0.5, obs_top - r, 0.5),
                      rot=(0.0, 0.0, math.sin(a / 2), math.cos(a / 2)), params0=(0.0, 1.0, 0.3, 0.0))
        centre = np.array([0.5, obs_top + gap * dx + half[1], 0.5])
        sc["xm"][:, :3] = centre + pts + jit
        sc["xm"][:, 3] = 1000.0 * h ** 3
        sc["v"][:, 1] = -0.5
        if kind == PB:
            sc["params"][:, 0] = 0.5 / dt
            sc["params"][:, 1] = 0.9
            sc["type"][:] = NEOHOOKEAN
        else:
            lam, mu = lame(E, nu)
            sc["params"][:, 0] = lam
            sc["params"][:, 1] = mu
            sc["type"][:] = COROTATIONAL
        sc["steps"] = steps
        scenes.append(sc)
    return scenes


def grf(rng, points, length=0.1, features=256):
    """Smooth Gaussian random field at the points (random Fourier features,
    Rahimi and Recht 2007): sqrt(2/K) sum_k cos(w_k . x + b_k), w_k ~ N(0, I/l^2),
    b_k ~ U(0, 2 pi); normals by Box-Muller from the SplitMix64 stream."""
    u1 = rng.uniform(3 * features)
    u2 = rng.uniform(3 * features)
    z = np.sqrt(-2.0 * np.log(np.maximum(u1, 1e-300))) * np.cos(2.0 * np.pi * u2)
    w = z.reshape(features, 3) / length
    b = rng.uniform(features, 0.0, 2.0 * np.pi)
    return np.sqrt(2.0 / features) * np.cos(np.asarray(points, np.float64) @ w.T + b).sum(1)


def gaussian_noise(n, seed, sigma):
    """(n, 3) Gaussian noise of standard deviation sigma, Box-Muller over the
    SplitMix64 stream (measurement noise on the differentiable block's targets)."""
    rng = SplitMix64(seed)
    u1 = np.maximum(rng.uniform(3 * n), 1e-300)
    u2 = rng.uniform(3 * n)
    return (sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).reshape(n, 3)


# targets of the differentiable block: final positions under the hidden log E
# field, plus 1 mm of measurement noise per coordinate
HIDDEN_FIELD_SEED = 20261203 + 44
TARGET_NOISE_SEED = 20261203 + 45
TARGET_NOISE = 1e-3


def config4(nx=64, ny=32, nz=32, steps=100, seed=20261203 + 4, field_seed=None):
    """Differentiable MLS block: nx x ny x nz = 65,536 particles at spacing dx/2
    on a 96^3-cell unit box (dx = 1/96), per-particle log E from a smooth
    Gaussian random field (length scale 0.1 m) mapped to [ln 5e3, ln 5e4],
    nu = 0.3, co-rotational (StVK as the reference evaluates it), v0 =
    (0, -0.5, 0) + 0.05 * GRF per component, dt = 1e-4.  field_seed picks
    the log E field (the hidden ground truth uses another seed); returns the
    scene and the per-particle log E."""
    rng = SplitMix64(seed)
    dx = 1.0 / 96
    h = dx / 2
    N = nx * ny * nz
    sc = new_scene(MLS, dx, (0, 0, 0), (1, 1, 1), N, 1e-4)
    size = np.array([nx, ny, nz]) * h
    origin = np.array([0.5, 0.5, 0.5]) - size / 2
    pts = lattice_block(rng, (nx, ny, nz), origin, (h, h, h), 0.0)
    sc["xm"][:, :3] = pts
    sc["xm"][:, 3] = 1000.0 * h ** 3
    frng = SplitMix64(seed + 1000 if field_seed is None else field_seed)
    f = grf(frng, pts)
    f = (f - f.min()) / max(f.max() - f.min(), 1e-12)
    log_e = math.log(5e3) + f * (math.log(5e4) - math.log(5e3))
    vrng = SplitMix64(seed + 2000)
    v0 = np.stack([grf(vrng, pts) for _ in range(3)], 1) * 0.05
    v0[:, 1] += -0.5
    sc["v"][:, :3] = v0
    lam, mu = lame(np.exp(log_e), 0.3)
    sc["params"][:, 0] = lam
    sc["params"][:, 1] = mu
    sc["type"][:] = COROTATIONAL
    sc["steps"] = steps
    return sc, log_e


def config5(nx=256, ny=32, nz=256, steps=500, seed=20261203 + 5, frame_substeps=10, speed=0.02):
    """MLS at scale with cutting: a 256 x 32 x 256 = 2,097,152-particle tissue
    block (spacing 5 mm, 8 particles per cell, 1.28 x 0.16 x 1.28 m) on a
    [0, 2.56] x [0, 1.28] x [0, 2.56] m grid with dx = 0.01 (257 x 129 x 257
    nodes), co-rotational (StVK) E = 2e4, nu = 0.3, dt = 1e-4.  A kinematic
    vertical capsule tool (radius 5 mm) 2 cm deep in the top surface sweeps
    along +x at 0.02 m/s through the block's mid plane z = 1.28; a kinematic
    quad slicer in that plane spans the swept part of the cut (its half
    extent grows with the swept length, one Geometry::setParams per frame).
    The block keeps 2.5 dx from the lower faces (base-cell weights)."""
    rng = SplitMix64(seed)
    dx = 0.01
    h = 0.005
    dt = 1e-4
    N = nx * ny * nz
    sc = new_scene(MLS, dx, (0, 0, 0), (2.56, 1.28, 2.56), N, dt)
    size = np.array([nx, ny, nz]) * h
    origin = np.array([1.28 - size[0] / 2, 2.5 * dx, 1.28 - size[2] / 2])
    sc["xm"][:, :3] = lattice_block(rng, (nx, ny, nz), origin, (h, h, h), 0.25 * h)
    sc["xm"][:, 3] = 1000.0 * h ** 3
    lam, mu = lame(2e4, 0.3)
    sc["params"][:, 0] = lam
    sc["params"][:, 1] = mu
    sc["type"][:] = COROTATIONAL
    top = origin[1] + size[1]
    depth, r, hh = 0.02, 0.005, 0.05
    x0 = origin[0] - 2 * r
    zc = 1.28
    tool = add_shape(sc, CAPSULE, (r, hh, 0, 0), pos=(x0, top - depth + r + hh, zc), stype=KINEMATIC,
                     params0=(0.0, 1.0, 0.3, 0.0), lin_vel=(speed, 0, 0))
    # quad in the local xz plane; +90 degrees about x puts its normal (local y) along z
    q = (math.sin(math.pi / 4), 0.0, 0.0, math.cos(math.pi / 4))
    slicer = add_shape(sc, QUAD_SLICER, (0.0, depth / 2 + 0.005, 0, 0), pos=(x0, top - depth / 2, zc), rot=q,
                       stype=KINEMATIC, params0=(0.0, 1.0, 0.0, 0.0))
    g_slicer = sc["shape_list"][slicer]["geom"]
    frames = []
    for f in range((steps + frame_substeps - 1) // frame_substeps):
        t_end = dt * frame_substeps * (f + 1)
        xt = x0 + speed * t_end
        frames.append(dict(substeps=min(frame_substeps, steps - f * frame_substeps),
                           moves=[(tool, (xt, top - depth + r + hh, zc), (speed, 0.0, 0.0)),
                                  (slicer, ((x0 + xt) / 2, top - depth / 2, zc), (speed / 2, 0.0, 0.0))],
                           geoms=[(g_slicer, ((xt - x0) / 2, depth / 2 + 0.005, 0.0, 0.0))]))
    sc["frames"] = frames
    sc["steps"] = steps
    return sc
