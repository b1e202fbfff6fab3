"""Seeded synthetic scenes for the contact hot path.

Meshes: icosphere (subdivision k), 24x12 torus, 12-triangle cube, and a
two-triangle ground plane. Bodies are set up as the reference's RigidBody
constructor does (src/physics/rigid_body.cpp:17-181): vertices centred on
the centre of mass and rotated into the principal-axis frame, mass and
principal moments from the closed triangle mesh, and the world pose
composed with that frame. The eigen-solver's axis choice for isotropic
bodies differs from Eigen's; the hot path consumes whatever body-local
frame the scene carries, so this only changes the scene, not the method.
"""
import types

import numpy as np


# ------------------------------------------------------------------ meshes
def icosphere(k, radius):
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]
    F = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
         [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    for _ in range(k):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = V[a] + V[b]
                V.append(p / np.linalg.norm(p))
                mid[key] = len(V) - 1
            return mid[key]
        nf = []
        for a, b, c in F:
            ab, bc, ca = m(a, b), m(b, c), m(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        F = nf
    return np.array(V) * radius, np.array(F, np.int64)


def torus(R, r, nu=24, nv=12):
    V, F = [], []
    for i in range(nu):
        u = 2 * np.pi * i / nu
        for j in range(nv):
            v = 2 * np.pi * j / nv
            V.append([(R + r * np.cos(v)) * np.cos(u), r * np.sin(v), (R + r * np.cos(v)) * np.sin(u)])
    for i in range(nu):
        for j in range(nv):
            a = i * nv + j
            b = ((i + 1) % nu) * nv + j
            c = ((i + 1) % nu) * nv + (j + 1) % nv
            d = i * nv + (j + 1) % nv
            F += [[a, c, b], [a, d, c]]
    return np.array(V), np.array(F, np.int64)


def cube(size=1.0):
    h = size / 2
    V = np.array([[x, y, z] for x in (-h, h) for y in (-h, h) for z in (-h, h)])
    F = [[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5], [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],
         [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]]
    return V, np.array(F, np.int64)


def plane(size=20.0):
    h = size / 2
    V = np.array([[-h, 0, -h], [h, 0, -h], [h, 0, h], [-h, 0, h]], float)
    F = np.array([[0, 2, 1], [0, 3, 2]], np.int64)
    return V, F


_MESH_CACHE = {}


def _cached(kind, arrays, fn):
    """fn(*arrays), memoised on the arrays' bytes (bodies share meshes)."""
    key = (kind,) + tuple((a.shape, a.dtype.str, a.tobytes()) for a in arrays)
    if key not in _MESH_CACHE:
        _MESH_CACHE[key] = fn(*arrays)
    return _MESH_CACHE[key]


def edges_of(F):
    return _cached("edges", (np.asarray(F),), _edges_of).copy()


def _edges_of(F):
    E = set()
    for f in F:
        for j in range(3):
            a, b = int(f[j]), int(f[(j + 1) % 3])
            E.add((min(a, b), max(a, b)))
    return np.array(sorted(E), np.int64).reshape(-1, 2)


# ---------------------------------------------------------- mass properties
def mass_properties(V, F):
    """Volume, centre of mass and inertia tensor (unit density) of a closed,
    outward-oriented triangle mesh (divergence theorem, per tetrahedron)."""
    a, b, c = V[F[:, 0]], V[F[:, 1]], V[F[:, 2]]
    vol6 = np.einsum("ij,ij->i", a, np.cross(b, c))
    vol = vol6.sum() / 6.0
    com = (vol6[:, None] * (a + b + c)).sum(0) / (24.0 * vol)
    # second moments of each tetrahedron (origin, a, b, c)
    C = np.zeros((3, 3))
    for i in range(3):
        for j in range(3):
            s = (a[:, i] * a[:, j] + b[:, i] * b[:, j] + c[:, i] * c[:, j]
                 + (a[:, i] + b[:, i] + c[:, i]) * (a[:, j] + b[:, j] + c[:, j]))
            C[i, j] = (vol6 * s).sum() / 120.0
    C -= vol * np.outer(com, com)
    inertia = np.trace(C) * np.eye(3) - C
    return vol, com, inertia


def open_mass_properties(V, F):
    """Mass, centre of mass and inertia tensor (unit density) of an open
    triangle mesh with fewer than 4 faces, as src/physics/mass.cpp:11-45
    gives them: lumped vertex masses from the mixed Voronoi areas (Meyer et
    al.; igl::massmatrix VORONOI: an obtuse corner takes half its triangle's
    area, the other two a quarter each), then point masses about the centre."""
    V = np.asarray(V, np.float64)
    m = np.zeros(len(V))
    for f in np.asarray(F):
        P = V[f]
        area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
        ang = []
        for c in range(3):
            u, w = P[(c + 1) % 3] - P[c], P[(c + 2) % 3] - P[c]
            ang.append(np.arctan2(np.linalg.norm(np.cross(u, w)), np.dot(u, w)))
        if max(ang) > np.pi / 2:
            ob = int(np.argmax(ang))
            for c in range(3):
                m[f[c]] += area / 2 if c == ob else area / 4
        else:
            for c in range(3):
                j, k = (c + 1) % 3, (c + 2) % 3
                ejc = np.dot(P[j] - P[c], P[j] - P[c])
                ekc = np.dot(P[k] - P[c], P[k] - P[c])
                m[f[c]] += (ejc / np.tan(ang[k]) + ekc / np.tan(ang[j])) / 8
    mass = m.sum()
    com = (m[:, None] * V).sum(0) / mass if mass > 0 else np.zeros(3)
    r = V - com
    inertia = np.zeros((3, 3))
    for mi, v in zip(m, r):
        inertia += mi * (np.dot(v, v) * np.eye(3) - np.outer(v, v))
    return mass, com, inertia


def rotvec_to_matrix(r):
    th = np.linalg.norm(r)
    if th == 0:
        return np.eye(3)
    k = r / th
    K = np.array([[0, -k[2], k[1]], [k[2], 0, -k[0]], [-k[1], k[0], 0]])
    return np.eye(3) + np.sin(th) * K + (1 - np.cos(th)) * K @ K


def matrix_to_rotvec(R):
    c = np.clip((np.trace(R) - 1) / 2, -1, 1)
    th = np.arccos(c)
    if th < 1e-12:
        return np.zeros(3)
    if np.pi - th < 1e-6:
        w, v = np.linalg.eigh((R + R.T) / 2)
        axis = v[:, np.argmax(w)]
        return axis * th
    axis = np.array([R[2, 1] - R[1, 2], R[0, 2] - R[2, 0], R[1, 0] - R[0, 1]]) / (2 * np.sin(th))
    return axis * th


def random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


# ------------------------------------------------------------------- scene
class Body:
    def __init__(self, V, F, R, p, v, w, density=1000.0, static=False, group=None, kinematic=False,
                 kinematic_max_time=-1.0, kinematic_poses=(), force=None):
        """V, F: mesh in its own frame; world pose = rotation R, translation p;
        v, w: world linear / angular velocity (rad/s). A kinematic body
        follows kinematic_poses (degrees of freedom as `pose`, one per step)
        or its velocity, until kinematic_max_time (< 0: never); force: its
        external force and torque (world), or None."""
        self.static = static
        self.group = group
        self.kinematic = bool(kinematic) and not static
        self.kinematic_max_time = float(kinematic_max_time)
        self.kinematic_poses = [np.asarray(x, np.float64).reshape(6) for x in kinematic_poses]
        self.force = None if force is None or static else np.asarray(force, np.float64).reshape(6)
        E = edges_of(F)
        if not static:
            props = mass_properties if len(F) >= 4 else open_mass_properties
            vol, com, I = _cached("mass" if len(F) >= 4 else "open_mass", (np.asarray(V, np.float64), np.asarray(F)),
                                  props)
            V = V - com
            p = p + R @ com
            evals, evecs = np.linalg.eigh(I)
            if np.linalg.det(evecs) < 0:
                evecs[:, 0] *= -1
            R0 = evecs
            self.mass = density * vol
            self.inertia = density * evals
        else:
            R0 = np.eye(3)
            self.mass = 0.0
            self.inertia = np.zeros(3)
        self.R0 = R0
        self.V = V @ R0  # body-local, principal frame
        self.F = F
        self.E = E
        Rw = R @ R0
        self.pose = np.concatenate([p, matrix_to_rotvec(Rw)])
        # angular velocity expressed in the body frame (rigid_body.cpp:161-163)
        self.velocity = np.concatenate([v, Rw.T @ w]) if not static else np.zeros(6)


def _face_edges(E, F):
    emap = {(int(a), int(c)): i for i, (a, c) in enumerate(E)}
    return np.array([[emap[(min(f[j], f[(j + 1) % 3]), max(f[j], f[(j + 1) % 3]))] for j in range(3)]
                     for f in F], np.int64).reshape(-1, 3)


def assemble(bodies):
    """Concatenate bodies into the arrays of include/rigc.h's rigc_scene."""
    V, E, F, FE = [], [], [], []
    bv, be, bf = [0], [0], [0]
    nv0 = ne0 = 0
    for b in bodies:
        fe = _cached("face_edges", (np.asarray(b.E), np.asarray(b.F)), _face_edges)
        V.append(b.V)
        E.append(b.E + nv0)
        F.append(b.F + nv0)
        FE.append(fe + ne0)
        bv.append(bv[-1] + len(b.V))
        be.append(be[-1] + len(b.E))
        bf.append(bf[-1] + len(b.F))
        nv0 += len(b.V)
        ne0 += len(b.E)
    groups = [b.group if b.group is not None else i for i, b in enumerate(bodies)]
    return types.SimpleNamespace(
        vertices=np.concatenate(V).astype(np.float64),
        edges=np.concatenate(E).astype(np.int32),
        faces=np.concatenate(F).astype(np.int32),
        face_edges=np.concatenate(FE).astype(np.int32),
        body_vertex_start=np.array(bv, np.int32), body_edge_start=np.array(be, np.int32),
        body_face_start=np.array(bf, np.int32), group_ids=np.array(groups, np.int32),
        is_static=np.array([b.static for b in bodies], np.uint8),
        poses=np.array([b.pose for b in bodies]), velocities=np.array([b.velocity for b in bodies]),
        mass=np.array([b.mass for b in bodies]), inertia=np.array([b.inertia for b in bodies]),
        bodies=bodies, **_kinematic_arrays(bodies),
    )


def _kinematic_arrays(bodies):
    """External forces and kinematic scripts of bodies, as the stepper takes
    them (none when no body has them)."""
    out = {}
    if any(getattr(b, "force", None) is not None for b in bodies):
        out["force"] = np.array([b.force if getattr(b, "force", None) is not None else np.zeros(6)
                                 for b in bodies])
    if any(getattr(b, "kinematic", False) for b in bodies):
        start, poses = [0], []
        for b in bodies:
            poses.extend(b.kinematic_poses if b.kinematic else [])
            start.append(len(poses))
        out["kinematic"] = np.array([int(b.kinematic) for b in bodies], np.int32)
        out["kinematic_max_time"] = np.array([b.kinematic_max_time for b in bodies])
        out["kinematic_pose_start"] = np.array(start, np.int32)
        out["kinematic_poses"] = np.array(poses).reshape(-1, 6)
    return out


def _ground():
    V, F = plane(20.0)
    return Body(V, F, np.eye(3), np.zeros(3), np.zeros(3), np.zeros(3), static=True)


def config1(seed=0, n=10):
    """10 subdivision-2 icospheres (r=0.5, rho=1000) with centres uniform in
    [0,3]^3 (rejecting overlaps with each other and the ground plane),
    v ~ N(0,1) m/s, w ~ N(0,2) rad/s, plus a static 20x20 plane at y=0."""
    rng = np.random.default_rng(seed)
    Vs, Fs = icosphere(2, 0.5)
    centres = []
    while len(centres) < n:
        c = rng.uniform(0, 3, size=3)
        if c[1] < 0.5 + 1e-3:
            continue
        if all(np.linalg.norm(c - d) > 1.0 + 1e-3 for d in centres):
            centres.append(c)
    bodies = [Body(Vs, Fs, np.eye(3), c, rng.normal(0, 1, 3), rng.normal(0, 2, 3)) for c in centres]
    bodies.append(_ground())
    sc = assemble(bodies)
    sc.dt, sc.dhat, sc.mu, sc.gravity, sc.steps = 0.01, 1e-3, 0.0, np.array([0, -9.81, 0]), 100
    sc.name = f"config1: {n} ico2 + plane"
    return sc


def _mixed_mesh(kind):
    if kind == 0:
        return icosphere(2, 0.5)
    if kind == 1:
        return torus(0.5, 0.15)
    return cube(1.0)


def lattice_bodies(rng, n, box, cell, lift, vel_sigma, omega_sigma, kinds=(0, 1, 2), meshes=None):
    """n bodies on a jittered lattice of `cell`-sized cells filling `box`
    (x, y, z extents), each with a uniform random orientation resampled until
    its AABB fits its cell (so bodies start overlap-free), lifted by `lift`."""
    nx, ny, nz = (int(box[0] // cell), int(box[1] // cell), int(box[2] // cell))
    assert nx * ny * nz >= n, "lattice too small"
    slots = rng.permutation(nx * ny * nz)[:n]
    meshes = meshes or {k: _mixed_mesh(k) for k in kinds}
    bodies = []
    margin = 1e-3
    for s in slots:
        i, j, k = s // (ny * nz), (s // nz) % ny, s % nz
        kind = kinds[int(rng.integers(0, len(kinds)))]
        V, F = meshes[kind]
        for _ in range(1000):
            R = random_rotation(rng)
            W = V @ R.T
            lo, hi = W.min(0), W.max(0)
            if np.all(hi - lo <= cell - 2 * margin):
                break
        centre_cell = np.array([(i + 0.5) * cell - box[0] / 2, (j + 0.5) * cell + lift, (k + 0.5) * cell - box[2] / 2])
        slack = (cell - 2 * margin - (hi - lo)) / 2
        off = rng.uniform(-1, 1, 3) * np.maximum(slack, 0)
        p = centre_cell + off - (lo + hi) / 2
        bodies.append(Body(V, F, R, p, rng.normal(0, vel_sigma, 3), rng.normal(0, omega_sigma, 3)))
    return bodies


def config2(seed=0, n=1000):
    """1000 bodies drawn uniformly from {ico2 r=0.5, 24x12 torus R=0.5
    r=0.15, 12-triangle unit cube}, uniform random orientations, placed
    overlap-free on a jittered 10x10x10 lattice of 1.5 m cells filling a
    15^3 box above the ground plane; v ~ N(0,0.5), w ~ N(0,1);
    dt=0.01, dhat=1e-3, mu=0.2, 50 steps."""
    rng = np.random.default_rng(seed)
    side = int(np.ceil(n ** (1 / 3) - 1e-9))
    bodies = lattice_bodies(rng, n, (side * 1.5,) * 3, 1.5, 0.01, 0.5, 1.0)
    bodies.append(_ground())
    sc = assemble(bodies)
    sc.dt, sc.dhat, sc.mu, sc.gravity, sc.steps = 0.01, 1e-3, 0.2, np.array([0, -9.81, 0]), 50
    sc.name = f"config2: {n} mixed bodies + plane"
    return sc


def config3_scene(i, base_seed=1000):
    """One of the 256 independent batching scenes: body count log-uniform
    in [8, 512], mix and distributions as config2, dt in {0.005, 0.01}."""
    rng = np.random.default_rng(base_seed + i)
    n = int(np.exp(rng.uniform(np.log(8), np.log(512))))
    side = int(np.ceil(n ** (1 / 3) - 1e-9))
    bodies = lattice_bodies(rng, n, (side * 1.5,) * 3, 1.5, 0.01, 0.5, 1.0)
    bodies.append(_ground())
    sc = assemble(bodies)
    sc.dt = float(rng.choice([0.005, 0.01]))
    sc.dhat, sc.mu, sc.gravity, sc.steps = 1e-3, 0.2, np.array([0, -9.81, 0]), 10
    sc.name = f"config3[{i}]: {n} bodies"
    return sc


def config4(seed=0):
    """64 subdivision-1 icospheres (r=0.3) on a 4^3 lattice (spacing 0.8)
    with jitter U(-0.05,0.05), dropped on the plane; 20 steps, dt=0.01, mu=0.5."""
    rng = np.random.default_rng(seed)
    V, F = icosphere(1, 0.3)
    bodies = []
    for i in range(4):
        for j in range(4):
            for k in range(4):
                c = np.array([i * 0.8, 0.4 + j * 0.8, k * 0.8]) + rng.uniform(-0.05, 0.05, 3)
                bodies.append(Body(V, F, np.eye(3), c, np.zeros(3), np.zeros(3)))
    bodies.append(_ground())
    sc = assemble(bodies)
    sc.dt, sc.dhat, sc.mu, sc.gravity, sc.steps = 0.01, 1e-3, 0.5, np.array([0, -9.81, 0]), 20
    sc.targets = np.array([b.pose[:3] for b in bodies[:-1]]) + rng.normal(0, 0.1, (64, 3))
    sc.name = "config4: 64 ico1 + plane"
    return sc


def _wall(dims, centre, group):
    V, F = cube(1.0)
    return Body(V * np.asarray(dims, float), F, np.eye(3), np.asarray(centre, float), np.zeros(3), np.zeros(3),
                static=True, group=group)


def config5(seed=0, n_spheres=9000, n_tori=1000, box=10.0):
    """Dense pile: 9,000 subdivision-3 icospheres (642 vertices, r 0.25) and
    1,000 24x12 tori (R 0.5, r 0.15) in uniform random orientations, stacked
    in a box x box column inside a box x box box with walls, above the
    plane, falling at v = (0, -2, 0) + N(0, 0.1) m/s per component; dt 0.005,
    dhat 1e-4, mu 0.5, 20 steps. Spheres sit in 0.52 m cells, tori in
    1.35 m cells (a torus's extent is at most 1.3 m in any orientation),
    jittered within them; layers of tori and layers of spheres alternate up
    the column in proportion to their numbers (with the default sizes 21
    torus layers and 25 sphere layers, about 42 m of the 60 m column). The
    plane and the four walls (0.5 m thick, 62 m tall) are static and share
    one group. 12.10M triangles with the defaults."""
    rng = np.random.default_rng(seed)
    Vs, Fs = icosphere(3, 0.25)
    Vt, Ft = torus(0.5, 0.15)
    half = box / 2.0
    layers = []
    for kind, n, cell in (("s", n_spheres, 0.52), ("t", n_tori, 1.35)):
        per_side = int(box // cell)
        per_layer = per_side * per_side
        n_layers = -(-n // per_layer)
        for k in range(n_layers):
            count = min(per_layer, n - k * per_layer)
            layers.append(((k + 0.5) / n_layers, kind, cell, per_side, count))
    layers.sort(key=lambda t: (t[0], t[1]))
    bodies = []
    y = 0.05
    for _, kind, cell, per_side, count in layers:
        V, F = (Vs, Fs) if kind == "s" else (Vt, Ft)
        slots = rng.permutation(per_side * per_side)[:count]
        off = (box - per_side * cell) / 2.0
        for slot in sorted(slots):
            i, j = divmod(int(slot), per_side)
            R = random_rotation(rng)
            W = V @ R.T
            lo, hi = W.min(0), W.max(0)
            slack = np.maximum(cell - 0.01 - (hi - lo), 0.0) / 2.0
            centre = np.array([-half + off + (i + 0.5) * cell, y + cell / 2.0, -half + off + (j + 0.5) * cell])
            p = centre + rng.uniform(-1, 1, 3) * slack - (lo + hi) / 2.0
            v = np.array([0.0, -2.0, 0.0]) + rng.normal(0.0, 0.1, 3)
            bodies.append(Body(V, F, R, p, v, np.zeros(3)))
        y += cell + 0.01
    statics = len(bodies)
    ground = _ground()
    ground.group = statics
    t = 0.5
    h = 62.0
    walls = [_wall((t, h, box + 2 * t), (s * (half + t / 2), h / 2, 0.0), statics) for s in (-1, 1)]
    walls += [_wall((box + 2 * t, h, t), (0.0, h / 2, s * (half + t / 2)), statics) for s in (-1, 1)]
    sc = assemble(bodies + [ground] + walls)
    sc.dt, sc.dhat, sc.mu, sc.gravity, sc.steps = 0.005, 1e-4, 0.5, np.array([0, -9.81, 0]), 20
    sc.column_height = y
    sc.name = f"config5: {n_spheres} ico3 + {n_tori} tori in a walled {box:g} x {box:g} box"
    return sc
