"""Seeded synthetic scene generators for the rigid-body step workloads.

Every random draw comes from a counter-based SplitMix64 stream keyed by
(master seed, object id, draw index), so a scene is a pure function of its
parameters and does not depend on library-specific random distributions.
The same arrays feed the GPU library and the reference CPU pipeline.

Geometry type codes follow PxGeometryType (physx/include/geometry/PxGeometry.h):
0 sphere, 1 plane, 2 capsule, 3 box, 5 convex mesh (dims[0] = the scene's
mesh index; the meshes are point clouds in hull_points / hull_counts).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SPHERE, PLANE, CAPSULE, BOX, CONVEX = 0, 1, 2, 3, 5
JOINT_FIXED, JOINT_SPHERICAL, JOINT_REVOLUTE, JOINT_PRISMATIC, JOINT_DISTANCE = 0, 1, 2, 3, 4

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _splitmix(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = x.astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, ids: np.ndarray, draw: int, lo=0.0, hi=1.0) -> np.ndarray:
    """U[lo,hi) float64 for each object id, draw index `draw` of its stream."""
    with np.errstate(over="ignore"):
        ctr = (np.asarray(ids, np.uint64) * np.uint64(64) + np.uint64(draw + 1)) * _GOLDEN
        x = _splitmix(np.uint64(seed) ^ ctr)
    u = (x >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


@dataclass
class Scene:
    """Flat body/plane arrays of one or more independent scenes (worlds)."""
    shape_type: np.ndarray          # (n,) int32: SPHERE, CAPSULE, BOX or CONVEX
    dims: np.ndarray                # (n,4) f32: radius; radius, half height; half extents; mesh index
    pose: np.ndarray                # (n,7) f32: p xyz, q xyzw
    lin_vel: np.ndarray             # (n,3) f32
    ang_vel: np.ndarray             # (n,3) f32
    mass: np.ndarray                # (n,) f32
    inertia: np.ndarray             # (n,3) f32, mass-space diagonal
    scene_id: np.ndarray            # (n,) int32 world index of each body
    planes: np.ndarray              # (m,4) f32: n xyz, d with n.x + d = 0
    plane_scene: np.ndarray         # (m,) int32 world index of each plane
    n_scenes: int = 1
    name: str = ""
    params: dict = field(default_factory=dict)
    hull_points: np.ndarray | None = None   # (P,3) f32: point clouds of the convex meshes, concatenated
    hull_counts: np.ndarray | None = None   # (H,) int32: points per mesh
    # joints: type (k,) int32 (JOINT_*), body0 / body1 (k,) int32 (-1: the
    # world frame), frame0 / frame1 (k,7) f32 local frames (p xyz, q xyzw)
    joints: dict | None = None
    # aggregates: list of (member body indices, self_collisions)
    aggregates: list | None = None
    # static triangle meshes, at most one per world: list of dicts with
    # "verts" (nv,3) f32, "tris" (nt,3) int32 (counter-clockwise about the
    # outward normal), "pose" (7,) f32 (p xyz, q xyzw), "world" int
    meshes: list | None = None

    @property
    def n(self) -> int:
        return int(self.shape_type.shape[0])

    def subset(self, world: int) -> "Scene":
        """The bodies and planes of one world, renumbered as world 0."""
        return self.subset_range(world, world + 1)

    def subset_range(self, w0: int, w1: int) -> "Scene":
        """The bodies and planes of worlds [w0, w1), renumbered from 0 (in order)."""
        b = (self.scene_id >= w0) & (self.scene_id < w1)
        p = (self.plane_scene >= w0) & (self.plane_scene < w1)
        name = f"{self.name}[{w0}]" if w1 == w0 + 1 else f"{self.name}[{w0}:{w1}]"
        joints = None
        if self.joints is not None:
            # joints of bodies in the range (a joint to the world belongs to
            # its body's world), body indices renumbered
            new = np.full(self.n, -1, np.int64)
            new[b] = np.arange(int(b.sum()))
            j0, j1 = self.joints["body0"], self.joints["body1"]
            keep = np.where(j0 >= 0, b[np.maximum(j0, 0)], True) & np.where(j1 >= 0, b[np.maximum(j1, 0)], True)
            keep &= (j0 >= 0) | (j1 >= 0)
            keep &= np.where(j0 >= 0, b[np.maximum(j0, 0)], b[np.maximum(j1, 0)])
            joints = {"type": self.joints["type"][keep].copy(),
                      "body0": np.where(j0[keep] >= 0, new[np.maximum(j0[keep], 0)], -1).astype(np.int32),
                      "body1": np.where(j1[keep] >= 0, new[np.maximum(j1[keep], 0)], -1).astype(np.int32),
                      "frame0": self.joints["frame0"][keep].copy(), "frame1": self.joints["frame1"][keep].copy()}
            if self.joints.get("params") is not None:
                joints["params"] = self.joints["params"][keep].copy()
        meshes = None
        if self.meshes:
            meshes = [dict(m, world=int(m["world"]) - w0) for m in self.meshes if w0 <= int(m["world"]) < w1]
        return Scene(self.shape_type[b].copy(), self.dims[b].copy(), self.pose[b].copy(),
                     self.lin_vel[b].copy(), self.ang_vel[b].copy(), self.mass[b].copy(),
                     self.inertia[b].copy(), (self.scene_id[b] - w0).astype(np.int32),
                     self.planes[p].copy(), (self.plane_scene[p] - w0).astype(np.int32), w1 - w0,
                     name, dict(self.params), self.hull_points, self.hull_counts, joints, None, meshes)


def mass_props(shape_type: np.ndarray, dims: np.ndarray, density: float):
    """Mass and mass-space inertia of solid spheres, capsules and boxes (closed
    forms). A capsule (radius r, half height h along x) is a cylinder of length
    2h and two hemispherical caps; each cap's inertia about the capsule centre
    perpendicular to the axis is m_cap (2/5 r^2 + h^2 + 3 h r / 4) (its own
    centroid sits 3r/8 from the flat face)."""
    dims = dims.astype(np.float64)
    n = shape_type.shape[0]
    mass = np.zeros(n)
    inertia = np.zeros((n, 3))
    s = shape_type == SPHERE
    r = dims[s, 0]
    mass[s] = density * (4.0 / 3.0) * math.pi * r ** 3
    inertia[s] = (0.4 * mass[s] * r * r)[:, None]
    c = shape_type == CAPSULE
    r, hh = dims[c, 0], dims[c, 1]
    m_cyl = density * math.pi * r * r * 2.0 * hh
    m_sph = density * (4.0 / 3.0) * math.pi * r ** 3
    mass[c] = m_cyl + m_sph
    inertia[c, 0] = 0.5 * m_cyl * r * r + 0.4 * m_sph * r * r
    perp = m_cyl * (0.25 * r * r + (2.0 * hh) ** 2 / 12.0) + m_sph * (0.4 * r * r + hh * hh + 0.75 * hh * r)
    inertia[c, 1] = perp
    inertia[c, 2] = perp
    b = shape_type == BOX
    h = dims[b, :3]
    mass[b] = density * 8.0 * h[:, 0] * h[:, 1] * h[:, 2]
    m3 = mass[b] / 3.0
    inertia[b, 0] = m3 * (h[:, 1] ** 2 + h[:, 2] ** 2)
    inertia[b, 1] = m3 * (h[:, 0] ** 2 + h[:, 2] ** 2)
    inertia[b, 2] = m3 * (h[:, 0] ** 2 + h[:, 1] ** 2)
    return mass.astype(np.float32), inertia.astype(np.float32)


def _yaw_quat(yaw: np.ndarray) -> np.ndarray:
    q = np.zeros((yaw.shape[0], 4))
    q[:, 1] = np.sin(0.5 * yaw)
    q[:, 3] = np.cos(0.5 * yaw)
    return q


def _finish(shape_type, dims, pos, quat, scene_id, planes, plane_scene, n_scenes, name, params, density=1000.0):
    n = shape_type.shape[0]
    mass, inertia = mass_props(shape_type, dims, density)
    pose = np.concatenate([pos, quat], 1).astype(np.float32)
    return Scene(shape_type.astype(np.int32), dims.astype(np.float32), pose,
                 np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), mass, inertia,
                 scene_id.astype(np.int32), np.asarray(planes, np.float32).reshape(-1, 4),
                 np.asarray(plane_scene, np.int32), n_scenes, name, params)


def box_grid_drop(nx=16, ny=16, nz=16, pitch=1.2, jitter=0.05, hmin=0.25, hmax=0.5, seed=1) -> Scene:
    """4,096 boxes (16^3) with half extents U[0.25,0.5], random yaw, dropped
    from a jittered 1.2 m grid onto the plane y=0."""
    n = nx * ny * nz
    ids = np.arange(n)
    ix, iy, iz = ids % nx, (ids // nx) % ny, ids // (nx * ny)
    dims = np.zeros((n, 4))
    for k in range(3):
        dims[:, k] = uniform(seed, ids, k, hmin, hmax)
    pos = np.stack([(ix - (nx - 1) / 2) * pitch, hmax + 0.1 + iy * pitch, (iz - (nz - 1) / 2) * pitch], 1)
    for k in range(3):
        pos[:, k] += uniform(seed, ids, 3 + k, -jitter, jitter)
    quat = _yaw_quat(uniform(seed, ids, 6, 0.0, 2 * math.pi))
    st = np.full(n, BOX)
    return _finish(st, dims, pos, quat, np.zeros(n), [[0, 1, 0, 0]], [0], 1, "box_grid_drop",
                   dict(nx=nx, ny=ny, nz=nz, pitch=pitch, jitter=jitter, seed=seed))


def sphere_box_bin(n_bodies=262144, bin_half=20.0, rmin=0.1, rmax=0.3, pitch_xz=0.9, pitch_y=0.65,
                   jitter=0.02, seed=2) -> Scene:
    """Spheres (r U[0.1,0.3]) and boxes (half extents U[0.1,0.3], random yaw)
    alternating on a jittered grid, poured into a walled bin: floor y=0 and
    four walls at |x|,|z| = bin_half."""
    per_row = int((2 * bin_half - 0.5) // pitch_xz)
    n = n_bodies
    ids = np.arange(n)
    ix, iz, iy = ids % per_row, (ids // per_row) % per_row, ids // (per_row * per_row)
    st = np.where((ix + iy + iz) % 2 == 0, SPHERE, BOX)
    dims = np.zeros((n, 4))
    r = uniform(seed, ids, 0, rmin, rmax)
    dims[:, 0] = r
    for k in range(3):
        hk = uniform(seed, ids, 1 + k, rmin, rmax)
        dims[:, k] = np.where(st == BOX, hk, dims[:, k])
    pos = np.stack([(ix - (per_row - 1) / 2) * pitch_xz, rmax + 0.05 + iy * pitch_y,
                    (iz - (per_row - 1) / 2) * pitch_xz], 1)
    for k in range(3):
        pos[:, k] += uniform(seed, ids, 4 + k, -jitter, jitter)
    quat = _yaw_quat(np.where(st == BOX, uniform(seed, ids, 7, 0.0, 2 * math.pi), 0.0))
    planes = [[0, 1, 0, 0], [1, 0, 0, bin_half], [-1, 0, 0, bin_half], [0, 0, 1, bin_half], [0, 0, -1, bin_half]]
    return _finish(st, dims, pos, quat, np.zeros(n), planes, [0] * 5, 1, "sphere_box_bin",
                   dict(n_bodies=n_bodies, bin_half=bin_half, pitch_xz=pitch_xz, pitch_y=pitch_y, seed=seed))


def batched_piles(n_scenes=8192, nmin=8, nmax=512, seed=4) -> Scene:
    """Independent worlds with body counts log-uniform in [nmin, nmax]; even
    worlds are box stacks (columns of boxes), odd worlds sphere/box piles on a
    jittered grid. Each world has its own floor plane y=0."""
    sid = np.arange(n_scenes)
    counts = np.floor(np.exp(uniform(seed, sid, 0, math.log(nmin), math.log(nmax + 1)))).astype(np.int64)
    counts = np.clip(counts, nmin, nmax)
    total = int(counts.sum())
    world = np.repeat(sid, counts)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    local = np.arange(total) - np.repeat(first, counts)
    gid = np.arange(total) + (seed << 32)
    stack = (world % 2) == 0
    # box stacks: columns of 8 boxes, columns on a 1.2 m grid
    side = np.ceil(np.sqrt(np.ceil(counts / 8.0))).astype(np.int64)[world]
    col = local // 8
    lvl = local % 8
    h = np.stack([uniform(seed, gid, 1 + k, 0.2, 0.3) for k in range(3)], 1)
    st = np.where(stack, BOX, np.where(local % 2 == 0, SPHERE, BOX))
    dims = np.zeros((total, 4))
    dims[:, :3] = h
    r = uniform(seed, gid, 4, 0.1, 0.3)
    dims[:, 0] = np.where(st == SPHERE, r, dims[:, 0])
    pos = np.zeros((total, 3))
    # stacks: boxes centred on the column, each level resting 1 cm above the one below
    pos[:, 0] = np.where(stack, (col % side - (side - 1) / 2.0) * 1.2, 0.0)
    pos[:, 2] = np.where(stack, (col // side - (side - 1) / 2.0) * 1.2, 0.0)
    pos[:, 1] = np.where(stack, 0.31 + lvl * 0.62, 0.0)
    # piles: jittered grid at 0.9 m
    pside = np.ceil(np.cbrt(counts)).astype(np.int64)[world]
    px, pz, py = local % pside, (local // pside) % pside, local // (pside * pside)
    for k, (a, c) in enumerate(((px, 0), (pz, 2))):
        pos[:, c] = np.where(stack, pos[:, c], (a - (pside - 1) / 2.0) * 0.9 + uniform(seed, gid, 5 + k, -0.02, 0.02))
    pos[:, 1] = np.where(stack, pos[:, 1], 0.35 + py * 0.9)
    quat = _yaw_quat(np.where(st == BOX, uniform(seed, gid, 8, 0.0, 2 * math.pi), 0.0) * np.where(stack, 0.05, 1.0))
    planes = np.tile(np.array([[0, 1, 0, 0]], np.float32), (n_scenes, 1))
    return _finish(st, dims, pos, quat, world, planes, sid, n_scenes, "batched_piles",
                   dict(n_scenes=n_scenes, nmin=nmin, nmax=nmax, seed=seed))


def _rand_quat(seed: int, ids: np.ndarray, draw: int) -> np.ndarray:
    """Uniform random rotations (Shoemake's subgroup algorithm), xyzw."""
    u1, u2, u3 = (uniform(seed, ids, draw + k) for k in range(3))
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    t1, t2 = 2 * math.pi * u2, 2 * math.pi * u3
    return np.stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)], 1)


def convex_layers(n_bodies=65536, per_side=64, pitch=1.3, n_points=24, amin=0.2, amax=0.6, density=1000.0,
                  seed=5) -> Scene:
    """Convex hulls of n_points points on seeded ellipsoids (semi-axes
    U[amin, amax] m, directions uniform on the sphere), dropped in layers of
    per_side x per_side on a 1.3 m grid onto the plane y=0 with random
    orientations. Each hull is expressed in its principal frame (points
    shifted by the hull's centre of mass and rotated onto its principal axes),
    so the body frame is the shape frame and the inertia is diagonal."""
    from . import cook_convex
    ids = np.arange(n_bodies)
    ax = np.stack([uniform(seed, ids, d, amin, amax) for d in range(3)], 1)
    gid = ((ids[:, None] << 8) + np.arange(n_points)[None, :]).ravel()
    z = uniform(seed + 1, gid, 0, -1.0, 1.0).reshape(n_bodies, n_points)
    phi = uniform(seed + 1, gid, 1, 0.0, 2 * math.pi).reshape(n_bodies, n_points)
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    pts = (np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], 2) * ax[:, None, :]).astype(np.float32)
    com = np.zeros((n_bodies, 3))
    tensor = np.zeros((n_bodies, 3, 3))
    vol = np.zeros(n_bodies)
    for i in range(n_bodies):
        h = cook_convex(pts[i])
        if h is None:
            raise RuntimeError(f"convex_layers: degenerate point set for hull {i}")
        com[i], tensor[i], vol[i] = h["com"], h["inertia"], h["volume"]
    w, vecs = np.linalg.eigh(tensor)
    flip = np.linalg.det(vecs) < 0
    vecs[flip, :, 2] *= -1.0
    local = np.einsum("nkc,ncd->nkd", pts.astype(np.float64) - com[:, None, :], vecs).astype(np.float32)
    mass = (density * vol).astype(np.float32)
    inertia = (density * w).astype(np.float32)
    per_layer = per_side * per_side
    lx, lz, ly = ids % per_side, (ids // per_side) % per_side, ids // per_layer
    pos = np.stack([(lx - (per_side - 1) / 2) * pitch, 0.7 + ly * pitch, (lz - (per_side - 1) / 2) * pitch], 1)
    quat = _rand_quat(seed + 2, ids, 0)
    dims = np.zeros((n_bodies, 4), np.float32)
    dims[:, 0] = ids
    pose = np.concatenate([pos, quat], 1).astype(np.float32)
    return Scene(np.full(n_bodies, CONVEX, np.int32), dims, pose, np.zeros((n_bodies, 3), np.float32),
                 np.zeros((n_bodies, 3), np.float32), mass, inertia, np.zeros(n_bodies, np.int32),
                 np.array([[0, 1, 0, 0]], np.float32), np.zeros(1, np.int32), 1, "convex_layers",
                 dict(n_bodies=n_bodies, per_side=per_side, pitch=pitch, n_points=n_points, seed=seed),
                 local.reshape(-1, 3), np.full(n_bodies, n_points, np.int32))


def mixed_pile(n_bodies=512, per_side=8, pitch=1.0, jitter=0.1, n_points=16, seed=9) -> Scene:
    """Spheres, capsules, boxes and convex hulls in equal shares, dropped in
    layers of per_side x per_side onto the plane y=0 (grid cell (x, z, layer)
    holds kind (x + 2 z + layer) mod 4, so neighbours along x, along z and
    above differ in kind) with seeded
    horizontal jitter U[-jitter, jitter] and random orientations: sphere
    radius U[0.1,0.3], capsule radius U[0.08,0.15] and half height
    U[0.05,0.25], box half extents U[0.1,0.3], hulls of n_points points on
    ellipsoids with semi-axes U[0.15,0.35] in their principal frames."""
    from . import cook_convex
    ids = np.arange(n_bodies)
    per_layer = per_side * per_side
    lx, lz, ly = ids % per_side, (ids // per_side) % per_side, ids // per_layer
    kind = (lx + 2 * lz + ly) % 4
    shape_type = np.array([SPHERE, CAPSULE, BOX, CONVEX], np.int32)[kind]
    dims = np.zeros((n_bodies, 4), np.float32)
    sp, ca, bx, cv = (kind == k for k in range(4))
    dims[sp, 0] = uniform(seed, ids[sp], 0, 0.1, 0.3)
    dims[ca, 0] = uniform(seed, ids[ca], 0, 0.08, 0.15)
    dims[ca, 1] = uniform(seed, ids[ca], 1, 0.05, 0.25)
    for d in range(3):
        dims[bx, d] = uniform(seed, ids[bx], d, 0.1, 0.3)
    cvx_ids = ids[cv]
    nc = cvx_ids.shape[0]
    ax = np.stack([uniform(seed, cvx_ids, 3 + d, 0.15, 0.35) for d in range(3)], 1)
    gid = ((cvx_ids[:, None] << 8) + np.arange(n_points)[None, :]).ravel()
    z = uniform(seed + 1, gid, 0, -1.0, 1.0).reshape(nc, n_points)
    phi = uniform(seed + 1, gid, 1, 0.0, 2 * math.pi).reshape(nc, n_points)
    rxy = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    pts = (np.stack([rxy * np.cos(phi), rxy * np.sin(phi), z], 2) * ax[:, None, :]).astype(np.float32)
    mass, inertia = mass_props(shape_type, dims, 1000.0)
    local = np.zeros_like(pts)
    for k in range(nc):
        h = cook_convex(pts[k])
        if h is None:
            raise RuntimeError(f"mixed_pile: degenerate point set for hull {k}")
        w, vecs = np.linalg.eigh(h["inertia"])
        if np.linalg.det(vecs) < 0:
            vecs[:, 2] *= -1.0
        local[k] = ((pts[k].astype(np.float64) - h["com"]) @ vecs).astype(np.float32)
        mass[cvx_ids[k]] = 1000.0 * h["volume"]
        inertia[cvx_ids[k]] = 1000.0 * w
    dims[cv, 0] = np.arange(nc)
    jx = uniform(seed + 3, ids, 0, -jitter, jitter)
    jz = uniform(seed + 3, ids, 1, -jitter, jitter)
    pos = np.stack([(lx - (per_side - 1) / 2) * pitch + jx, 0.45 + ly * pitch,
                    (lz - (per_side - 1) / 2) * pitch + jz], 1)
    quat = _rand_quat(seed + 2, ids, 0)
    pose = np.concatenate([pos, quat], 1).astype(np.float32)
    return Scene(shape_type, dims, pose, np.zeros((n_bodies, 3), np.float32), np.zeros((n_bodies, 3), np.float32),
                 mass.astype(np.float32), inertia.astype(np.float32), np.zeros(n_bodies, np.int32),
                 np.array([[0, 1, 0, 0]], np.float32), np.zeros(1, np.int32), 1, "mixed_pile",
                 dict(n_bodies=n_bodies, per_side=per_side, pitch=pitch, jitter=jitter, n_points=n_points, seed=seed),
                 local.reshape(-1, 3), np.full(nc, n_points, np.int32))


def _quat_axis_angle(axis, angle):
    axis = np.asarray(axis, np.float64)
    axis = axis / np.linalg.norm(axis, axis=-1, keepdims=True)
    h = 0.5 * np.asarray(angle, np.float64)[..., None]
    return np.concatenate([axis * np.sin(h), np.cos(h)], -1)


def joint_singles(n_groups=64, seed=5) -> Scene:
    """Independent jointed groups 3 m apart (no contacts): even groups
    one body joined to the world, odd groups two bodies joined to each other.
    Kinds cycle fixed, spherical, revolute, prismatic; bodies are boxes,
    spheres and capsules with random orientation and velocity, and the second
    frame is placed off the first by up to 5 cm / 0.1 rad so each joint starts
    with an error to correct. No gravity-free assumption: the step runs with
    gravity. Each body is in exactly one joint, so one step does not depend on
    the solver's constraint order."""
    types, dims, pos, quat, lv, av = [], [], [], [], [], []
    jt, jb0, jb1, jf0, jf1 = [], [], [], [], []
    for g in range(n_groups):
        ids = np.array([g])
        u = lambda d, lo=0.0, hi=1.0: float(uniform(seed, ids, d, lo, hi)[0])  # noqa: E731
        kind = g // 2 % 4
        nb = 1 if g % 2 == 0 else 2
        # near the origin: a joint's error over the step dt amplifies the
        # rounding of positions, which grows with their magnitude
        base = np.array([3.0 * (g % 8 - 3.5), 1.0, 3.0 * (g // 8 - 3.5)])
        first = len(types)
        for k in range(nb):
            st = [BOX, SPHERE, CAPSULE][(g + k) % 3]
            types.append(st)
            if st == BOX:
                dims.append([u(10 + k, 0.1, 0.4), u(12 + k, 0.1, 0.4), u(14 + k, 0.1, 0.4), 0.0])
            elif st == SPHERE:
                dims.append([u(10 + k, 0.15, 0.4), 0.0, 0.0, 0.0])
            else:
                dims.append([u(10 + k, 0.1, 0.25), u(12 + k, 0.15, 0.4), 0.0, 0.0])
            pos.append(base + np.array([1.2 * k, 0.0, 0.0]))
            quat.append(_quat_axis_angle([u(20 + k, -1, 1), u(22 + k, -1, 1), u(24 + k, -1, 1)], u(26 + k, 0, 3.0)))
            lv.append([u(30 + k, -2, 2), u(32 + k, -2, 2), u(34 + k, -2, 2)])
            av.append([u(36 + k, -3, 3), u(38 + k, -3, 3), u(40 + k, -3, 3)])
        # joint frame in the world: an anchor between / beside the bodies
        anchor = base + np.array([0.6 if nb == 2 else -0.5, 0.1, 0.0])
        jq = _quat_axis_angle([u(50, -1, 1), u(51, -1, 1), u(52, -1, 1)], u(53, 0, 3.0))
        err_p = np.array([u(54, -0.05, 0.05), u(55, -0.05, 0.05), u(56, -0.05, 0.05)])
        err_q = _quat_axis_angle([u(57, -1, 1), u(58, -1, 1), u(59, -1, 1)], u(60, 0, 0.1))

        def local(body, p_w, q_w):
            """the world frame (p_w, q_w) in body `body`'s frame"""
            qb = np.asarray(quat[body], np.float64)
            inv = np.array([-qb[0], -qb[1], -qb[2], qb[3]])
            return np.concatenate([_qrot(inv, p_w - pos[body]), _qmul(inv, q_w)])

        world_f = np.concatenate([anchor, jq])
        q_b = _qmul(err_q, jq)
        if nb == 1:
            if g % 4 == 0:      # the world as body 0
                jb0.append(-1); jb1.append(first)
                jf0.append(world_f); jf1.append(local(first, anchor + err_p, q_b))
            else:               # the world as body 1
                jb0.append(first); jb1.append(-1)
                jf0.append(local(first, anchor + err_p, q_b)); jf1.append(world_f)
        else:
            jb0.append(first); jb1.append(first + 1)
            jf0.append(local(first, anchor, jq)); jf1.append(local(first + 1, anchor + err_p, q_b))
        jt.append(kind)
    st = np.array(types, np.int32)
    sc = _finish(st, np.array(dims), np.array(pos), np.array(quat), np.zeros(len(types), np.int32),
                 np.zeros((0, 4)), np.zeros(0, np.int32), 1, "joint_singles", {"n_groups": n_groups, "seed": seed})
    sc.lin_vel = np.array(lv, np.float32)
    sc.ang_vel = np.array(av, np.float32)
    sc.joints = {"type": np.array(jt, np.int32), "body0": np.array(jb0, np.int32), "body1": np.array(jb1, np.int32),
                 "frame0": np.array(jf0, np.float32), "frame1": np.array(jf1, np.float32)}
    return sc


def limited_joints(n_groups=48, seed=7) -> Scene:
    """joint_singles' independent groups with revolute (even joints) and
    prismatic (odd) joints carrying hard limits and, on revolute joints,
    velocity drives: revolute limits [-U(0.05,0.5), U(0.05,0.5)] rad, every
    third of them driven at U(-3,3) rad/s with a force limit of U(0.2,3) or
    unbounded (every fourth drive freespin, gear ratio 1.5 between two
    bodies), prismatic limits [-U(0.01,0.1), U(0.01,0.1)] m, the limits of
    the joints to the world as body 1 with restitution 0.5 (bounce threshold
    0.1), every eighth joint with its limit locked (lower = upper = 0), every
    fifth otherwise soft (stiffness 20 N m/rad, damping 0.5 revolute; 200 N/m,
    5 prismatic). The random starting
    velocities drive the joints into their limits within a few steps."""
    sc = joint_singles(n_groups, seed)
    k = len(sc.joints["type"])
    ids = np.arange(k)
    kind = np.where(ids % 2 == 0, JOINT_REVOLUTE, JOINT_PRISMATIC).astype(np.int32)
    params = np.zeros((k, 10), np.float32)
    params[:, 7] = 1.0
    rev = kind == JOINT_REVOLUTE
    lo = np.where(rev, -uniform(seed, ids, 70, 0.05, 0.5), -uniform(seed, ids, 70, 0.01, 0.1))
    hi = np.where(rev, uniform(seed, ids, 71, 0.05, 0.5), uniform(seed, ids, 71, 0.01, 0.1))
    locked = ids % 8 == 6
    lo[locked] = 0.0
    hi[locked] = 0.0
    flags = np.ones(k, np.int32)
    drive = rev & (ids % 3 == 0) & ~locked
    flags[drive] |= 2
    flags[drive & (ids % 4 == 0)] |= 4
    params[:, 0] = flags
    params[:, 1] = lo
    params[:, 2] = hi
    # restitution on joints to the world as body 1 (the reference's 4-wide
    # joint preparation projects body 1's angular velocity differently from
    # its one-joint preparation when it computes a bounce)
    bouncy = (ids % 4 == 2) & (sc.joints["body1"] < 0)
    params[bouncy, 3] = 0.5
    params[bouncy, 4] = 0.1
    params[:, 5] = np.where(drive, uniform(seed, ids, 72, -3.0, 3.0), 0.0)
    params[:, 6] = np.where(drive & (ids % 5 == 0), np.inf, uniform(seed, ids, 73, 0.2, 3.0))
    both = (sc.joints["body0"] >= 0) & (sc.joints["body1"] >= 0)
    params[drive & both, 7] = 1.5
    # soft limits (springs beyond the stops) on every fifth joint
    soft = (ids % 5 == 4) & ~locked & ~bouncy
    params[soft, 8] = np.where(rev[soft], 20.0, 200.0)
    params[soft, 9] = np.where(rev[soft], 0.5, 5.0)
    sc.joints["type"] = kind
    sc.joints["params"] = params
    sc.name = "limited_joints"
    sc.params = {"n_groups": n_groups, "seed": seed}
    return sc


def cone_limited_joints(n_groups=48, seed=7) -> Scene:
    """joint_singles' groups with spherical joints whose cone limits have
    y / z angles U(0.1, 0.8) rad (every fifth soft: stiffness 30 N m/rad,
    damping 0.5); the random starting velocities swing them into the cones'
    rims within a few steps."""
    sc = joint_singles(n_groups, seed)
    k = len(sc.joints["type"])
    ids = np.arange(k)
    p = np.zeros((k, 10), np.float32)
    p[:, 0] = 1
    p[:, 7] = 1.0
    p[:, 1] = uniform(seed, ids, 80, 0.1, 0.8)
    p[:, 2] = uniform(seed, ids, 81, 0.1, 0.8)
    p[ids % 5 == 4, 8] = 30.0
    p[ids % 5 == 4, 9] = 0.5
    sc.joints["type"] = np.full(k, JOINT_SPHERICAL, np.int32)
    sc.joints["params"] = p
    sc.name = "cone_limited_joints"
    sc.params = {"n_groups": n_groups, "seed": seed}
    return sc


def distance_joints(n_groups=48, seed=7, equal=True) -> Scene:
    """joint_singles' groups with distance joints: the first frames' anchors
    kept at most U(0.05, 0.3) m apart (every third also at least that less
    0.05 m; with equal, every seventh exactly that distance; every fifth a
    spring of 100 N/m, damping 2), tolerance 0.01 m."""
    sc = joint_singles(n_groups, seed)
    k = len(sc.joints["type"])
    ids = np.arange(k)
    p = np.zeros((k, 10), np.float32)
    p[:, 7] = 1.0
    mx = uniform(seed, ids, 82, 0.05, 0.3)
    p[:, 0] = 1
    p[:, 2] = mx
    p[ids % 3 == 0, 0] = 3
    p[ids % 3 == 0, 1] = mx[ids % 3 == 0] - 0.05
    eq = (ids % 7 == 2) & equal
    p[eq, 0] = 3
    p[eq, 1] = mx[eq]
    p[:, 3] = 0.01
    spring = ids % 5 == 4
    p[spring, 0] = p[spring, 0] + 4
    p[spring, 8] = 100.0
    p[spring, 9] = 2.0
    sc.joints["type"] = np.full(k, JOINT_DISTANCE, np.int32)
    sc.joints["params"] = p
    sc.name = "distance_joints"
    sc.params = {"n_groups": n_groups, "seed": seed}
    return sc


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + bw * ax + ay * bz - by * az, aw * by + bw * ay + az * bx - bz * ax,
                     aw * bz + bw * az + ax * by - bx * ay, aw * bw - ax * bx - ay * by - az * bz])


def _qrot(q, v):
    u = np.asarray(q[:3], np.float64)
    w = q[3]
    v = np.asarray(v, np.float64)
    return v + 2.0 * w * np.cross(u, v) + 2.0 * np.cross(u, np.cross(u, v))


def jointed_chains(n_scenes=8, links=12, drops=6, seed=3) -> Scene:
    """Chains of boxes (half extents 0.25 x 0.08 x 0.08 m) laid along x at
    2.5 m above a floor, the first link joined to the world at its free end,
    each link to the next at their shared end (seeded jitter of 2 mm / 0.02
    rad per link, so joints start with small errors). World w's joints are
    all of kind w % 4 (fixed: a sagging cantilever; spherical; revolute about
    z; prismatic along x, which lets the chain slide). Each world also drops
    `drops` boxes onto its chain, so contacts and joints act on the same
    bodies. Worlds are independent."""
    types, dims, pos, quat, wid = [], [], [], [], []
    jt, jb0, jb1, jf0, jf1 = [], [], [], [], []
    qz = np.array([0.0, -math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4)])   # frame x axis -> world z
    planes, pscene = [], []
    for w in range(n_scenes):
        kind = w % 4
        # revolute: twist (frame x) about world z; the others: identity frames
        jq = qz if kind == JOINT_REVOLUTE else np.array([0.0, 0.0, 0.0, 1.0])
        first = len(types)
        for i in range(links):
            ids = np.array([w * 1000 + i])
            jit = np.array([uniform(seed, ids, d, -0.002, 0.002)[0] for d in range(3)])
            types.append(BOX)
            dims.append([0.25, 0.08, 0.08, 0.0])
            pos.append(np.array([0.5 * i, 2.5, 0.0]) + jit)
            quat.append(_quat_axis_angle([0.0, 0.0, 1.0], uniform(seed, ids, 3, -0.02, 0.02)[0]))
            wid.append(w)
        for i in range(links):
            if i == 0:
                jb0.append(-1); jb1.append(first)
                jf0.append(np.concatenate([[-0.25, 2.5, 0.0], jq])); jf1.append(np.concatenate([[-0.25, 0.0, 0.0], jq]))
            else:
                jb0.append(first + i - 1); jb1.append(first + i)
                jf0.append(np.concatenate([[0.25, 0.0, 0.0], jq])); jf1.append(np.concatenate([[-0.25, 0.0, 0.0], jq]))
            jt.append(kind)
        for k in range(drops):
            ids = np.array([w * 1000 + 500 + k])
            types.append(BOX)
            h = [uniform(seed, ids, d, 0.1, 0.2)[0] for d in range(3)]
            dims.append(h + [0.0])
            pos.append(np.array([0.5 + 0.9 * k, 3.2 + 0.5 * (k % 2), uniform(seed, ids, 4, -0.05, 0.05)[0]]))
            quat.append(_quat_axis_angle([0.0, 1.0, 0.0], uniform(seed, ids, 5, 0.0, 3.0)[0]))
            wid.append(w)
        planes.append([0.0, 1.0, 0.0, 0.0])
        pscene.append(w)
    st = np.array(types, np.int32)
    sc = _finish(st, np.array(dims), np.array(pos), np.array(quat), np.array(wid), np.array(planes),
                 np.array(pscene), n_scenes, "jointed_chains",
                 {"n_scenes": n_scenes, "links": links, "drops": drops, "seed": seed})
    sc.joints = {"type": np.array(jt, np.int32), "body0": np.array(jb0, np.int32), "body1": np.array(jb1, np.int32),
                 "frame0": np.array(jf0, np.float32), "frame1": np.array(jf1, np.float32)}
    return sc


def aggregate_pile(n_groups=64, per_group=4, seed=7) -> Scene:
    """Groups of per_group boxes whose members overlap (half extents 0.15 to
    0.25 m, 0.3 m apart along x), laid out 1.4 m apart in a grid above a
    plane, each group an aggregate; every other group with self-collisions
    on. Without self-collisions a group's members pass through each other;
    groups collide with each other and the floor."""
    n = n_groups * per_group
    ids = np.arange(n)
    grp = ids // per_group
    k = ids % per_group
    side = int(math.ceil(math.sqrt(n_groups)))
    gx, gz = grp % side, grp // side
    h = np.stack([uniform(seed, ids, d, 0.15, 0.25) for d in range(3)], 1)
    pos = np.stack([(gx - side / 2) * 1.4 + 0.3 * k - 0.45, 0.6 + 0.2 * (k % 2) + uniform(seed, ids, 3, 0, 0.1),
                    (gz - side / 2) * 1.4 + uniform(seed, ids, 4, -0.05, 0.05)], 1)
    quat = _quat_axis_angle(np.stack([uniform(seed, ids, 5, -1, 1), uniform(seed, ids, 6, -1, 1),
                                      uniform(seed, ids, 7, -1, 1)], 1), uniform(seed, ids, 8, 0, 0.6))
    dims = np.concatenate([h, np.zeros((n, 1))], 1)
    sc = _finish(np.full(n, BOX, np.int32), dims, pos, quat, np.zeros(n, np.int32), [[0, 1, 0, 0]], [0], 1,
                 "aggregate_pile", {"n_groups": n_groups, "per_group": per_group, "seed": seed})
    sc.aggregates = [(np.arange(g * per_group, (g + 1) * per_group, dtype=np.int32), g % 2 == 1)
                     for g in range(n_groups)]
    return sc


def terrain_mesh(cells=16, size=16.0, amp=0.4, seed=1):
    """A terrain as a triangle mesh: a (cells+1)^2 grid of heights over a
    size x size square centred at the origin, y = amp * (0.6 sin(a x + p)
    cos(b z + q) + 0.4 sin(c (x + z) + r)) with seeded wave numbers and
    phases, two triangles per cell wound counter-clockwise about +y.
    Returns (verts (nv,3) f32, tris (nt,3) int32)."""
    k = np.arange(6)
    a, b, c = (uniform(seed, k[i:i + 1], 0, 0.4, 1.2)[0] for i in range(3))
    p, q, r = (uniform(seed, k[i:i + 1], 1, 0.0, 2 * math.pi)[0] for i in range(3, 6))
    xs = np.linspace(-size / 2, size / 2, cells + 1)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    Y = amp * (0.6 * np.sin(a * X + p) * np.cos(b * Z + q) + 0.4 * np.sin(c * (X + Z) + r))
    verts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1).astype(np.float32)
    i, j = np.meshgrid(np.arange(cells), np.arange(cells), indexing="ij")
    v00 = (i * (cells + 1) + j).ravel()
    v01, v10 = v00 + 1, v00 + cells + 1
    v11 = v10 + 1
    tris = np.concatenate([np.stack([v00, v01, v10], 1), np.stack([v01, v11, v10], 1)]).astype(np.int32)
    return verts, tris


def terrain_drop(n_scenes=1, per_side=6, layers=1, cells=16, size=16.0, amp=0.4, seed=9) -> Scene:
    """mixed_pile's spheres, capsules, boxes and hulls (per_side^2 x layers per
    world) dropped onto a terrain mesh instead of a plane: one terrain_mesh per
    world (the same mesh, identity pose), bodies raised by amp + 0.1 m."""
    base = mixed_pile(n_bodies=per_side * per_side * layers, per_side=per_side, seed=seed)
    verts, tris = terrain_mesh(cells, size, amp, seed)
    n1 = base.n
    rep = lambda a: np.concatenate([a] * n_scenes)  # noqa: E731
    pose = rep(base.pose).copy()
    pose[:, 1] += amp + 0.1
    sc = Scene(rep(base.shape_type), rep(base.dims), pose, rep(base.lin_vel), rep(base.ang_vel), rep(base.mass),
               rep(base.inertia), np.repeat(np.arange(n_scenes, dtype=np.int32), n1), np.zeros((0, 4), np.float32),
               np.zeros(0, np.int32), n_scenes, "terrain_drop",
               dict(n_scenes=n_scenes, per_side=per_side, layers=layers, cells=cells, size=size, amp=amp, seed=seed),
               base.hull_points, base.hull_counts)
    ident = np.array([0, 0, 0, 0, 0, 0, 1], np.float32)
    sc.meshes = [{"verts": verts, "tris": tris, "pose": ident, "world": w} for w in range(n_scenes)]
    return sc


def heightfield_mesh(heights, tess=None, height_scale=1.0, row_scale=1.0, column_scale=1.0):
    """The triangle mesh of a heightfield's cells, as pxh_scene_add_heightfield
    builds it (Gu::HeightField::getTriangleVertexIndices): vertex row * cols +
    col at (row * row_scale, height * height_scale, col * column_scale); a
    cell's diagonal runs from its zeroth vertex when that sample's
    tessellation flag is set. Returns (verts (nv,3) f32, tris (nt,3) int32)."""
    h = np.asarray(heights, np.int16)
    rows, cols = h.shape
    r, c = np.meshgrid(np.arange(rows), np.arange(cols), indexing="ij")
    verts = np.stack([r.astype(np.float32) * np.float32(row_scale), h.astype(np.float32) * np.float32(height_scale),
                      c.astype(np.float32) * np.float32(column_scale)], -1).reshape(-1, 3).astype(np.float32)
    tris = []
    for rr in range(rows - 1):
        for cc in range(cols - 1):
            cell = rr * cols + cc
            if tess is not None and tess[rr][cc]:
                tris += [[cell + cols, cell, cell + cols + 1], [cell + 1, cell + cols + 1, cell]]
            else:
                tris += [[cell, cell + 1, cell + cols], [cell + cols + 1, cell + cols, cell + 1]]
    return verts, np.array(tris, np.int32)


def terrain_heights(rows=17, cols=17, amp=0.4, seed=9, height_scale=0.01):
    """int16 samples of terrain_mesh's height function over a rows x cols grid
    at 1 m spacing (heights in units of height_scale) and a seeded
    checkerboard of tessellation flags"""
    k = np.arange(6)
    a, b, c = (uniform(seed, k[i:i + 1], 0, 0.4, 1.2)[0] for i in range(3))
    p, q, r = (uniform(seed, k[i:i + 1], 1, 0.0, 2 * math.pi)[0] for i in range(3, 6))
    X, Z = np.meshgrid(np.arange(rows, dtype=np.float64), np.arange(cols, dtype=np.float64), indexing="ij")
    Y = amp * (0.6 * np.sin(a * X + p) * np.cos(b * Z + q) + 0.4 * np.sin(c * (X + Z) + r))
    heights = np.round(Y / height_scale).astype(np.int16)
    tess = ((np.arange(rows)[:, None] + np.arange(cols)[None, :] + seed) % 2).astype(np.uint8)
    return heights, tess


def heightfield_drop(n_scenes=1, per_side=6, layers=1, rows=17, cols=17, amp=0.4, seed=9) -> Scene:
    """terrain_drop over a heightfield instead: terrain_heights samples at
    0.01 m, 1 m rows and columns, the field centred on the origin (pose
    (-(rows-1)/2, 0, -(cols-1)/2)); the scene's meshes carry both the
    heightfield and its triangles"""
    sc = terrain_drop(n_scenes, per_side, layers, 16, 16.0, amp, seed)
    heights, tess = terrain_heights(rows, cols, amp, seed)
    hf = {"heights": heights, "tess": tess, "height_scale": 0.01, "row_scale": 1.0, "column_scale": 1.0}
    verts, tris = heightfield_mesh(heights, tess, 0.01, 1.0, 1.0)
    pose = np.array([-(rows - 1) / 2, 0, -(cols - 1) / 2, 0, 0, 0, 1], np.float32)
    sc.meshes = [{"verts": verts, "tris": tris, "pose": pose, "world": w, "heightfield": hf} for w in range(n_scenes)]
    sc.name = "heightfield_drop"
    return sc
