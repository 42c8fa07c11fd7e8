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
        return Scene(self.shape_type[b].copy(), self.dims[b].copy(), self.pose[b].copy(),
                     self.lin_vel[b].copy(), self.ang_vel[b].copy(), self.mass[b].copy(),
                     self.inertia[b].copy(), (self.scene_id[b] - w0).astype(np.int32),
                     self.planes[p].copy(), (self.plane_scene[p] - w0).astype(np.int32), w1 - w0,
                     name, dict(self.params), self.hull_points, self.hull_counts)


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
