"""ctypes bindings of the test-side checkers: the plain-C oracle
(oracle/libpxo.so) and the reference PhysX CPU SDK harness
(oracle/_ref/libpxref.so, built from the reference sources when present).
Only tests, smoke() and bench.py's CPU-baseline leg use these."""
from __future__ import annotations

import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PXO_PATH = os.environ.get("PXO_LIB", os.path.join(ROOT, "oracle", "libpxo.so"))
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "libpxref.so")

F = ctypes.POINTER(ctypes.c_float)
I32 = ctypes.POINTER(ctypes.c_int32)
U32 = ctypes.POINTER(ctypes.c_uint32)
I64 = ctypes.POINTER(ctypes.c_longlong)
D = ctypes.POINTER(ctypes.c_double)
VP = ctypes.c_void_p


def fp(a):
    return a.ctypes.data_as(F)


def ip(a):
    return a.ctypes.data_as(I32)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class PxoDesc(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_float * 3), ("solver", ctypes.c_int), ("pos_iters", ctypes.c_int),
                ("vel_iters", ctypes.c_int), ("contact_offset", ctypes.c_float), ("rest_offset", ctypes.c_float),
                ("static_friction", ctypes.c_float), ("dynamic_friction", ctypes.c_float),
                ("restitution", ctypes.c_float), ("linear_damping", ctypes.c_float),
                ("angular_damping", ctypes.c_float), ("max_lin_vel_sq", ctypes.c_float),
                ("max_ang_vel_sq", ctypes.c_float), ("bounce_threshold", ctypes.c_float),
                ("friction_offset_threshold", ctypes.c_float), ("correlation_distance", ctypes.c_float),
                ("tolerance_length", ctypes.c_float), ("order_salt", ctypes.c_uint),
                ("enable_sleeping", ctypes.c_int), ("sleep_threshold", ctypes.c_float)]


_pxo = None
_ref = None


def pxo():
    global _pxo
    if _pxo is None:
        lib = ctypes.CDLL(PXO_PATH)
        lib.pxo_generate_contacts.restype = ctypes.c_int
        lib.pxo_bp_pairs.restype = ctypes.c_longlong
        lib.pxo_world_create.restype = VP
        lib.pxo_world_step.restype = ctypes.c_int
        lib.pxo_world_constraints.restype = ctypes.c_longlong
        lib.pxo_world_contacts.restype = ctypes.c_longlong
        lib.pxo_hull_add.restype = ctypes.c_int
        for name in ("pxo_world_release", "pxo_world_add_bodies", "pxo_world_add_planes", "pxo_world_step",
                     "pxo_world_get_bodies", "pxo_world_set_bodies", "pxo_world_stats", "pxo_world_constraints",
                     "pxo_world_contacts"):
            getattr(lib, name).argtypes = None
        _pxo = lib
    return _pxo


def have_ref():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(REF_PATH)
        lib.pxref_generate_contacts.restype = ctypes.c_int
        lib.pxref_bp_pairs.restype = ctypes.c_longlong
        lib.pxref_scene_create.restype = VP
        lib.pxref_scene_create2.restype = VP
        lib.pxref_scene_step.restype = ctypes.c_double
        lib.pxref_dynamic_group.restype = ctypes.c_uint32
        lib.pxref_static_group.restype = ctypes.c_uint32
        lib.pxref_convex_create.restype = ctypes.c_int
        _ref = lib
    return _ref


REFGPU_PATH = os.path.join(ROOT, "oracle", "_ref", "gpu", "libpxrefgpu.so")
_refgpu = None


def have_refgpu():
    return os.path.exists(REFGPU_PATH)


def refgpu():
    """The reference SDK built with its GPU support on and the HIP GPU plug-in
    linked in (static plug-in mode), with the broadphase harness."""
    global _refgpu
    if _refgpu is None:
        import physx_amd
        physx_amd.load()   # libpxhip.so first: the plug-in's dependency
        lib = ctypes.CDLL(REFGPU_PATH)
        lib.pxrefgpu_bp_create.restype = VP
        for name in ("pxrefgpu_static_group", "pxrefgpu_dynamic_group", "pxrefgpu_kinematic_group"):
            getattr(lib, name).restype = ctypes.c_uint32
        _refgpu = lib
    return _refgpu


class SdkBroadPhase:
    """PxCreateBroadPhase of the GPU-enabled reference SDK: bp_type 2 eABP,
    3 ePABP, 0 eSAP (CPU), or 4 eGPU (through PxCreatePhysXGpu to the HIP
    plug-in). update() takes the SDK's incremental form: handle lists plus
    persistent bounds / groups / distances arrays of the capacity."""

    SAP, MBP, ABP, PABP, GPU = 0, 1, 2, 3, 4

    def __init__(self, bp_type):
        self.lib = refgpu()
        self.h = VP(self.lib.pxrefgpu_bp_create(int(bp_type), 0, 0))
        if not self.h.value:
            raise RuntimeError(f"PxCreateBroadPhase failed for type {bp_type}")

    def update(self, created, updated, removed, bounds, groups, distances):
        u32 = lambda a: np.ascontiguousarray(a, dtype=np.uint32)  # noqa: E731
        c, u, r, g = u32(created), u32(updated), u32(removed), u32(groups)
        b, d = f32(bounds), f32(distances)
        up = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
        assert self.lib.pxrefgpu_bp_update(self.h, up(c), len(c), up(u), len(u), up(r), len(r), up(b), up(g), up(d),
                                           int(b.shape[0])) == 0
        cnt = np.zeros(2, np.int32)
        self.lib.pxrefgpu_bp_results(self.h, None, 0, None, 0, cnt.ctypes.data_as(ctypes.c_void_p))
        cr = np.zeros((max(cnt[0], 1), 2), np.uint32)
        dl = np.zeros((max(cnt[1], 1), 2), np.uint32)
        self.lib.pxrefgpu_bp_results(self.h, up(cr), int(cnt[0]), up(dl), int(cnt[1]),
                                     cnt.ctypes.data_as(ctypes.c_void_p))
        return cr[: cnt[0]], dl[: cnt[1]]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pxrefgpu_bp_release(self.h)
            self.h = VP()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def contacts(lib_name: str, t0, g0, p0, t1, g1, p1, cd=0.04, tol=1.0):
    """(n,8) contacts: point xyz, separation, normal xyz, 0."""
    out = np.zeros(64 * 8, np.float32)
    g0 = f32(np.resize(np.asarray(g0, np.float32), 4))
    g1 = f32(np.resize(np.asarray(g1, np.float32), 4))
    p0, p1 = f32(p0), f32(p1)
    if lib_name == "pxo":
        n = pxo().pxo_generate_contacts(t0, fp(g0), fp(p0), t1, fp(g1), fp(p1), ctypes.c_float(cd),
                                        ctypes.c_float(tol), fp(out))
    else:
        n = ref().pxref_generate_contacts(t0, fp(g0), fp(p0), t1, fp(g1), fp(p1), ctypes.c_float(cd),
                                          ctypes.c_float(tol), fp(out), 64)
    assert n >= 0
    return out[: 8 * n].reshape(n, 8)


def bp_pairs_pxo(bounds, dist, world=None):
    b, d = f32(bounds), f32(dist)
    n = b.shape[0]
    w = i32(world) if world is not None else None
    cap = max(1024, 64 * n)
    out = np.zeros(2 * cap, np.uint32)
    cnt = pxo().pxo_bp_pairs(fp(b), fp(d), ip(w) if w is not None else None, n, out.ctypes.data_as(U32),
                             ctypes.c_longlong(cap))
    assert cnt <= cap
    return out[: 2 * cnt].reshape(-1, 2)


def bp_pairs_ref(bounds, dist, groups=None, pabp=False):
    lib = ref()
    b, d = f32(bounds), f32(dist)
    n = b.shape[0]
    if groups is None:
        groups = np.array([lib.pxref_dynamic_group(i) for i in range(n)], np.uint32)
    g = np.ascontiguousarray(groups, np.uint32)
    cap = max(1024, 64 * n)
    out = np.zeros(2 * cap, np.uint32)
    cnt = lib.pxref_bp_pairs(fp(b), fp(d), g.ctypes.data_as(U32), n, out.ctypes.data_as(U32),
                             ctypes.c_longlong(cap), int(pabp))
    assert 0 <= cnt <= cap
    p = out[: 2 * cnt].reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def _vp(a):
    return a.ctypes.data_as(VP)


def pxo_add_hull(h):
    """Register a cooked hull (dict of verts, planes, polys, vrefs, internal
    extents) with the oracle; returns its index."""
    v, pl = f32(h["verts"]), f32(h["planes"])
    po, vr = i32(h["polys"]), i32(h["vrefs"])
    ie = f32(h["internal_extents"])
    return pxo().pxo_hull_add(int(v.shape[0]), _vp(v), int(pl.shape[0]), _vp(pl), _vp(po), _vp(vr), _vp(ie))


_TRIMESH_IDS = {}


def ref_trimesh(verts, tris):
    """Cook a triangle mesh with the reference cooker (cached by content);
    returns its mesh index (geometry code 8, dims[0])."""
    v, t = f32(verts).reshape(-1, 3), i32(tris).reshape(-1, 3)
    key = (v.tobytes(), t.tobytes())
    if key not in _TRIMESH_IDS:
        lib = ref()
        lib.pxref_trimesh_create.restype = ctypes.c_int
        mid = lib.pxref_trimesh_create(_vp(v), int(v.shape[0]), _vp(t), int(t.shape[0]))
        assert mid >= 0
        _TRIMESH_IDS[key] = mid
    return _TRIMESH_IDS[key]


_HF_IDS = {}


def ref_heightfield_geom(hf):
    """A heightfield (dict: heights (rows, cols) int16, tess or None,
    height_scale, row_scale, column_scale) created in the reference SDK
    (cached by content); returns its geometry row (index, height scale, row
    scale, column scale) for code 9."""
    h = np.ascontiguousarray(hf["heights"], np.int16)
    t = np.ascontiguousarray(hf["tess"], np.uint8) if hf.get("tess") is not None else None
    key = (h.tobytes(), None if t is None else t.tobytes(), h.shape)
    if key not in _HF_IDS:
        lib = ref()
        lib.pxref_heightfield_create.restype = ctypes.c_int
        hid = lib.pxref_heightfield_create(int(h.shape[0]), int(h.shape[1]), _vp(h), _vp(t) if t is not None else None)
        assert hid >= 0
        _HF_IDS[key] = hid
    return np.array([_HF_IDS[key], hf["height_scale"], hf["row_scale"], hf["column_scale"]], np.float32)


def ref_cook(points):
    """Cook a hull with the reference cooker; returns (mesh index, hull dict)."""
    lib = ref()
    pts = f32(points).reshape(-1, 3)
    mid = lib.pxref_convex_create(_vp(pts), int(pts.shape[0]))
    assert mid >= 0
    V = np.zeros((255, 3), np.float32)
    Pl = np.zeros((255, 4), np.float32)
    Po = np.zeros((255, 2), np.int32)
    R = np.zeros(1530, np.int32)
    M = np.zeros(13, np.float64)
    S = np.zeros(10, np.float32)
    C = np.zeros(3, np.int32)
    lib.pxref_convex_data(mid, _vp(V), _vp(Pl), _vp(Po), _vp(R), _vp(M), _vp(S), _vp(C))
    nv, npol, nr = (int(x) for x in C)
    return mid, dict(verts=V[:nv].copy(), planes=Pl[:npol].copy(), polys=Po[:npol].copy(), vrefs=R[:nr].copy(),
                     poly_counts=Po[:npol, 1].copy(), mass=M[0], com=M[1:4].copy(),
                     inertia_origin=M[4:13].reshape(3, 3).copy(), bounds_center=S[0:3].copy(),
                     bounds_extents=S[3:6].copy(), internal_radius=float(S[6]), internal_extents=S[7:10].copy())


def _hull_clouds(scene):
    if scene.hull_counts is None or not len(scene.hull_counts):
        return []
    off = np.concatenate([[0], np.cumsum(scene.hull_counts)])
    return [scene.hull_points[off[i]:off[i + 1]] for i in range(len(scene.hull_counts))]


def default_desc(pos_iters=4, vel_iters=1, gravity=(0.0, -9.81, 0.0)):
    d = PxoDesc()
    pxo().pxo_default_desc(ctypes.byref(d))
    d.pos_iters, d.vel_iters = pos_iters, vel_iters
    d.gravity[0], d.gravity[1], d.gravity[2] = gravity
    return d


class OracleWorld:
    """The plain-C oracle's world (TGS or PGS)."""

    def __init__(self, scene, pos_iters=4, vel_iters=1, gravity=(0.0, -9.81, 0.0), friction=0.5, solver="tgs",
                 order_salt=0, sleep_threshold=0.0):
        self.lib = pxo()
        self.desc = default_desc(pos_iters, vel_iters, gravity)
        self.desc.static_friction = self.desc.dynamic_friction = friction
        self.desc.solver = 1 if solver == "tgs" else 0
        self.desc.order_salt = order_salt
        self.desc.enable_sleeping = int(sleep_threshold > 0.0)
        self.desc.sleep_threshold = sleep_threshold
        self.h = VP(self.lib.pxo_world_create(ctypes.byref(self.desc)))
        self.n = scene.n
        dims = f32(scene.dims).copy()
        clouds = _hull_clouds(scene)
        if clouds:
            # the hulls the GPU library cooks from the same point clouds
            import physx_amd as P
            ids = np.array([pxo_add_hull(P.cook_convex(c)) for c in clouds], np.float32)
            cvx = scene.shape_type == 5
            dims[cvx, 0] = ids[dims[cvx, 0].astype(np.int64)]
        self._keep = [i32(scene.shape_type), dims, f32(scene.pose), f32(scene.lin_vel),
                      f32(scene.ang_vel), f32(scene.mass), f32(scene.inertia), i32(scene.scene_id)]
        t, dm, po, lv, av, m, inn, sc = self._keep
        self.lib.pxo_world_add_bodies(self.h, self.n, ip(t), fp(dm), fp(po), fp(lv), fp(av), fp(m), fp(inn), ip(sc))
        pl, ps = f32(scene.planes), i32(scene.plane_scene)
        self._keep += [pl, ps]
        self.lib.pxo_world_add_planes(self.h, int(pl.shape[0]), fp(pl), ip(ps))
        for members, self_coll in (getattr(scene, "aggregates", None) or []):
            m = i32(members)
            assert self.lib.pxo_world_add_aggregate(self.h, int(m.size), ip(m), int(bool(self_coll))) >= 0
        for m in (getattr(scene, "meshes", None) or []):
            v, t, po = f32(m["verts"]), i32(m["tris"]), f32(m["pose"])
            self._keep += [v, t, po]
            assert self.lib.pxo_world_add_trimesh(self.h, int(v.shape[0]), fp(v), int(t.shape[0]), ip(t), fp(po),
                                                  int(m["world"])) >= 0
        j = getattr(scene, "joints", None)
        if j is not None and len(j["type"]):
            arrs = [i32(j["type"]), i32(j["body0"]), i32(j["body1"]), f32(j["frame0"]), f32(j["frame1"])]
            self._keep += arrs
            r = self.lib.pxo_world_add_joints(self.h, int(arrs[0].shape[0]), ip(arrs[0]), ip(arrs[1]), ip(arrs[2]),
                                              fp(arrs[3]), fp(arrs[4]))
            assert r >= 0, "oracle rejected the joints"
            if j.get("params") is not None:
                jj, pp = i32(np.arange(len(j["type"]))), f32(j["params"])
                self._keep += [jj, pp]
                assert self.lib.pxo_world_set_joint_params(self.h, int(jj.size), ip(jj), fp(pp)) == 0

    def step(self, dt, n=1):
        for _ in range(n):
            r = self.lib.pxo_world_step(self.h, ctypes.c_float(dt))
            assert r == 0, f"oracle step failed ({r})"

    def bodies(self):
        pose = np.zeros((self.n, 7), np.float32)
        lv = np.zeros((self.n, 3), np.float32)
        av = np.zeros((self.n, 3), np.float32)
        self.lib.pxo_world_get_bodies(self.h, fp(pose), fp(lv), fp(av))
        return pose, lv, av

    def sleep_state(self):
        """(wake counter, asleep) per body"""
        wc = np.zeros(self.n, np.float32)
        asleep = np.zeros(self.n, np.int32)
        self.lib.pxo_world_get_sleep(self.h, fp(wc), ip(asleep))
        return wc, asleep.astype(bool)

    def wake(self, bodies):
        b = i32(bodies)
        self.lib.pxo_world_wake(self.h, int(b.size), ip(b))

    def set_bodies(self, pose, lv, av):
        """pose / linear / angular velocity arrays (any may be None: kept)"""
        arrs = [f32(a) if a is not None else None for a in (pose, lv, av)]
        self.lib.pxo_world_set_bodies(self.h, *[fp(a) if a is not None else None for a in arrs])

    def set_forces(self, force=None, torque=None):
        """addForce / addTorque (eFORCE) for the next step, (n,3) arrays."""
        f = f32(force) if force is not None else None
        t = f32(torque) if torque is not None else None
        self.lib.pxo_world_set_forces(self.h, fp(f) if f is not None else None, fp(t) if t is not None else None)

    def stats(self):
        out = np.zeros(6, np.int64)
        self.lib.pxo_world_stats(self.h, out.ctypes.data_as(I64))
        return dict(bp_pairs=int(out[0]), plane_tests=int(out[1]), constraints=int(out[2]), contacts=int(out[3]),
                    colors=int(out[4]), anchors=int(out[5]))

    def constraints(self):
        cap = 1 << 22
        out = np.zeros((cap, 4), np.int32)
        n = self.lib.pxo_world_constraints(self.h, ip(out), ctypes.c_longlong(cap))
        return out[:n]

    def contact_data(self):
        """(k, 32) records like the library's contact data (oracle constraint order)."""
        cap = 1 << 21
        out = np.zeros((cap, 32), np.float32)
        n = self.lib.pxo_world_contacts(self.h, fp(out), ctypes.c_longlong(cap))
        return out[:n].copy()

    def __del__(self):
        try:
            self.lib.pxo_world_release(self.h)
        except Exception:
            pass


class ChunkedOracle:
    """The oracle over a scene of independent worlds split into chunks of
    whole worlds, each chunk an OracleWorld stepped on its own thread (the
    library call releases the GIL). Every stage of the oracle keys its work by
    world (world-local colour keys and pair keys, per-world broadphase), so
    the chunks together give exactly what one OracleWorld over all worlds
    gives. Scenes with convex meshes are not split (the hull registry is
    shared)."""

    def __init__(self, scene, chunks=16, **kw):
        from concurrent.futures import ThreadPoolExecutor
        assert scene.hull_counts is None or len(scene.hull_counts) == 0
        nw = scene.n_scenes
        bounds = np.linspace(0, nw, min(chunks, nw) + 1).astype(int)
        self.parts = [scene.subset_range(int(a), int(b)) for a, b in zip(bounds[:-1], bounds[1:])]
        self.body0 = np.concatenate([[0], np.cumsum([p.n for p in self.parts])])
        self.plane0 = np.concatenate([[0], np.cumsum([p.planes.shape[0] for p in self.parts])])
        self.pool = ThreadPoolExecutor(max_workers=len(self.parts))
        self.worlds = list(self.pool.map(lambda p: OracleWorld(p, **kw), self.parts))
        self.n = scene.n

    def step(self, dt, n=1):
        list(self.pool.map(lambda w: w.step(dt, n), self.worlds))

    def bodies(self):
        parts = [w.bodies() for w in self.worlds]
        return tuple(np.concatenate([p[k] for p in parts]) for k in range(3))

    def stats(self):
        out = {}
        for w in self.worlds:
            for k, v in w.stats().items():
                out[k] = max(out.get(k, 0), v) if k == "colors" else out.get(k, 0) + v
        return out

    def _shift(self, ids, c):
        """chunk-local body (>= 0) and plane (-1 - p) ids to scene ids"""
        return np.where(ids >= 0, ids + self.body0[c], ids - self.plane0[c]).astype(np.int32)

    def constraints(self):
        out = []
        for c, w in enumerate(self.worlds):
            t = w.constraints().copy()
            t[:, 0] = self._shift(t[:, 0], c)
            t[:, 1] = self._shift(t[:, 1], c)
            out.append(t)
        return np.concatenate(out)

    def contact_data(self):
        out = []
        for c, w in enumerate(self.worlds):
            r = w.contact_data()
            ids = r[:, :2].copy().view(np.int32)
            r[:, :2] = self._shift(ids, c).view(np.float32)
            out.append(r)
        return np.concatenate(out)


class RefScene:
    """The reference PhysX CPU PxScene (TGS or PGS) on the same arrays."""

    def __init__(self, scene, solver="tgs", pos_iters=4, vel_iters=1, threads=8, gravity=(0.0, -9.81, 0.0),
                 pabp=False, friction=0.5, report_contacts=False, sleep_threshold=0.0):
        self.lib = ref()
        g = f32(gravity)
        self.h = VP(self.lib.pxref_scene_create2(fp(g), 1 if solver == "tgs" else 0, int(pabp), int(threads),
                                                 int(report_contacts)))
        assert self.h.value
        self.lib.pxref_scene_set_material(self.h, ctypes.c_float(friction), ctypes.c_float(friction),
                                          ctypes.c_float(0.0))
        self.n = scene.n
        dims = f32(scene.dims).copy()
        clouds = _hull_clouds(scene)
        if clouds:
            ids = np.array([ref_cook(c)[0] for c in clouds], np.float32)
            cvx = scene.shape_type == 5
            dims[cvx, 0] = ids[dims[cvx, 0].astype(np.int64)]
        self._keep = [i32(scene.shape_type), dims, f32(scene.pose), f32(scene.lin_vel),
                      f32(scene.ang_vel), f32(scene.mass), f32(scene.inertia)]
        t, dm, po, lv, av, m, inn = self._keep
        self.lib.pxref_scene_add_bodies(self.h, self.n, ip(t), fp(dm), fp(po), fp(lv), fp(av), fp(m), fp(inn),
                                        pos_iters, vel_iters, ctypes.c_float(0.02), ctypes.c_float(0.0),
                                        ctypes.c_float(0.0), ctypes.c_float(0.05))
        pl = f32(scene.planes)
        self._keep.append(pl)
        assert self.lib.pxref_scene_add_planes(self.h, int(pl.shape[0]), fp(pl), ctypes.c_float(0.02)) == 0
        for members, self_coll in (getattr(scene, "aggregates", None) or []):
            m = i32(members)
            assert self.lib.pxref_scene_add_aggregate(self.h, int(m.size), ip(m), int(bool(self_coll))) == 0
        for m in (getattr(scene, "meshes", None) or []):
            po = f32(m["pose"])
            self._keep.append(po)
            if "heightfield" in m:
                g = ref_heightfield_geom(m["heightfield"])
                self._keep.append(g)
                assert self.lib.pxref_scene_add_heightfield(self.h, fp(g), fp(po), ctypes.c_float(0.02)) == 0
            else:
                mid = ref_trimesh(m["verts"], m["tris"])
                assert self.lib.pxref_scene_add_trimesh(self.h, mid, fp(po), ctypes.c_float(0.02)) == 0
        j = getattr(scene, "joints", None)
        if j is not None and len(j["type"]):
            arrs = [i32(j["type"]), i32(j["body0"]), i32(j["body1"]), f32(j["frame0"]), f32(j["frame1"])]
            self._keep += arrs
            assert self.lib.pxref_scene_add_joints(self.h, int(arrs[0].shape[0]), ip(arrs[0]), ip(arrs[1]),
                                                   ip(arrs[2]), fp(arrs[3]), fp(arrs[4])) == arrs[0].shape[0]
            if j.get("params") is not None:
                jj, pp = i32(np.arange(len(j["type"]))), f32(j["params"])
                self._keep += [jj, pp]
                assert self.lib.pxref_scene_set_joint_params(self.h, int(jj.size), ip(jj), fp(pp)) == 0
        if sleep_threshold:
            assert self.lib.pxref_scene_set_sleep_threshold(self.h, ctypes.c_float(sleep_threshold)) == 0

    def sleep_state(self):
        """(wake counter, asleep) per body: PxRigidDynamic::getWakeCounter,
        isSleeping"""
        wc = np.zeros(self.n, np.float32)
        asleep = np.zeros(self.n, np.int32)
        self.lib.pxref_scene_sleep_state(self.h, fp(wc), ip(asleep))
        return wc, asleep.astype(bool)

    def add_forces(self, force=None, torque=None):
        """PxRigidDynamic::addForce / addTorque (eFORCE) on every body, (n,3) arrays."""
        f = f32(force) if force is not None else None
        t = f32(torque) if torque is not None else None
        self.lib.pxref_scene_add_forces(self.h, fp(f) if f is not None else None, fp(t) if t is not None else None)

    def set_threads(self, threads):
        """Dispatcher of `threads` workers for the following steps
        (PxTaskManager::setCpuDispatcher between steps)."""
        assert self.lib.pxref_scene_set_threads(self.h, int(threads)) == 0

    def step(self, dt, n=1):
        ms = np.zeros(n, np.float64)
        total = self.lib.pxref_scene_step(self.h, ctypes.c_float(dt), n, ms.ctypes.data_as(D))
        return total, ms

    def bodies(self):
        pose = np.zeros((self.n, 7), np.float32)
        lv = np.zeros((self.n, 3), np.float32)
        av = np.zeros((self.n, 3), np.float32)
        self.lib.pxref_scene_get_bodies(self.h, fp(pose), fp(lv), fp(av))
        return pose, lv, av

    def contacts(self):
        """(k, 10) contact rows of the last step: actor a, actor b (body i -> i + 1,
        plane j -> -(j + 1)), point xyz, separation, normal xyz (from b towards
        a), impulse magnitude."""
        cap = 1 << 20
        out = np.zeros((cap, 10), np.float32)
        k = self.lib.pxref_scene_contacts(self.h, fp(out), cap)
        assert k >= 0, "scene created without contact reports"
        return out[: min(k, cap)].copy()

    def stats(self):
        out = np.zeros(4, np.int64)
        self.lib.pxref_scene_stats(self.h, out.ctypes.data_as(I64))
        return dict(pairs=int(out[0]), bp_adds=int(out[1]), active=int(out[2]), touching=int(out[3]))

    def __del__(self):
        try:
            self.lib.pxref_scene_release(self.h)
        except Exception:
            pass


CONVEX_CLASSES = ("convex-convex", "box-convex", "sphere-convex", "plane-convex", "capsule-convex", "capsule-box")


class ConvexGolden:
    """tests/golden/convex_pairs.npz: touching pairs of piles formed in the
    reference SDK with PxGenerateContacts' fresh-cache contacts, and the
    reference-cooked hulls they use (tests/golden/make_golden.py)."""

    def __init__(self):
        # every array read once (an npz decompresses an array at each access)
        with np.load(os.path.join(ROOT, "tests", "golden", "convex_pairs.npz")) as z:
            self.g = {k: z[k] for k in z.files}
        g = self.g
        self.nv, self.np_, self.nr = g["hull_nv"], g["hull_np"], g["hull_nr"]
        self._ov = np.concatenate([[0], np.cumsum(self.nv)])
        self._op = np.concatenate([[0], np.cumsum(self.np_)])
        self._or = np.concatenate([[0], np.cumsum(self.nr)])

    @property
    def n_hulls(self):
        return int(self.nv.shape[0])

    def hull(self, h):
        g = self.g
        return dict(verts=g["hull_verts"][self._ov[h]:self._ov[h + 1]],
                    planes=g["hull_planes"][self._op[h]:self._op[h + 1]],
                    polys=g["hull_polys"][self._op[h]:self._op[h + 1]].astype(np.int32),
                    vrefs=g["hull_vrefs"][self._or[h]:self._or[h + 1]].astype(np.int32),
                    internal_extents=g["hull_ie"][h], com=g["hull_com"][h])

    def pairs(self, cls):
        """(types (k,2), geoms (k,8), poses (k,14), counts (k,), contacts (k,4,8))"""
        k = cls.replace("-", "_")
        g = self.g
        return g[k + "_t"], g[k + "_g"], g[k + "_p"], g[k + "_n"], g[k + "_c"]

    def register_with_oracle(self):
        """every hull with the oracle, hull h as oracle hull h"""
        pxo().pxo_hull_clear()
        for h in range(self.n_hulls):
            assert pxo_add_hull(self.hull(h)) == h


def dedupe(c):
    """contacts as a set: coincident points (to 1e-5) counted once"""
    return np.unique(np.round(c, 5), axis=0) if len(c) else c


def same_contacts(a, b):
    """None when the contact sets agree (1e-4 m points and separations,
    1e-3 on the normal), else what differs: 'count', 'normal' or 'point'"""
    if len(a) != len(b):
        return "count"
    if len(a) == 0:
        return None
    if np.abs(a[0, 4:7] - b[0, 4:7]).max() > 1e-3:
        return "normal"
    used = set()
    for r in a:
        d = np.abs(b[:, :4] - r[:4]).max(1)
        for j in used:
            d[j] = 9.0
        j = int(np.argmin(d))
        if d[j] > 1e-4:
            return "point"
        used.add(j)
    return None
