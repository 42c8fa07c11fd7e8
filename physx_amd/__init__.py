"""physx_amd: MI355X (gfx950) GPU rigid-body step for PhysX scenes.

Python mirror of the reference's scene-level interface for the GPU path
(PxScene::simulate/fetchResults, PxDirectGPUAPI, PxCreateBroadPhase,
PxGenerateContacts) over the C ABI of include/pxhip.h, implemented by the
HIP library physx_amd/libpxhip.so. There is no CPU fallback: loading fails
loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import scenes  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpxhip.so")

PXH_GEOM_SPHERE, PXH_GEOM_PLANE, PXH_GEOM_CAPSULE, PXH_GEOM_BOX, PXH_GEOM_CONVEXMESH = 0, 1, 2, 3, 5
PXH_SOLVER_PGS, PXH_SOLVER_TGS = 0, 1
PXH_SOLVE_LAYOUT_AUTO, PXH_SOLVE_LAYOUT_COLORS, PXH_SOLVE_LAYOUT_WORLDS = 0, 1, 2
PXH_DATA_GLOBAL_POSE, PXH_DATA_LINEAR_VELOCITY, PXH_DATA_ANGULAR_VELOCITY = 0, 1, 2
PXH_READ_LINEAR_ACCELERATION, PXH_READ_ANGULAR_ACCELERATION = 3, 4
PXH_WRITE_FORCE, PXH_WRITE_TORQUE = 3, 4
PXH_BP_NOALLOC_MAX_BOXES, PXH_BP_NOALLOC_MAX_PAIRS = 65536, 262144
PXH_JOINT_FIXED, PXH_JOINT_SPHERICAL, PXH_JOINT_REVOLUTE, PXH_JOINT_PRISMATIC = 0, 1, 2, 3

_F = ctypes.POINTER(ctypes.c_float)
_I32 = ctypes.POINTER(ctypes.c_int32)
_I64 = ctypes.POINTER(ctypes.c_int64)
_VP = ctypes.c_void_p


class PxhError(RuntimeError):
    pass


class SceneDesc(ctypes.Structure):
    """pxh_scene_desc (include/pxhip.h), PxSceneDesc/PxRigidDynamic defaults."""
    _fields_ = [("gravity", ctypes.c_float * 3), ("solver_type", ctypes.c_int32),
                ("position_iterations", ctypes.c_int32), ("velocity_iterations", ctypes.c_int32),
                ("contact_offset", ctypes.c_float), ("rest_offset", ctypes.c_float),
                ("static_friction", ctypes.c_float), ("dynamic_friction", ctypes.c_float),
                ("restitution", ctypes.c_float), ("linear_damping", ctypes.c_float),
                ("angular_damping", ctypes.c_float), ("max_linear_velocity_sq", ctypes.c_float),
                ("max_angular_velocity_sq", ctypes.c_float), ("bounce_threshold_velocity", ctypes.c_float),
                ("friction_offset_threshold", ctypes.c_float), ("friction_correlation_distance", ctypes.c_float),
                ("tolerance_length", ctypes.c_float), ("max_pairs_per_body", ctypes.c_int32),
                ("enable_body_accelerations", ctypes.c_int32), ("report_touch_events", ctypes.c_int32),
                ("enable_sleeping", ctypes.c_int32), ("sleep_threshold", ctypes.c_float)]


EXPORTED = ("pxh_get_version", "pxh_last_error", "pxh_default_scene_desc", "pxh_scene_create", "pxh_scene_release",
            "pxh_scene_add_bodies", "pxh_scene_add_planes", "pxh_scene_simulate", "pxh_scene_fetch_results",
            "pxh_scene_get_rigid_dynamic_data", "pxh_scene_set_rigid_dynamic_data", "pxh_scene_get_stats",
            "pxh_scene_get_constraints", "pxh_scene_get_stream", "pxh_bp_create", "pxh_bp_update",
            "pxh_bp_fetch_results", "pxh_bp_release", "pxh_generate_contacts", "pxh_scene_set_profiling",
            "pxh_scene_get_profile", "pxh_scene_remove_worlds", "pxh_scene_world_count", "pxh_scene_world_range",
            "pxh_server_create", "pxh_server_release", "pxh_server_last_error", "pxh_server_submit",
            "pxh_server_step", "pxh_server_result", "pxh_server_finished", "pxh_server_stats", "pxh_convex_hull",
            "pxh_scene_add_convex_meshes", "pxh_scene_copy_contact_data", "pxh_scene_set_tuning",
            "pxh_scene_get_tuning", "pxh_scene_get_tuning_log", "pxh_bp_update_handles", "pxh_scene_get_touch_events",
            "pxh_scene_add_cooked_convex", "pxh_bp_update_noalloc",
            "pxh_scene_set_profiling_stride", "pxh_scene_add_joints", "pxh_scene_set_joint_params", "pxh_scene_add_aggregate",
            "pxh_scene_get_sleep_state", "pxh_scene_wake_bodies", "pxh_scene_add_triangle_mesh",
            "pxh_scene_add_heightfield")

_lib = None


def load():
    """Load libpxhip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        # share the process's HIP runtime with torch (same SONAME, loaded first)
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise PxhError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       "(make -C physx_amd/csrc)")
    lib = ctypes.CDLL(LIB_PATH)
    lib.pxh_last_error.restype = ctypes.c_char_p
    lib.pxh_server_last_error.restype = ctypes.c_char_p
    lib.pxh_scene_get_stream.restype = _VP
    lib.pxh_scene_get_constraints.restype = ctypes.c_int64
    for name in ("pxh_scene_create", "pxh_scene_release", "pxh_scene_add_bodies", "pxh_scene_add_planes",
                 "pxh_scene_simulate", "pxh_scene_fetch_results", "pxh_scene_get_rigid_dynamic_data",
                 "pxh_scene_set_rigid_dynamic_data", "pxh_scene_get_stats", "pxh_bp_create", "pxh_bp_update",
                 "pxh_bp_fetch_results", "pxh_bp_release", "pxh_generate_contacts", "pxh_scene_set_profiling",
                 "pxh_scene_get_profile"):
        getattr(lib, name).restype = ctypes.c_int32
    _lib = lib
    return lib


def _check(rc):
    if rc < 0:
        raise PxhError(f"pxhip error {rc}: {load().pxh_last_error().decode()}")
    return rc


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def default_desc(**kw) -> SceneDesc:
    d = SceneDesc()
    load().pxh_default_scene_desc(ctypes.byref(d))
    for k, v in kw.items():
        if k == "gravity":
            d.gravity[0], d.gravity[1], d.gravity[2] = v
        else:
            setattr(d, k, v)
    return d


def _dptr(t):
    return ctypes.c_void_p(t.data_ptr())


class Scene:
    """A GPU scene (one or more independent worlds) stepped on the device.

    Mirrors PxScene for this path: simulate(dt) / fetch_results(), and the
    Direct GPU API body-state getters/setters on device buffers."""

    def __init__(self, scene=None, desc: SceneDesc | None = None, **kw):
        self.lib = load()
        self.desc = desc if desc is not None else default_desc(**kw)
        h = ctypes.c_void_p()
        _check(self.lib.pxh_scene_create(ctypes.byref(self.desc), ctypes.byref(h)))
        self.h = h
        self.n = 0
        if scene is not None:
            self.add(scene)

    def add(self, scene):
        t, dm, po = _i32(scene.shape_type), _f32(scene.dims), _f32(scene.pose)
        if scene.hull_counts is not None and len(scene.hull_counts):
            # the scene's convex meshes, renumbered after those already registered
            hc, hp = _i32(scene.hull_counts), _f32(scene.hull_points)
            first = ctypes.c_int32()
            _check(self.lib.pxh_scene_add_convex_meshes(self.h, int(hc.size), hc.ctypes.data_as(_I32),
                                                        hp.ctypes.data_as(_F), ctypes.byref(first)))
            dm = dm.copy()
            cvx = t == PXH_GEOM_CONVEXMESH
            dm[cvx, 0] += first.value
        lv, av, m, inn = _f32(scene.lin_vel), _f32(scene.ang_vel), _f32(scene.mass), _f32(scene.inertia)
        w = _i32(scene.scene_id)
        fp = lambda a: a.ctypes.data_as(_F)  # noqa: E731
        ip = lambda a: a.ctypes.data_as(_I32)  # noqa: E731
        base = _check(self.lib.pxh_scene_add_bodies(self.h, scene.n, ip(t), fp(dm), fp(po), fp(lv), fp(av), fp(m),
                                                    fp(inn), ip(w)))
        pl, ps = _f32(scene.planes), _i32(scene.plane_scene)
        _check(self.lib.pxh_scene_add_planes(self.h, int(pl.shape[0]), fp(pl), ip(ps)))
        self.n += scene.n
        for members, self_coll in (getattr(scene, "aggregates", None) or []):
            self.add_aggregate(np.asarray(members) + base, self_coll)
        for m in (getattr(scene, "meshes", None) or []):
            if "heightfield" in m:
                hf = m["heightfield"]
                self.add_heightfield(hf["heights"], hf.get("tess"), hf["height_scale"], hf["row_scale"],
                                     hf["column_scale"], m["pose"], int(m["world"]))
            else:
                self.add_triangle_mesh(m["verts"], m["tris"], m["pose"], int(m["world"]))
        j = getattr(scene, "joints", None)
        if j is not None and len(j["type"]):
            first = self.add_joints(j["type"], np.where(j["body0"] >= 0, j["body0"] + base, -1),
                                    np.where(j["body1"] >= 0, j["body1"] + base, -1), j["frame0"], j["frame1"])
            if j.get("params") is not None:
                self.set_joint_params(first + np.arange(len(j["type"])), j["params"])

    def add_triangle_mesh(self, verts, tris, pose=(0, 0, 0, 0, 0, 0, 1), world: int = 0) -> int:
        """pxh_scene_add_triangle_mesh: a static triangle mesh (verts (nv,3),
        tris (nt,3) counter-clockwise about the outward normal) at pose in a
        world; returns the mesh index."""
        v, t, po = _f32(verts).reshape(-1, 3), _i32(tris).reshape(-1, 3), _f32(pose).reshape(7)
        return _check(self.lib.pxh_scene_add_triangle_mesh(self.h, int(v.shape[0]), v.ctypes.data_as(_F),
                                                           int(t.shape[0]), t.ctypes.data_as(_I32),
                                                           po.ctypes.data_as(_F), int(world)))

    def add_heightfield(self, heights, tess=None, height_scale=1.0, row_scale=1.0, column_scale=1.0,
                        pose=(0, 0, 0, 0, 0, 0, 1), world: int = 0) -> int:
        """pxh_scene_add_heightfield: heights (rows, cols) int16 samples, tess
        (rows, cols) tessellation flags or None; returns the mesh index."""
        h = np.ascontiguousarray(heights, dtype=np.int16)
        rows, cols = h.shape
        t = np.ascontiguousarray(tess, dtype=np.uint8) if tess is not None else None
        po = _f32(pose).reshape(7)
        return _check(self.lib.pxh_scene_add_heightfield(
            self.h, int(rows), int(cols), h.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
            t.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)) if t is not None else None, ctypes.c_float(height_scale),
            ctypes.c_float(row_scale), ctypes.c_float(column_scale), po.ctypes.data_as(_F), int(world)))

    def add_aggregate(self, bodies, self_collisions: bool = False) -> int:
        """pxh_scene_add_aggregate (PxAggregate with the bodies): returns its id."""
        b = _i32(bodies)
        return _check(self.lib.pxh_scene_add_aggregate(self.h, int(b.size), b.ctypes.data_as(_I32),
                                                       int(bool(self_collisions))))

    def add_joints(self, kind, body0, body1, frame0, frame1) -> int:
        """pxh_scene_add_joints (PxFixedJointCreate / PxSphericalJointCreate /
        PxRevoluteJointCreate / PxPrismaticJointCreate): host arrays; body -1
        is the world frame; frames (k,7) p xyz, q xyzw. Returns the first
        joint's index."""
        t, b0, b1 = _i32(kind), _i32(body0), _i32(body1)
        f0, f1 = _f32(frame0).reshape(-1, 7), _f32(frame1).reshape(-1, 7)
        return _check(self.lib.pxh_scene_add_joints(self.h, int(t.size), t.ctypes.data_as(_I32),
                                                    b0.ctypes.data_as(_I32), b1.ctypes.data_as(_I32),
                                                    f0.ctypes.data_as(_F), f1.ctypes.data_as(_F)))

    def set_joint_params(self, joints, params):
        """pxh_scene_set_joint_params: params (k, 10) per joint: flags (1
        limit, 2 drive, 4 freespin), lower, upper, restitution, bounce
        threshold, drive velocity, drive force limit, gear ratio, limit
        stiffness, limit damping."""
        jj = _i32(joints)
        p = _f32(params).reshape(-1, 10)
        _check(self.lib.pxh_scene_set_joint_params(self.h, int(jj.size), jj.ctypes.data_as(_I32),
                                                   p.ctypes.data_as(_F)))

    def add_cooked_convex(self, verts, planes, polys, indices, com) -> int:
        """pxh_scene_add_cooked_convex: a hull as the reference cooker wrote it
        (PxConvexMesh getVertices / getPolygonData / getIndexBuffer and the
        local centre of mass); returns the mesh index bodies name in dims[0]."""
        v, pl = _f32(verts).reshape(-1, 3), _f32(planes).reshape(-1, 4)
        po = _i32(polys).reshape(-1, 2)
        ix = np.ascontiguousarray(indices, dtype=np.uint8)
        c = _f32(com).reshape(3)
        out = ctypes.c_int32()
        _check(self.lib.pxh_scene_add_cooked_convex(self.h, int(v.shape[0]), v.ctypes.data_as(_F), int(pl.shape[0]),
                                                    pl.ctypes.data_as(_F), po.ctypes.data_as(_I32),
                                                    ix.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                                    c.ctypes.data_as(_F), ctypes.byref(out)))
        return out.value

    def simulate(self, dt: float):
        _check(self.lib.pxh_scene_simulate(self.h, ctypes.c_float(dt)))

    def fetch_results(self, block: bool = True) -> bool:
        return _check(self.lib.pxh_scene_fetch_results(self.h, int(block))) == 0

    def step(self, dt: float, n: int = 1):
        for _ in range(n):
            self.simulate(dt)
            self.fetch_results(True)

    @property
    def stream(self) -> int:
        return self.lib.pxh_scene_get_stream(self.h) or 0

    def get_rigid_dynamic_data(self, out, kind: int, index=None, n: int | None = None, stream: int = 0):
        """PxDirectGPUAPI::getRigidDynamicData into a device tensor `out`."""
        n = self.n if n is None else n
        _check(self.lib.pxh_scene_get_rigid_dynamic_data(self.h, _dptr(out), _dptr(index) if index is not None else None,
                                                         kind, n, ctypes.c_void_p(stream) if stream else None))

    def set_rigid_dynamic_data(self, data, kind: int, index=None, n: int | None = None, stream: int = 0):
        n = self.n if n is None else n
        _check(self.lib.pxh_scene_set_rigid_dynamic_data(self.h, _dptr(data), _dptr(index) if index is not None else None,
                                                         kind, n, ctypes.c_void_p(stream) if stream else None))

    def bodies(self):
        """Host copies (pose p xyz + q xyzw, linear velocity, angular velocity)."""
        import torch
        dev = torch.device("cuda")
        pose = torch.empty((self.n, 7), dtype=torch.float32, device=dev)
        lv = torch.empty((self.n, 3), dtype=torch.float32, device=dev)
        av = torch.empty((self.n, 3), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        self.get_rigid_dynamic_data(pose, PXH_DATA_GLOBAL_POSE)
        self.get_rigid_dynamic_data(lv, PXH_DATA_LINEAR_VELOCITY)
        self.get_rigid_dynamic_data(av, PXH_DATA_ANGULAR_VELOCITY)
        p = pose.cpu().numpy()
        # PxTransform is (q, p): reorder to p xyz + q xyzw
        return np.concatenate([p[:, 4:7], p[:, 0:4]], 1), lv.cpu().numpy(), av.cpu().numpy()

    def set_forces(self, force=None, torque=None, index=None):
        """Direct GPU API eFORCE / eTORQUE writes from host (k,3) arrays, for the
        bodies in `index` (all bodies when None)."""
        import torch
        dev = torch.device("cuda")
        idx = torch.from_numpy(_i32(index)).to(dev) if index is not None else None
        for kind, arr in ((PXH_WRITE_FORCE, force), (PXH_WRITE_TORQUE, torque)):
            if arr is not None:
                t = torch.from_numpy(_f32(arr)).to(dev)
                torch.cuda.synchronize()
                self.set_rigid_dynamic_data(t, kind, index=idx, n=int(t.shape[0]))

    def accelerations(self):
        """Linear and angular accelerations of the last step (needs
        enable_body_accelerations), host (n,3) arrays."""
        import torch
        dev = torch.device("cuda")
        la = torch.empty((self.n, 3), dtype=torch.float32, device=dev)
        aa = torch.empty((self.n, 3), dtype=torch.float32, device=dev)
        torch.cuda.synchronize()
        self.get_rigid_dynamic_data(la, PXH_READ_LINEAR_ACCELERATION)
        self.get_rigid_dynamic_data(aa, PXH_READ_ANGULAR_ACCELERATION)
        return la.cpu().numpy(), aa.cpu().numpy()

    def set_bodies(self, pose, lin_vel, ang_vel):
        import torch
        dev = torch.device("cuda")
        pose = _f32(pose)
        tr = np.concatenate([pose[:, 3:7], pose[:, 0:3]], 1)
        self.set_rigid_dynamic_data(torch.from_numpy(_f32(tr)).to(dev), PXH_DATA_GLOBAL_POSE)
        self.set_rigid_dynamic_data(torch.from_numpy(_f32(lin_vel)).to(dev), PXH_DATA_LINEAR_VELOCITY)
        self.set_rigid_dynamic_data(torch.from_numpy(_f32(ang_vel)).to(dev), PXH_DATA_ANGULAR_VELOCITY)

    def stats(self) -> dict:
        out = np.zeros(13, np.int64)
        _check(self.lib.pxh_scene_get_stats(self.h, out.ctypes.data_as(_I64), 13))
        keys = ("bodies", "bp_pairs", "plane_pairs", "constraints", "contacts", "colors", "steps", "friction_rows",
                "body_body_constraints", "color_rounds", "body_body_friction_rows", "device_bytes", "solve_launches")
        return dict(zip(keys, map(int, out)))

    def set_tuning(self, layout=None, world_workgroups=None, lds_bodies=None, autotune_steps=None):
        """Launch shapes (pxh_scene_set_tuning); None keeps a setting. Results
        do not depend on them. autotune_steps > 0 tries every candidate shape
        for that many steps each and keeps the fastest."""
        t = self.tuning()
        for i, v in enumerate((layout, world_workgroups, lds_bodies, autotune_steps)):
            if v is not None:
                t[i] = int(v)
        arr = np.array(t, np.int32)
        _check(self.lib.pxh_scene_set_tuning(self.h, arr.ctypes.data_as(_I32), 4))

    def tuning(self) -> list:
        arr = np.zeros(4, np.int32)
        _check(self.lib.pxh_scene_get_tuning(self.h, arr.ctypes.data_as(_I32), 4))
        return [int(x) for x in arr]

    def tuning_log(self) -> list:
        """The last autotuning: [(layout, workgroups, lds bodies, solve ms per step)]."""
        k = _check(self.lib.pxh_scene_get_tuning_log(self.h, None, None, 0))
        shapes = np.zeros((max(k, 1), 3), np.int32)
        ms = np.zeros(max(k, 1), np.float64)
        self.lib.pxh_scene_get_tuning_log(self.h, shapes.ctypes.data_as(_I32),
                                          ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), k)
        return [(int(a), int(b), int(c), float(m)) for (a, b, c), m in zip(shapes[:k], ms[:k])]

    STAGES = ("bounds", "broadphase", "np_pairs", "np_planes", "prep", "solve", "integrate", "unused")

    def set_profiling(self, enable):
        """True: time every stage; False: off; an int: bit mask of the stages
        (index in STAGES) to time."""
        mask = 0xFF if enable is True else (0 if enable is False else int(enable))
        _check(self.lib.pxh_scene_set_profiling(self.h, mask))

    def set_profiling_stride(self, stride: int):
        """Time one in `stride` of the solve launches (counted across steps)."""
        _check(self.lib.pxh_scene_set_profiling_stride(self.h, int(stride)))

    def profile(self) -> dict:
        ms = np.zeros(8, np.float64)
        cnt = np.zeros(8, np.int64)
        _check(self.lib.pxh_scene_get_profile(self.h, ms.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                              cnt.ctypes.data_as(_I64), 8))
        return {name: (float(ms[i]), int(cnt[i])) for i, name in enumerate(self.STAGES) if cnt[i] > 0}

    def contact_data(self):
        """PxDirectGPUAPI::copyContactData: the last step's touching pairs as a
        (k, 32) float32 array; columns 0-2 hold int32 body0, body1 (-1 - plane),
        count; 4-6 the normal; 8-23 four (point, separation); 24-27 normal impulses."""
        import torch
        k = self.stats()["constraints"]
        buf = torch.zeros((max(k, 1), 32), dtype=torch.float32, device="cuda")
        nb = torch.zeros(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        _check(self.lib.pxh_scene_copy_contact_data(self.h, _dptr(buf), _dptr(nb), k, None))
        out = buf.cpu().numpy()[: int(nb.cpu().item())]
        return out

    def touch_events(self):
        """Touch found / lost pairs of the last step (needs report_touch_events):
        two (k,2) int32 arrays of (body a, body b or -1 - plane)."""
        cnt = np.zeros(2, np.int64)
        _check(self.lib.pxh_scene_get_touch_events(self.h, None, 0, None, 0, cnt.ctypes.data_as(_I64)))
        f = np.zeros((max(int(cnt[0]), 1), 2), np.int32)
        lo = np.zeros((max(int(cnt[1]), 1), 2), np.int32)
        _check(self.lib.pxh_scene_get_touch_events(self.h, f.ctypes.data_as(_I32), ctypes.c_int64(int(cnt[0])),
                                                   lo.ctypes.data_as(_I32), ctypes.c_int64(int(cnt[1])),
                                                   cnt.ctypes.data_as(_I64)))
        return f[: cnt[0]], lo[: cnt[1]]

    def sleep_state(self):
        """(wake counter, asleep) of every body (PxRigidDynamic::getWakeCounter,
        isSleeping); needs enable_sleeping."""
        wc = np.zeros(self.n, np.float32)
        asleep = np.zeros(self.n, np.int32)
        _check(self.lib.pxh_scene_get_sleep_state(self.h, wc.ctypes.data_as(_F), asleep.ctypes.data_as(_I32),
                                                  self.n))
        return wc, asleep.astype(bool)

    def wake_bodies(self, bodies):
        """PxRigidDynamic::wakeUp on each listed body."""
        b = _i32(bodies)
        _check(self.lib.pxh_scene_wake_bodies(self.h, b.ctypes.data_as(_I32), int(b.size)))

    def remove_worlds(self, worlds):
        """PxScene::removeActors for every actor of the listed worlds; the
        remaining worlds are renumbered in order and their bodies compacted."""
        w = _i32(worlds)
        _check(self.lib.pxh_scene_remove_worlds(self.h, w.ctypes.data_as(_I32), int(w.size)))
        self.n = sum(self.world_range(i)[1] for i in range(self.world_count()))

    def world_count(self) -> int:
        return _check(self.lib.pxh_scene_world_count(self.h))

    def world_range(self, world: int):
        first, count = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.pxh_scene_world_range(self.h, world, ctypes.byref(first), ctypes.byref(count)))
        return first.value, count.value

    def constraints(self) -> np.ndarray:
        cap = 1 << 23
        out = np.zeros((cap, 4), np.int32)
        n = self.lib.pxh_scene_get_constraints(self.h, out.ctypes.data_as(_I32), ctypes.c_int64(cap))
        _check(int(n))
        return out[: min(n, cap)].copy()

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pxh_scene_release(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cook_convex(points):
    """Convex hull cooking on the host (pxh_convex_hull): a dict of the hull's
    vertices, planes, polygon loops, edges, mass properties at unit density and
    internal extents, or None for a degenerate point cloud."""
    lib = load()
    pts = _f32(points).reshape(-1, 3)
    V = np.zeros((255, 3), np.float32)
    Pl = np.zeros((255, 4), np.float32)
    Po = np.zeros((255, 2), np.int32)
    R = np.zeros(1530, np.int32)
    E = np.zeros((765, 4), np.int32)
    M = np.zeros(13, np.float64)
    S = np.zeros(10, np.float32)
    C = np.zeros(4, np.int32)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    rc = lib.pxh_convex_hull(vp(pts), int(pts.shape[0]), vp(V), vp(Pl), vp(Po), vp(R), vp(E), vp(M), vp(S), vp(C))
    if rc != 0:
        return None
    nv, npol, nr, ne = (int(x) for x in C)
    return dict(verts=V[:nv].copy(), planes=Pl[:npol].copy(), polys=Po[:npol].copy(), vrefs=R[:nr].copy(),
                edges=E[:ne].copy(), poly_counts=Po[:npol, 1].copy(), volume=float(M[0]), com=M[1:4].copy(),
                inertia=M[4:13].reshape(3, 3).copy(), bounds_center=S[0:3].copy(), bounds_extents=S[3:6].copy(),
                internal_radius=float(S[6]), internal_extents=S[7:10].copy())


class Server:
    """Serving loop over one batched scene (pxh_server_*): submitted scenes are
    admitted as worlds while max_worlds and the HBM budget allow, stepped
    together, and retired when their requested steps are done."""

    STATS = ("queued", "active", "active_bodies", "active_bytes", "completed", "steps", "body_steps", "device_bytes")

    def __init__(self, hbm_budget_bytes: int, max_worlds: int, bytes_per_body: int = 8192,
                 desc: SceneDesc | None = None, **kw):
        self.lib = load()
        self.desc = desc if desc is not None else default_desc(**kw)
        h = ctypes.c_void_p()
        self._check(self.lib.pxh_server_create(ctypes.byref(self.desc), ctypes.c_int64(hbm_budget_bytes), max_worlds,
                                               ctypes.c_int64(bytes_per_body), ctypes.byref(h)))
        self.h = h
        self.sizes = {}

    def _check(self, rc):
        if rc < 0:
            raise PxhError(f"pxhip server error {rc}: {self.lib.pxh_server_last_error().decode()}")
        return rc

    def submit(self, scene, steps: int) -> int:
        """Queue one scene (a scenes.SceneArrays of a single world) for `steps` steps."""
        t, dm, po = _i32(scene.shape_type), _f32(scene.dims), _f32(scene.pose)
        lv, av, m, inn = _f32(scene.lin_vel), _f32(scene.ang_vel), _f32(scene.mass), _f32(scene.inertia)
        pl = _f32(scene.planes).reshape(-1, 4)
        fp = lambda a: a.ctypes.data_as(_F)  # noqa: E731
        rid = ctypes.c_int64()
        self._check(self.lib.pxh_server_submit(self.h, scene.n, t.ctypes.data_as(_I32), fp(dm), fp(po), fp(lv), fp(av),
                                               fp(m), fp(inn), int(pl.shape[0]), fp(pl), steps, ctypes.byref(rid)))
        self.sizes[rid.value] = scene.n
        return rid.value

    def step(self, dt: float):
        """One serving step; returns (admitted, completed)."""
        a, c = ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.pxh_server_step(self.h, ctypes.c_float(dt), ctypes.byref(a), ctypes.byref(c)))
        return a.value, c.value

    def result(self, rid: int):
        """(pose (n,7) p xyz + q xyzw, lin (n,3), ang (n,3), times (submit, admit, finish))."""
        n = self.sizes[rid]
        pose = np.zeros((n, 7), np.float32)
        lv = np.zeros((n, 3), np.float32)
        av = np.zeros((n, 3), np.float32)
        tm = np.zeros(3, np.float64)
        self._check(self.lib.pxh_server_result(self.h, ctypes.c_int64(rid), pose.ctypes.data_as(_F),
                                               lv.ctypes.data_as(_F), av.ctypes.data_as(_F),
                                               tm.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        del self.sizes[rid]
        return pose, lv, av, tm

    def finished(self):
        """Ids of finished scenes whose results are not read yet."""
        ids = np.zeros(max(1, len(self.sizes)), np.int64)
        k = self._check(self.lib.pxh_server_finished(self.h, ids.ctypes.data_as(_I64), int(ids.size)))
        return [int(i) for i in ids[:k]]

    def stats(self) -> dict:
        out = np.zeros(len(self.STATS), np.int64)
        self._check(self.lib.pxh_server_stats(self.h, out.ctypes.data_as(_I64), len(self.STATS)))
        return dict(zip(self.STATS, map(int, out)))

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pxh_server_release(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BroadPhase:
    """Standalone GPU broadphase (PxCreateBroadPhase / PxBroadPhase)."""

    def __init__(self, capacity: int = 0):
        self.lib = load()
        h = ctypes.c_void_p()
        _check(self.lib.pxh_bp_create(capacity, ctypes.byref(h)))
        self.h = h

    def update(self, bounds, distances=None, groups=None, stream: int = 0):
        """bounds: (n,6) float32 device tensor; distances (n,), groups (n,) uint32/int32 device tensors."""
        n = int(bounds.shape[0])
        _check(self.lib.pxh_bp_update(self.h, _dptr(bounds), _dptr(groups) if groups is not None else None,
                                      _dptr(distances) if distances is not None else None, n,
                                      ctypes.c_void_p(stream) if stream else None))

    def fetch_results(self):
        """(created, deleted) as (k,2) uint32 numpy arrays, each sorted."""
        import torch  # noqa: F401
        cp, dp = ctypes.c_void_p(), ctypes.c_void_p()
        nc, nd = ctypes.c_int32(), ctypes.c_int32()
        _check(self.lib.pxh_bp_fetch_results(self.h, ctypes.byref(cp), ctypes.byref(nc), ctypes.byref(dp),
                                             ctypes.byref(nd)))
        return _copy_pairs(cp.value, nc.value), _copy_pairs(dp.value, nd.value)

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pxh_bp_release(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bp_update_noalloc(bounds, distances=None, groups=None, max_created=None, max_deleted=None, stream: int = 0):
    """pxh_bp_update_noalloc: the standalone broadphase update without any
    device allocation in the library, diffed against the previous call on
    this GPU. bounds (n,6) float32 device tensor, distances (n,) and groups
    (n,) device tensors. Returns (created, deleted) as sorted (k,2) uint32
    numpy arrays."""
    import torch
    lib = load()
    n = int(bounds.shape[0])
    mc = max(1, max_created if max_created is not None else 4 * max(n, 1))
    md = max(1, max_deleted if max_deleted is not None else PXH_BP_NOALLOC_MAX_PAIRS)
    cr = torch.empty((mc, 2), dtype=torch.int32, device=bounds.device)
    dl = torch.empty((md, 2), dtype=torch.int32, device=bounds.device)
    nc, nd = ctypes.c_int32(), ctypes.c_int32()
    _check(lib.pxh_bp_update_noalloc(_dptr(bounds), _dptr(groups) if groups is not None else None,
                                     _dptr(distances) if distances is not None else None, n, _dptr(cr), mc,
                                     ctypes.byref(nc), _dptr(dl), md, ctypes.byref(nd),
                                     ctypes.c_void_p(stream) if stream else None))
    return (cr[: nc.value].cpu().numpy().view(np.uint32).reshape(-1, 2),
            dl[: nd.value].cpu().numpy().view(np.uint32).reshape(-1, 2))


def _copy_pairs(ptr, n):
    out = np.zeros((n, 2), np.uint32)
    if n > 0:
        import torch
        torch.cuda.synchronize()
        hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime already in the process
        hip.hipMemcpy.restype = ctypes.c_int
        rc = hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ptr), ctypes.c_size_t(out.nbytes), 2)
        if rc != 0:
            raise PxhError(f"hipMemcpy failed ({rc})")
    return out


def generate_contacts(type0, geom0, pose0, type1, geom1, pose1, contact_distance=0.04, tolerance_length=1.0):
    """Batched PxGenerateContacts on the device. Host numpy inputs; returns
    (counts (n,), contacts (n,4,8)) as numpy arrays."""
    import torch
    dev = torch.device("cuda")
    n = len(type0)
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(dev)  # noqa: E731
    T0, T1 = t(type0, np.int32), t(type1, np.int32)
    G0, G1 = t(np.asarray(geom0).reshape(n, 4), np.float32), t(np.asarray(geom1).reshape(n, 4), np.float32)
    P0, P1 = t(np.asarray(pose0).reshape(n, 7), np.float32), t(np.asarray(pose1).reshape(n, 7), np.float32)
    out = torch.zeros((n, 4, 8), dtype=torch.float32, device=dev)
    cnt = torch.zeros(n, dtype=torch.int32, device=dev)
    torch.cuda.synchronize()
    _check(load().pxh_generate_contacts(n, _dptr(T0), _dptr(G0), _dptr(P0), _dptr(T1), _dptr(G1), _dptr(P1),
                                        ctypes.c_float(contact_distance), ctypes.c_float(tolerance_length),
                                        _dptr(out), _dptr(cnt), None))
    return cnt.cpu().numpy(), out.cpu().numpy()
