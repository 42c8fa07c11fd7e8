"""Pins the oracle's static triangle meshes (oracle/pxo.c: tm_body, tm_sphere,
tm_capsule, tm_poly) before they check the GPU: contacts of spheres,
capsules, boxes and hulls resting on a terrain mesh against the reference
SDK's PxGenerateContacts on the same cooked mesh (PxCreateTriangleMesh,
oracle/_ref), bodies dropped on the terrain against the reference scene over
240 steps, and closed forms on flat meshes."""
import ctypes

import numpy as np
import pytest

import oracle_libs as L
from physx_amd import scenes as S

DT = 1.0 / 60.0
IDENT = np.array([0, 0, 0, 0, 0, 0, 1], np.float32)
need_ref = pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")


def mesh_height(V, T, x, z):
    """height of the mesh under (x, z) (the triangle whose xz projection holds it)"""
    a, b, c = V[T[:, 0]], V[T[:, 1]], V[T[:, 2]]
    v0 = np.stack([b[:, 0] - a[:, 0], b[:, 2] - a[:, 2]], 1).astype(np.float64)
    v1 = np.stack([c[:, 0] - a[:, 0], c[:, 2] - a[:, 2]], 1).astype(np.float64)
    v2 = np.stack([x - a[:, 0], z - a[:, 2]], 1).astype(np.float64)
    den = v0[:, 0] * v1[:, 1] - v1[:, 0] * v0[:, 1]
    u = (v2[:, 0] * v1[:, 1] - v1[:, 0] * v2[:, 1]) / den
    v = (v0[:, 0] * v2[:, 1] - v2[:, 0] * v0[:, 1]) / den
    k = np.where((u >= -1e-9) & (v >= -1e-9) & (u + v <= 1 + 1e-9))[0][0]
    return float(a[k, 1] + u[k] * (b[k, 1] - a[k, 1]) + v[k] * (c[k, 1] - a[k, 1]))


def _lowest(base, i, q):
    """depth of body i's lowest point below its centre for orientation q"""
    t, d = int(base.shape_type[i]), base.dims[i]
    x, y, z, w = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    if t == S.SPHERE:
        return d[0]
    if t == S.CAPSULE:
        return d[0] + abs(R[1, 0]) * d[1]
    if t == S.BOX:
        return abs(R[1]) @ d[:3]
    pts = base.hull_points.reshape(len(base.hull_counts), -1, 3)[int(d[0])]
    return -(pts @ R.T)[:, 1].min()


def resting_on_terrain(n=128, seed=3):
    """mixed_pile's bodies (a quarter of each kind), one per world, each over
    a random point of the terrain with its lowest point 2 cm below to 1 cm
    above the surface there; one terrain mesh per world"""
    base = S.mixed_pile(n_bodies=n, per_side=4, seed=5)
    V, T = S.terrain_mesh(16, 16.0, 0.4, 9)
    rng = np.random.default_rng(seed)
    pose = base.pose.copy()
    for i in range(n):
        x, z = rng.uniform(-6, 6, 2)
        pose[i, 0], pose[i, 2] = x, z
        pose[i, 1] = mesh_height(V, T, x, z) + _lowest(base, i, pose[i, 3:]) + rng.uniform(-0.02, 0.01)
    sc = S.Scene(base.shape_type, base.dims, pose, np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32),
                 base.mass, base.inertia, np.arange(n, dtype=np.int32), np.zeros((0, 4), np.float32),
                 np.zeros(0, np.int32), n, "resting_on_terrain", {}, base.hull_points, base.hull_counts)
    sc.meshes = [{"verts": V, "tris": T, "pose": IDENT, "world": w} for w in range(n)]
    return sc


def oracle_contacts(sc):
    """per body: (k, 8) contacts (point, separation, normal) the oracle's first
    step generates"""
    o = L.OracleWorld(sc)
    o.step(DT)
    rec = o.contact_data()
    ids = rec[:, :3].copy().view(np.int32)
    out = {}
    for r, (a, b, k) in zip(rec, ids):
        c = np.zeros((k, 8), np.float32)
        c[:, :4] = r[8:8 + 4 * k].reshape(k, 4)
        c[:, 4:7] = r[4:7]
        out.setdefault(int(a), []).append(c)
    return {i: np.concatenate(v) for i, v in out.items()}


def reference_contacts(sc, i):
    lib = L.ref()
    V, T = sc.meshes[0]["verts"], sc.meshes[0]["tris"]
    mid = L.ref_trimesh(V, T)
    dims = L.f32(sc.dims[i]).copy()
    if sc.shape_type[i] == S.CONVEX:
        dims[0] = L.ref_cook(L._hull_clouds(sc)[int(sc.dims[i, 0])])[0]
    out = np.zeros((64, 8), np.float32)
    g1 = np.array([mid, 0, 0, 0], np.float32)
    k = lib.pxref_generate_contacts(int(sc.shape_type[i]), L.fp(dims), L.fp(L.f32(sc.pose[i])), 8, L.fp(g1),
                                    L.fp(IDENT), ctypes.c_float(0.04), ctypes.c_float(1.0), L.fp(out), 64)
    return out[:k]


# per class: share of bodies whose deepest contact separation is within 1e-5
# m of the reference's, and share whose deepest contact normal is within 5
# degrees (ties between equally deep contacts of different patches decide
# which normal is the deepest's)
BOUNDS = {S.SPHERE: (1.0, 1.0), S.CAPSULE: (0.95, 0.95), S.BOX: (0.95, 0.95), S.CONVEX: (0.9, 0.9)}
# largest deepest-separation difference per class: a box or hull a few
# centimetres into a slope, where the reference's separating-axis search
# settles on one of its own faces while this build keeps the triangle normal
MAX_SEP = {S.SPHERE: 1e-3, S.CAPSULE: 1e-3, S.BOX: 5e-3, S.CONVEX: 5e-3}


@need_ref
def test_terrain_contacts_vs_reference():
    """Every body touching the terrain in the reference touches it in the
    oracle and the other way round; every deepest separation within MAX_SEP
    of the reference's, and per class the shares of BOUNDS within 1e-5 m and
    5 degrees. The sphere path restates the reference's own (closest point,
    face contacts first, deferred edge and vertex contacts with feature
    caches)."""
    sc = resting_on_terrain()
    ora = oracle_contacts(sc)
    per = {}
    for i in range(sc.n):
        ref = reference_contacts(sc, i)
        oc = ora.get(i, np.zeros((0, 8), np.float32))
        assert (len(ref) > 0) == (len(oc) > 0), (i, len(ref), len(oc))
        if len(ref) == 0:
            continue
        rd, od = ref[np.argmin(ref[:, 3])], oc[np.argmin(oc[:, 3])]
        sep = abs(float(rd[3] - od[3]))
        ang = float(np.degrees(np.arccos(np.clip(np.dot(rd[4:7], od[4:7]), -1.0, 1.0))))
        assert sep < MAX_SEP[int(sc.shape_type[i])], (i, int(sc.shape_type[i]), sep)
        per.setdefault(int(sc.shape_type[i]), []).append((sep, ang))
    for t, rows in per.items():
        seps = np.array([r[0] for r in rows])
        angs = np.array([r[1] for r in rows])
        share_sep, share_ang = BOUNDS[t]
        print(t, len(rows), (seps < 1e-5).mean(), (angs < 5.0).mean())
        assert (seps < 1e-5).mean() >= share_sep, (t, (seps < 1e-5).mean())
        assert (angs < 5.0).mean() >= share_ang, (t, (angs < 5.0).mean())


@need_ref
def test_terrain_drop_long_run_vs_reference():
    """36 bodies of every kind dropped on the terrain, 240 steps side by side
    with the reference scene (the same mesh as a static PxTriangleMesh
    actor): median height difference within 3 mm, 80% of the bodies within
    2 cm; no body sinks more than 3 cm into the terrain in either."""
    sc = S.terrain_drop(n_scenes=1, per_side=6, layers=1, seed=9)
    o = L.OracleWorld(sc)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1)
    V, T = sc.meshes[0]["verts"], sc.meshes[0]["tris"]
    for _ in range(240):
        o.step(DT)
        r.step(DT)
    op, rp = o.bodies()[0], r.bodies()[0]
    d = np.abs(op[:, 1] - rp[:, 1])
    print("median %.4f 80th %.4f max %.4f" % (np.median(d), np.percentile(d, 80), d.max()))
    assert np.median(d) < 3e-3 and np.percentile(d, 80) < 2e-2
    for p in (op, rp):
        for i in range(sc.n):
            if abs(p[i, 0]) < 7.5 and abs(p[i, 2]) < 7.5:
                assert p[i, 1] > mesh_height(V, T, p[i, 0], p[i, 2]) - 0.03, (i, p[i, :3])


def _flat_grid(n=8, size=8.0, h=0.0):
    V, T = S.terrain_mesh(n, size, 0.0, 1)
    V = V.copy()
    V[:, 1] = h
    return V, T


def test_sphere_rolls_across_flat_mesh_edges():
    """Closed form: a sphere rolling at 2 m/s over a flat triangulated floor
    crosses many triangle edges and vertices without a bump: its height stays
    at the resting height (within 0.1 mm) and it keeps rolling straight."""
    V, T = _flat_grid()
    st = np.array([S.SPHERE], np.int32)
    dims = np.array([[0.25, 0, 0, 0]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    sc = S.Scene(st, dims, np.array([[-2.5, 0.25, -0.3, 0, 0, 0, 1]], np.float32), np.array([[2.0, 0, 0.4]], np.float32),
                 np.zeros((1, 3), np.float32), m, inn, np.zeros(1, np.int32), np.zeros((0, 4), np.float32),
                 np.zeros(0, np.int32), 1, "roll", {})
    sc.meshes = [{"verts": V, "tris": T, "pose": IDENT, "world": 0}]
    o = L.OracleWorld(sc)
    ys, zs = [], []
    for _ in range(120):
        o.step(DT)
        p = o.bodies()[0]
        ys.append(p[0, 1])
        zs.append(p[0, 2])
    ys = np.array(ys[10:])
    assert np.abs(ys - ys.mean()).max() < 1e-4 and abs(ys.mean() - 0.25) < 2e-3, (ys.min(), ys.max())
    assert abs(o.bodies()[1][0, 2] / o.bodies()[1][0, 0] - 0.2) < 0.02


def test_box_rests_on_flat_mesh_like_a_plane():
    """Closed form: a box resting across several triangles of a flat mesh
    behaves as on the plane y=0: over 60 steps its height stays within
    0.1 mm of the plane run's and it stays where it is (within 0.1 mm;
    its velocity jitters below 1 cm/s as the four kept contacts change)."""
    V, T = _flat_grid()
    st = np.array([S.BOX], np.int32)
    dims = np.array([[0.6, 0.3, 0.45, 0]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    q = np.array([0, np.sin(0.3), 0, np.cos(0.3)], np.float32)
    pose = np.array([[0.3, 0.3, 0.2, *q]], np.float32)
    base = (st, dims, pose, np.zeros((1, 3), np.float32), np.zeros((1, 3), np.float32), m, inn,
            np.zeros(1, np.int32))
    on_mesh = S.Scene(*base, np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "mesh", {})
    on_mesh.meshes = [{"verts": V, "tris": T, "pose": IDENT, "world": 0}]
    on_plane = S.Scene(*base, np.array([[0, 1, 0, 0]], np.float32), np.zeros(1, np.int32), 1, "plane", {})
    a, b = L.OracleWorld(on_mesh), L.OracleWorld(on_plane)
    for _ in range(60):
        a.step(DT)
        b.step(DT)
    pa, la, _ = a.bodies()
    pb, _, _ = b.bodies()
    assert abs(pa[0, 1] - pb[0, 1]) < 1e-4 and np.abs(la).max() < 1e-2, (pa, pb, la)
    assert np.abs(pa[0, [0, 2]] - pose[0, [0, 2]]).max() < 1e-4, pa


def test_mesh_is_one_sided():
    """A body whose centre lies behind the mesh passes up through it: a box
    launched upward from below a flat mesh crosses it untouched, while the
    same box falling from above lands on it."""
    V, T = _flat_grid()
    st = np.array([S.BOX, S.BOX], np.int32)
    dims = np.array([[0.2, 0.2, 0.2, 0]] * 2, np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    pose = np.array([[0, -0.5, 0, 0, 0, 0, 1], [1.5, 1.0, 1.5, 0, 0, 0, 1]], np.float32)
    lv = np.array([[0, 6.0, 0], [0, 0, 0]], np.float32)
    sc = S.Scene(st, dims, pose, lv, np.zeros((2, 3), np.float32), m, inn, np.zeros(2, np.int32),
                 np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "one-sided", {})
    sc.meshes = [{"verts": V, "tris": T, "pose": IDENT, "world": 0}]
    o = L.OracleWorld(sc)
    for _ in range(30):
        o.step(DT)
    p = o.bodies()[0]
    assert p[0, 1] > 0.3, p[0]          # passed up through the mesh
    assert abs(p[1, 1] - 0.2) < 5e-3, p[1]   # landed on it


def resting_on_heightfield(n=128, seed=4):
    """resting_on_terrain's bodies over a heightfield (terrain_heights,
    0.01 m height steps, 1 m cells, centred on the origin)"""
    base = S.mixed_pile(n_bodies=n, per_side=4, seed=5)
    heights, tess = S.terrain_heights(17, 17, 0.4, 9)
    hf = {"heights": heights, "tess": tess, "height_scale": 0.01, "row_scale": 1.0, "column_scale": 1.0}
    V, T = S.heightfield_mesh(heights, tess, 0.01, 1.0, 1.0)
    pose_hf = np.array([-8, 0, -8, 0, 0, 0, 1], np.float32)
    Vw = V + pose_hf[:3]
    rng = np.random.default_rng(seed)
    pose = base.pose.copy()
    for i in range(n):
        x, z = rng.uniform(-6, 6, 2)
        pose[i, 0], pose[i, 2] = x, z
        pose[i, 1] = mesh_height(Vw, T, x, z) + _lowest(base, i, pose[i, 3:]) + rng.uniform(-0.02, 0.01)
    sc = S.Scene(base.shape_type, base.dims, pose, np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32),
                 base.mass, base.inertia, np.arange(n, dtype=np.int32), np.zeros((0, 4), np.float32),
                 np.zeros(0, np.int32), n, "resting_on_heightfield", {}, base.hull_points, base.hull_counts)
    sc.meshes = [{"verts": V, "tris": T, "pose": pose_hf, "world": w, "heightfield": hf} for w in range(n)]
    return sc


@need_ref
def test_heightfield_contacts_vs_reference():
    """Bodies resting on a heightfield: the oracle (the heightfield as the
    triangles of its cells) against PxGenerateContacts with the reference's
    PxHeightFieldGeometry: touching agreement for every body, deepest
    separation within MAX_SEP for all and within 1e-5 m for the BOUNDS shares,
    deepest normals within 5 degrees for the same shares."""
    sc = resting_on_heightfield()
    ora = oracle_contacts(sc)
    lib = L.ref()
    g1 = L.ref_heightfield_geom(sc.meshes[0]["heightfield"])
    ph = L.f32(sc.meshes[0]["pose"])
    per = {}
    for i in range(sc.n):
        dims = L.f32(sc.dims[i]).copy()
        if sc.shape_type[i] == S.CONVEX:
            dims[0] = L.ref_cook(L._hull_clouds(sc)[int(sc.dims[i, 0])])[0]
        out = np.zeros((64, 8), np.float32)
        k = lib.pxref_generate_contacts(int(sc.shape_type[i]), L.fp(dims), L.fp(L.f32(sc.pose[i])), 9, L.fp(g1),
                                        L.fp(ph), ctypes.c_float(0.04), ctypes.c_float(1.0), L.fp(out), 64)
        ref = out[:k]
        oc = ora.get(i, np.zeros((0, 8), np.float32))
        assert (len(ref) > 0) == (len(oc) > 0), (i, len(ref), len(oc))
        if len(ref) == 0:
            continue
        rd, od = ref[np.argmin(ref[:, 3])], oc[np.argmin(oc[:, 3])]
        sep = abs(float(rd[3] - od[3]))
        ang = float(np.degrees(np.arccos(np.clip(np.dot(rd[4:7], od[4:7]), -1.0, 1.0))))
        assert sep < MAX_SEP[int(sc.shape_type[i])], (i, int(sc.shape_type[i]), sep)
        per.setdefault(int(sc.shape_type[i]), []).append((sep, ang))
    for t, rows in per.items():
        seps = np.array([r[0] for r in rows])
        angs = np.array([r[1] for r in rows])
        share_sep, share_ang = BOUNDS[t]
        print(t, len(rows), (seps < 1e-5).mean(), (angs < 5.0).mean())
        assert (seps < 1e-5).mean() >= share_sep, (t, (seps < 1e-5).mean())
        assert (angs < 5.0).mean() >= share_ang, (t, (angs < 5.0).mean())


@need_ref
def test_heightfield_drop_long_run_vs_reference():
    """36 bodies dropped on a heightfield, 240 steps beside the reference
    scene with a PxHeightFieldGeometry actor: median height difference
    within 5 mm, 80% within 2 cm."""
    sc = S.heightfield_drop(n_scenes=1, per_side=6, seed=9)
    o = L.OracleWorld(sc)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1)
    for _ in range(240):
        o.step(DT)
        r.step(DT)
    d = np.abs(o.bodies()[0][:, 1] - r.bodies()[0][:, 1])
    print("median %.4f 80th %.4f max %.4f" % (np.median(d), np.percentile(d, 80), d.max()))
    assert np.median(d) < 5e-3 and np.percentile(d, 80) < 2e-2
