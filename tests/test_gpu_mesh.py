"""GPU static triangle meshes (pxh_scene_add_triangle_mesh: tree midphase and
the body-mesh contacts of pxh_mesh.h) against the oracle, which
tests/test_cpu_mesh.py pins to the reference SDK: the contacts of 128 bodies
resting on a terrain, one step from a mid-run state of bodies dropped on it,
the run side by side, a world's bits alone and in a batch and after world
removal, sleeping on a mesh, and the triangle-count limit."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L
from physx_amd import scenes as S
from test_cpu_mesh import IDENT, _flat_grid, resting_on_terrain

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def _by_body(rec):
    out = {}
    ids = rec[:, :3].copy().view(np.int32)
    for r, (a, b, k) in zip(rec, ids):
        c = np.zeros((k, 8), np.float32)
        c[:, :4] = r[8:8 + 4 * k].reshape(k, 4)
        c[:, 4:7] = r[4:7]
        out.setdefault(int(a), []).append(c)
    return {i: np.concatenate(v) for i, v in out.items()}


def test_mesh_contacts_match_oracle(cuda):
    """128 bodies of every kind resting on a terrain mesh (one world each):
    the GPU's contact sets equal the oracle's (points and separations within
    1e-4, normals within 1e-3) for every body."""
    sc = resting_on_terrain()
    g = P.Scene(sc)
    g.step(DT)
    o = L.OracleWorld(sc)
    o.step(DT)
    gs, os_ = g.stats(), o.stats()
    assert gs["constraints"] == os_["constraints"] > 100, (gs, os_)
    gc, oc = _by_body(g.contact_data()), _by_body(o.contact_data())
    assert sorted(gc) == sorted(oc)
    bad = [i for i in gc if L.same_contacts(L.dedupe(gc[i]), L.dedupe(oc[i]))]
    assert not bad, bad[:8]


def _close(g, o, tol):
    gp, gl, ga = g
    op, ol, oa = o
    lt = tol * np.maximum(np.abs(ol).max(1), 9.81 * DT)
    at = tol * np.maximum(np.abs(oa).max(1), 1.0)
    bl = np.where(np.abs(gl - ol).max(1) > lt)[0]
    ba = np.where(np.abs(ga - oa).max(1) > at)[0]
    assert bl.size == 0 and ba.size == 0, (bl[:5], ba[:5])


def test_terrain_drop_one_step_from_mid_run_state(cuda):
    """Four worlds of bodies dropped on the terrain: the GPU runs 60 steps,
    then the GPU and the oracle each take one step from that state: the same
    constraints, velocities within 1e-5 of max(|v|, g dt)."""
    sc = S.terrain_drop(n_scenes=4, per_side=6, layers=1, seed=9)
    g = P.Scene(sc)
    g.step(DT, 60)
    pose, lv, av = g.bodies()
    g.close()
    st = S.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes, sc.plane_scene,
                 sc.n_scenes, sc.name, {}, sc.hull_points, sc.hull_counts, None, None, sc.meshes)
    g = P.Scene(st)
    o = L.OracleWorld(st)
    g.step(DT)
    o.step(DT)
    assert g.stats()["constraints"] == o.stats()["constraints"]
    gc, oc = g.constraints(), o.constraints()
    assert sorted(map(tuple, gc[:, :2].tolist())) == sorted(map(tuple, oc[:, :2].tolist()))
    _close(g.bodies(), o.bodies(), 1e-5)


def test_terrain_drop_tracks_oracle(cuda):
    """The drop stepped side by side on the GPU and in the oracle for 60
    steps: every step's constraint count equal; at the end 90% of the bodies
    within 0.1 mm of the oracle's and all within 1 cm. Bodies tumbling down
    slopes make the run chaotic: in the oracle alone, moving every body by
    1e-6 m gives a 6 mm largest difference by step 60 while 90% stay within
    3.4e-5 m, the spread the GPU's rounding (sincosf, the symmetric
    square-root inverse inertia) meets too."""
    sc = S.terrain_drop(n_scenes=2, per_side=6, layers=1, seed=9)
    g = P.Scene(sc)
    o = L.OracleWorld(sc)
    for s in range(60):
        g.step(DT)
        o.step(DT)
        assert g.stats()["constraints"] == o.stats()["constraints"], s
    d = np.abs(g.bodies()[0][:, :3] - o.bodies()[0][:, :3]).max(1)
    print(f"position difference after 60 steps: largest {d.max():.2e}, 90th percentile {np.percentile(d, 90):.2e}")
    assert np.percentile(d, 90) < 1e-4 and d.max() < 1e-2


def test_mesh_world_alone_in_batch_and_after_removal(cuda):
    """A terrain world gives the same bits stepped alone and inside a batch of
    six, and after another world (and its mesh) is removed."""
    sc = S.terrain_drop(n_scenes=6, per_side=4, layers=1, seed=9)
    alone = P.Scene(sc.subset(4))
    batch = P.Scene(sc)
    for _ in range(30):
        alone.step(DT)
        batch.step(DT)
    f, c = batch.world_range(4)
    for x, y in zip(alone.bodies(), batch.bodies()):
        assert np.array_equal(x, y[f:f + c])
    batch.remove_worlds([1])
    f2, c2 = batch.world_range(3)
    for _ in range(30):
        alone.step(DT)
        batch.step(DT)
    for x, y in zip(alone.bodies(), batch.bodies()):
        assert np.array_equal(x, y[f2:f2 + c2])


def test_mesh_sleeping_matches_oracle(cuda):
    """A box and a sphere coming to rest on a flat mesh with sleeping on:
    every step's sleeping flags and wake counters equal the oracle's (the
    mesh counts among each body's interactions)."""
    V, T = _flat_grid()
    st = np.array([S.BOX, S.SPHERE], np.int32)
    dims = np.array([[0.4, 0.3, 0.5, 0], [0.3, 0, 0, 0]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    pose = np.array([[0.3, 0.6, 0.2, 0, 0, 0, 1], [-1.7, 0.5, 1.3, 0, 0, 0, 1]], np.float32)
    sc = S.Scene(st, dims, pose, np.zeros((2, 3), np.float32), np.zeros((2, 3), np.float32), m, inn,
                 np.zeros(2, np.int32), np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "rest", {})
    sc.meshes = [{"verts": V, "tris": T, "pose": IDENT, "world": 0}]
    g = P.Scene(sc, enable_sleeping=1, sleep_threshold=0.005)
    o = L.OracleWorld(sc, sleep_threshold=0.005)
    for s in range(90):
        g.step(DT)
        o.step(DT)
        gw, ga = g.sleep_state()
        ow, oa = o.sleep_state()
        assert np.array_equal(ga, oa) and np.abs(gw - ow).max() <= 1e-5, (s, gw, ow, ga, oa)
    assert ga.all()


def test_too_many_triangles_under_a_body(cuda):
    """A 2 m box over a mesh of 5 cm triangles meets more than 128 of them:
    the step reports PXH_ERR_CAPACITY."""
    V, T = S.terrain_mesh(80, 4.0, 0.0, 1)
    st = np.array([S.BOX], np.int32)
    dims = np.array([[1.0, 0.2, 1.0, 0]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    sc = S.Scene(st, dims, np.array([[0, 0.19, 0, 0, 0, 0, 1]], np.float32), np.zeros((1, 3), np.float32),
                 np.zeros((1, 3), np.float32), m, inn, np.zeros(1, np.int32), np.zeros((0, 4), np.float32),
                 np.zeros(0, np.int32), 1, "fine", {})
    sc.meshes = [{"verts": V, "tris": T, "pose": IDENT, "world": 0}]
    g = P.Scene(sc)
    with pytest.raises(P.PxhError, match="128 mesh triangles"):
        g.simulate(DT)


def test_heightfield_equals_its_triangle_mesh_and_oracle(cuda):
    """A heightfield added through pxh_scene_add_heightfield steps bit for bit
    like the same cells added as a triangle mesh (the library builds the
    triangles itself), and one step of its bodies matches the oracle's
    (constraint count, velocities within 1e-5)."""
    sc = S.heightfield_drop(n_scenes=2, per_side=6, seed=9)
    as_mesh = S.Scene(**{**sc.__dict__})
    as_mesh.meshes = [{k: v for k, v in m.items() if k != "heightfield"} for m in sc.meshes]
    a = P.Scene(sc)
    b = P.Scene(as_mesh)
    for _ in range(40):
        a.step(DT)
        b.step(DT)
    for x, y in zip(a.bodies(), b.bodies()):
        assert np.array_equal(x, y)
    pose, lv, av = a.bodies()
    st = S.Scene(**{**sc.__dict__})
    st.pose, st.lin_vel, st.ang_vel = pose, lv, av
    g = P.Scene(st)
    o = L.OracleWorld(st)
    g.step(DT)
    o.step(DT)
    assert g.stats()["constraints"] == o.stats()["constraints"] > 30
    _close(g.bodies(), o.bodies(), 1e-5)
