"""Pins the C oracle (oracle/pxo.c) before it is trusted as the GPU checker:
against golden vectors written by the reference PhysX 5.6.1 CPU SDK
(tests/golden/make_golden.py), against closed forms and brute force, and -
when oracle/_ref/libpxref.so is built - against the live reference on fresh
seeded inputs."""
import math

import numpy as np
import pytest

import oracle_libs as L
from physx_amd import scenes as S

GOLDEN = L.os.path.join(L.ROOT, "tests", "golden")
KINDS = ("ss", "sp", "sb", "pb", "bb")


def golden(name):
    return np.load(L.os.path.join(GOLDEN, name))


@pytest.mark.parametrize("kind", KINDS)
def test_contacts_vs_reference_golden(kind):
    """Contacts within 1e-4 m (points, separations) and 1e-3 (normal).
    Box-box pairs that face clipping leaves empty go through the reference's
    GJK/EPA fallback, restated here by a SAT/closest-edge computation; up to 3%
    of box-box pairs may differ there (parity of that sub-case is partial)."""
    g = golden("contacts.npz")
    sel = np.nonzero(g["kind"] == KINDS.index(kind))[0]
    bad = []
    for i in sel:
        got = L.contacts("pxo", int(g["t0"][i]), g["g0"][i], g["p0"][i], int(g["t1"][i]), g["g1"][i], g["p1"][i])
        n = int(g["count"][i])
        ref = g["contacts"][i][:n]
        if got.shape[0] != n:
            bad.append((i, "count", got.shape[0], n))
            continue
        if n and (np.abs(got[:, :4] - ref[:, :4]).max() > 1e-4 or np.abs(got[:, 4:7] - ref[:, 4:7]).max() > 1e-3):
            bad.append((i, "value", np.abs(got - ref).max()))
    allowed = 0 if kind != "bb" else int(0.03 * len(sel))
    assert len(bad) <= allowed, bad[:5]


def test_broadphase_vs_reference_golden():
    g = golden("broadphase.npz")
    got = L.bp_pairs_pxo(g["bounds"], g["distances"])
    assert np.array_equal(got, g["pairs"])


def test_broadphase_brute_force():
    rng = np.random.default_rng(5)
    n = 600
    c = rng.uniform(0, 6, (n, 3)).astype(np.float32)
    h = rng.uniform(0.05, 0.3, (n, 3)).astype(np.float32)
    b = np.concatenate([c - h, c + h], 1)
    d = rng.uniform(0, 0.05, n).astype(np.float32)
    # boxes touching exactly: inclusive test must report them
    b[1, :3] = b[0, 3:] + 2 * d[0]
    b[1, :3] -= d[0] + d[1]
    b[1, 3:] = b[1, :3] + 0.1
    lo = b[:, :3] - d[:, None]
    hi = b[:, 3:] + d[:, None]
    ov = np.all((lo[:, None] <= hi[None]) & (lo[None] <= hi[:, None]), axis=2)
    iu = np.triu_indices(n, 1)
    expect = np.stack([iu[0][ov[iu]], iu[1][ov[iu]]], 1)
    got = L.bp_pairs_pxo(b, d)
    assert np.array_equal(got, expect)
    world = (np.arange(n) % 3).astype(np.int32)
    got_w = L.bp_pairs_pxo(b, d, world)
    assert np.array_equal(got_w, expect[world[expect[:, 0]] == world[expect[:, 1]]])


def _single(name_row, g, i):
    t = int(g["type"][i])
    dims = g["dims"][i][None].astype(np.float32)
    m, inertia = S.mass_props(np.array([t]), dims, 1000.0)
    return S.Scene(np.array([t], np.int32), dims, g["pose0"][i][None], g["lv0"][i][None], g["av0"][i][None], m,
                   inertia, np.zeros(1, np.int32), np.array([[0, 1, 0, 0]], np.float32), np.zeros(1, np.int32))


def test_one_step_vs_reference_golden():
    """Single bodies on a plane, where the result does not depend on the
    constraint order: velocities within 1e-4 of max(|v|, g*dt) (linear) and of
    max(|w|, 1 rad/s) (angular) after one TGS step."""
    g = golden("step.npz")
    dt = float(g["dt"])
    for i in range(len(g["name"])):
        sc = _single(g["name"][i], g, i)
        o = L.OracleWorld(sc, int(g["pos_iters"]), int(g["vel_iters"]))
        o.step(dt)
        p, lv, av = o.bodies()
        lin_tol = 1e-4 * max(np.linalg.norm(g["lv1"][i]), 9.81 * dt)
        ang_tol = 1e-4 * max(np.linalg.norm(g["av1"][i]), 1.0)
        assert np.abs(lv[0] - g["lv1"][i]).max() <= lin_tol, (g["name"][i], lv[0], g["lv1"][i])
        assert np.abs(av[0] - g["av1"][i]).max() <= ang_tol, (g["name"][i], av[0], g["av1"][i])
        assert np.abs(p[0, :3] - g["pose1"][i][:3]).max() <= 1e-5, g["name"][i]


def test_sphere_sphere_closed_form():
    rng = np.random.default_rng(3)
    for _ in range(200):
        r0, r1 = rng.uniform(0.1, 0.3, 2)
        p0, p1 = rng.uniform(-0.3, 0.3, 3), rng.uniform(-0.3, 0.3, 3)
        d = np.linalg.norm(p0 - p1)
        c = L.contacts("pxo", 0, [r0], np.r_[p0, 0, 0, 0, 1], 0, [r1], np.r_[p1, 0, 0, 0, 1])
        if d >= r0 + r1 + 0.04:
            assert c.shape[0] == 0
            continue
        n = (p0 - p1) / d
        assert c.shape[0] == 1
        assert np.allclose(c[0, 4:7], n, atol=1e-5)
        assert np.allclose(c[0, :3], p1 + n * r1, atol=1e-5)
        assert abs(c[0, 3] - (d - r0 - r1)) < 1e-5


def test_box_face_on_plane_closed_form():
    """A box resting face-down touches the plane at its 4 lower corners with
    equal separation."""
    for h, y in (((0.3, 0.2, 0.25), 0.19), ((0.5, 0.1, 0.4), 0.12)):
        c = L.contacts("pxo", 1, [0], [0, 0, 0, 0, 0, math.sqrt(.5), math.sqrt(.5)], 3, h, [1, y, -2, 0, 0, 0, 1])
        assert c.shape[0] == 4
        assert np.allclose(c[:, 3], y - h[1], atol=1e-6)
        corners = sorted((round(float(x), 4), round(float(z), 4)) for x, z in c[:, [0, 2]])
        expect = sorted((round(1 + sx * h[0], 4), round(-2 + sz * h[2], 4)) for sx in (1, -1) for sz in (1, -1))
        assert corners == expect
        assert np.allclose(c[:, 4:7], [0, -1, 0], atol=1e-6)


def test_free_flight_closed_forms():
    """No contacts: v' = (v + g dt) (linear damping 0), w' = w (1 - 0.05 dt),
    p' = p + v' dt, and the orientation turns by |w'| dt about w' (closed-form
    quaternion step, DyBodyCoreIntegrator.h:142-167)."""
    dt = 1.0 / 60.0
    dims = np.array([[0.2, 0.3, 0.4, 0]], np.float32)
    m, inertia = S.mass_props(np.array([3]), dims, 1000.0)
    w0 = np.array([0.0, 2.0, 0.0], np.float32)
    sc = S.Scene(np.array([3], np.int32), dims, np.array([[0, 10, 0, 0, 0, 0, 1]], np.float32),
                 np.array([[1, 2, 3]], np.float32), w0[None], m, inertia, np.zeros(1, np.int32),
                 np.zeros((0, 4), np.float32), np.zeros(0, np.int32))
    o = L.OracleWorld(sc, 4, 1)
    o.step(dt)
    p, lv, av = o.bodies()
    v = np.array([1, 2 - 9.81 * dt, 3])
    assert np.allclose(lv[0], v, atol=1e-6)
    w = w0 * (1 - 0.05 * dt)
    assert np.allclose(av[0], w, atol=1e-6)
    assert np.allclose(p[0, :3], np.array([0, 10, 0]) + v * dt, atol=1e-5)
    ang = np.linalg.norm(w) * dt
    assert np.allclose(p[0, 3:], [0, math.sin(ang / 2), 0, math.cos(ang / 2)], atol=1e-6)


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_contacts_vs_live_reference():
    import sys
    sys.path.insert(0, GOLDEN)
    import make_golden as MG
    rng = np.random.default_rng(777)
    for kind in KINDS:
        bad = 0
        n = 300
        for _ in range(n):
            t0, g0, p0, t1, g1, p1 = MG.pair(rng, kind)
            a = L.contacts("pxo", t0, g0, p0, t1, g1, p1)
            b = L.contacts("ref", t0, g0, p0, t1, g1, p1)
            if a.shape != b.shape or (a.shape[0] and (np.abs(a[:, :4] - b[:, :4]).max() > 1e-4 or
                                                      np.abs(a[:, 4:7] - b[:, 4:7]).max() > 1e-3)):
                bad += 1
        assert bad <= (0 if kind != "bb" else int(0.03 * n)), (kind, bad)


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_broadphase_vs_live_reference_abp_and_pabp():
    rng = np.random.default_rng(31)
    n = 5000
    c = rng.uniform(0, 20, (n, 3)).astype(np.float32)
    h = rng.uniform(0.05, 0.4, (n, 3)).astype(np.float32)
    b = np.concatenate([c - h, c + h], 1)
    d = np.full(n, 0.02, np.float32)
    got = L.bp_pairs_pxo(b, d)
    assert np.array_equal(got, L.bp_pairs_ref(b, d))
    assert np.array_equal(got, L.bp_pairs_ref(b, d, pabp=True))


def _single_column(height):
    n = height
    st = np.full(n, S.BOX)
    dims = np.zeros((n, 4))
    dims[:, :3] = 0.5
    pos = np.zeros((n, 3))
    pos[:, 1] = 0.5 + np.arange(n)
    quat = np.tile([0, 0, 0, 1.0], (n, 1))
    return S._finish(st, dims, pos, quat, np.zeros(n), [[0, 1, 0, 0]], [0], 1, "column", {})


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_plane_constraints_solved_after_body_constraints_vs_live_reference():
    """A 10-box column: the reference places each body's static constraints in
    the partitions after its body-body ones, which keeps the bottom box on the
    floor (solved in any other order it sinks by ~2 mm under the column)."""
    sc = _single_column(10)
    o = L.OracleWorld(sc, 4, 1)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1)
    for _ in range(100):
        o.step(1 / 60)
        r.step(1 / 60)
    po, _, _ = o.bodies()
    pr, _, _ = r.bodies()
    assert abs(po[0, 1] - 0.5) < 2e-4 and abs(pr[0, 1] - 0.5) < 2e-4, (po[0, 1], pr[0, 1])
    assert np.abs(po[:, 1] - pr[:, 1]).max() < 2e-3


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
@pytest.mark.parametrize("vx", [0.05, 0.3])
def test_friction_anchors_persist_vs_live_reference(vx):
    """A box on the floor pushed sideways: below the static limit the kept
    friction anchor pulls it back towards where it first touched (a fresh
    anchor every step would leave it ~1e-4 m away); above it the box slides
    and stops. Both follow the reference within 1e-5 m."""
    sc = _single_column(1)
    sc.lin_vel[:, 0] = vx
    o = L.OracleWorld(sc, 4, 1)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1)
    for _ in range(40):
        o.step(1 / 60)
        r.step(1 / 60)
    po, _, _ = o.bodies()
    pr, _, _ = r.bodies()
    assert abs(po[0, 0] - pr[0, 0]) < 1e-5, (po[0, 0], pr[0, 0])


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
@pytest.mark.parametrize("grid,seed", [((6, 3, 6), 1), ((8, 4, 8), 2)])
def test_long_run_statistics_vs_live_reference(grid, seed):
    """300 steps of a settling box pile: trajectories diverge chaotically, the
    pile statistics stay close: height and potential energy within 5%, resting
    floor penetration within 0.3 mm, resting jitter within 10%."""
    import pile_stats as PS
    sc = S.box_grid_drop(nx=grid[0], ny=grid[1], nz=grid[2], seed=seed)
    a = PS.run_stats(L.OracleWorld(sc, 4, 1), sc, 300)
    b = PS.run_stats(L.RefScene(sc, "tgs", 4, 1, threads=8), sc, 300)
    PS.assert_stats_close(a, b)


def _settled(sc, solver, steps):
    """State of `sc` after `steps` reference steps (bodies touching the floor
    and each other), as a new scene."""
    r = L.RefScene(sc, solver, 4, 1, threads=8)
    r.step(1 / 60, steps)
    pose, lv, av = r.bodies()
    return S.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes,
                   sc.plane_scene, 1, sc.name + "_settled")


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
@pytest.mark.parametrize("solver", ["pgs", "tgs"])
def test_one_step_from_contact_rich_state_vs_live_reference(solver):
    """From a contact-rich state (a box pile after 40 reference steps, handed
    to fresh worlds on both sides so neither has friction history) one step of
    coupled contacts depends on the Gauss-Seidel order, which the reference
    takes from its island graph and the colouring here from constraint keys.
    The oracle's deviation from the reference stays within the spread that
    reordering alone produces in the oracle (same constraints, four other
    colour orders): largest and mean deviation within 1.5 times the largest of
    those. The constraint set is the reference's."""
    sc = _settled(S.box_grid_drop(nx=6, ny=3, nz=6, seed=3), solver, 40)
    worlds = [L.OracleWorld(sc, 4, 1, solver=solver, order_salt=k) for k in range(5)]
    r = L.RefScene(sc, solver, 4, 1, threads=1)
    for w in worlds + [r]:
        w.step(1 / 60)
    assert worlds[0].stats()["constraints"] == r.stats()["touching"] > sc.n
    v = [w.bodies()[1] for w in worlds]
    rl = r.bodies()[1]
    scale = np.maximum(np.linalg.norm(rl, axis=1), 9.81 / 60)
    dev_ref = np.abs(v[0] - rl).max(1) / scale
    dev_order = [np.abs(v[0] - x).max(1) / scale for x in v[1:]]
    assert max(d.max() for d in dev_order) > 1e-2  # the order matters for this state
    assert dev_ref.max() <= 1.5 * max(d.max() for d in dev_order)
    assert dev_ref.mean() <= 1.5 * max(d.mean() for d in dev_order)


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
@pytest.mark.parametrize("solver", ["pgs", "tgs"])
def test_steps_of_separate_piles_vs_live_reference(solver):
    """Where the order cannot matter - bodies of a sphere/box pile still
    falling apart, each touching at most the floor - 5 steps agree with the
    reference within 1e-4 of max(|v|, g dt)."""
    sc = S.batched_piles(n_scenes=4, nmin=64, nmax=256, seed=21).subset(1)
    o = L.OracleWorld(sc, 4, 1, solver=solver)
    r = L.RefScene(sc, solver, 4, 1, threads=1)
    o.step(1 / 60, 5)
    r.step(1 / 60, 5)
    _, ol, _ = o.bodies()
    _, rl, _ = r.bodies()
    scale = np.maximum(np.linalg.norm(rl, axis=1), 9.81 / 60)
    assert (np.abs(ol - rl).max(1) / scale).max() < 1e-4


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_pgs_long_run_statistics_vs_live_reference():
    """The box-drop configuration under PGS 4/1 (8x4x8 grids, four seeds, 300
    steps each): PGS leaves more penetration and single runs scatter more, so
    the statistics are compared as means over the seeds - height and potential
    energy within 5%, resting penetration and jitter within 15%."""
    import pile_stats as PS
    a, b = [], []
    for seed in (1, 2, 3, 4):
        sc = S.box_grid_drop(nx=8, ny=4, nz=8, seed=seed)
        a.append(PS.run_stats(L.OracleWorld(sc, 4, 1, solver="pgs"), sc, 300))
        b.append(PS.run_stats(L.RefScene(sc, "pgs", 4, 1, threads=8), sc, 300))
    mean = lambda runs, k: float(np.mean([x[k] for x in runs]))  # noqa: E731
    for k, rel in (("height", 0.05), ("potential", 0.05), ("penetration", 0.15), ("jitter", 0.15)):
        assert abs(mean(a, k) - mean(b, k)) <= rel * abs(mean(b, k)), (k, mean(a, k), mean(b, k))


def test_chunked_oracle_equals_whole_scene_oracle():
    """The full-size batched GPU check splits the oracle into chunks of whole
    worlds stepped on threads: after three steps the chunks give the same
    bodies bit for bit, the same counts and the same constraint and contact
    tables as one oracle over every world."""
    from physx_amd import scenes as S
    sc = S.batched_piles(n_scenes=24, nmin=8, nmax=64, seed=41)
    whole = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    parts = L.ChunkedOracle(sc, chunks=5, pos_iters=4, vel_iters=1)
    whole.step(1.0 / 120.0, 3)
    parts.step(1.0 / 120.0, 3)
    for a, b in zip(whole.bodies(), parts.bodies()):
        assert np.array_equal(a, b)
    ws, ps = whole.stats(), parts.stats()
    for k in ("bp_pairs", "constraints", "contacts"):
        assert ws[k] == ps[k], (k, ws, ps)
    key = lambda c: set(map(tuple, c.tolist()))  # noqa: E731
    assert key(whole.constraints()) == key(parts.constraints())
    rec = lambda r: sorted(map(tuple, r.view(np.int32)[:, :3].tolist()))  # noqa: E731
    wc, pc = whole.contact_data(), parts.contact_data()
    assert rec(wc) == rec(pc)
    order = lambda r: np.lexsort((r.view(np.int32)[:, 1], r.view(np.int32)[:, 0]))  # noqa: E731
    assert np.array_equal(wc[order(wc)].view(np.int32), pc[order(pc)].view(np.int32))   # bit for bit


def _free_boxes(n=24, seed=51):
    """Boxes far apart in the air, random orientations and velocities."""
    rng = np.random.default_rng(seed)
    st = np.full(n, S.BOX)
    dims = np.zeros((n, 4))
    dims[:, :3] = rng.uniform(0.2, 0.6, (n, 3))
    pos = np.stack([np.arange(n) * 3.0, np.full(n, 20.0), np.zeros(n)], 1)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    sc = S._finish(st, dims, pos, q, np.zeros(n), [[0, 1, 0, 0]], [0], 1, "free_boxes", {})
    sc.lin_vel[:] = rng.uniform(-2, 2, (n, 3))
    sc.ang_vel[:] = rng.uniform(-2, 2, (n, 3))
    force = (rng.uniform(-30, 30, (n, 3)) * sc.mass[:, None]).astype(np.float32)
    torque = (rng.uniform(-5, 5, (n, 3)) * sc.inertia).astype(np.float32)
    return sc, force, torque


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
@pytest.mark.parametrize("solver", ["tgs", "pgs"])
def test_forces_and_torques_match_reference(solver):
    """PxRigidDynamic::addForce / addTorque (eFORCE) on free bodies: one step
    of the oracle against the reference CPU SDK within 1e-5 of max(|v|, g dt),
    the linear velocity against the closed form v0 + (F/m + g) dt (before
    damping, which is 0 here), and the forces lasting one step only."""
    sc, force, torque = _free_boxes()
    dt = 1.0 / 60.0
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1, solver=solver)
    r = L.RefScene(sc, solver, 4, 1, threads=2)
    o.set_forces(force, torque)
    r.add_forces(force, torque)
    o.step(dt)
    r.step(dt)
    _, ol, oa = o.bodies()
    _, rl, ra = r.bodies()
    scale = np.maximum(np.linalg.norm(rl, axis=1), 9.81 * dt)[:, None]
    assert np.all(np.abs(ol - rl) <= 1e-5 * scale), np.abs(ol - rl).max()
    assert np.all(np.abs(oa - ra) <= 1e-5 * np.maximum(np.linalg.norm(ra, axis=1), 1.0)[:, None])
    want = sc.lin_vel + (force / sc.mass[:, None] + [0.0, -9.81, 0.0]) * dt
    np.testing.assert_allclose(ol, want, rtol=0, atol=2e-5)
    # the angular change is the world-space I^-1 tau dt (then angular damping 0.05)
    assert np.abs(oa - sc.ang_vel).max() > 0.01
    o.step(dt)
    r.step(dt)
    _, ol2, _ = o.bodies()
    _, rl2, _ = r.bodies()
    np.testing.assert_allclose(ol2, ol + [0.0, -9.81 * dt, 0.0], rtol=0, atol=2e-5)
    assert np.all(np.abs(ol2 - rl2) <= 1e-5 * np.maximum(np.linalg.norm(rl2, axis=1), 9.81 * dt)[:, None])
