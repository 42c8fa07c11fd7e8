"""Convex meshes on the GPU: the HIP path against the oracle on the same
cooked hulls (the library's own cooking of the scene's point clouds), and
against the reference CPU SDK (which cooks the same point clouds itself) by
pile statistics. Tolerances as for boxes: one step within 1e-5 of
max(|v|, g*dt); long runs by height, potential energy, resting penetration
and jitter."""
import numpy as np
import pytest

import oracle_libs as L
import physx_amd as P
from physx_amd import scenes as S

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _mixed(n=192, per_side=8, seed=6, n_points=24):
    """Convex layers with every third body a box and every third a sphere, so
    all convex pair classes (sphere-, box-, convex-convex and plane) occur."""
    sc = S.convex_layers(n_bodies=n, per_side=per_side, seed=seed, n_points=n_points)
    rng = np.random.default_rng(seed)
    t = sc.shape_type.copy()
    dims = sc.dims.copy()
    t[1::3] = S.BOX
    t[2::3] = S.SPHERE
    b, s = t == S.BOX, t == S.SPHERE
    dims[b, :3] = rng.uniform(0.15, 0.4, (int(b.sum()), 3))
    dims[s, 0] = rng.uniform(0.15, 0.35, int(s.sum()))
    dims[s, 1:3] = 0.0
    m, inn = S.mass_props(t, dims, 1000.0)
    mass, inertia = sc.mass.copy(), sc.inertia.copy()
    mass[b | s], inertia[b | s] = m[b | s], inn[b | s]
    q = sc.pose.copy()
    return S.Scene(t, dims, q, sc.lin_vel, sc.ang_vel, mass, inertia, sc.scene_id, sc.planes, sc.plane_scene, 1,
                   "mixed_convex", {}, sc.hull_points, sc.hull_counts)


def _contact_map(rec):
    ids = rec[:, :3].copy().view(np.int32)
    return {(int(a), int(b)): r for (a, b, _), r in zip(ids, rec)}


def _compare_contacts(g, o, tol_p=1e-4, tol_n=1e-3):
    """Every touching pair with the same contact count, normal within tol_n and
    points / separations within tol_p (points matched within the pair)."""
    gm, om = _contact_map(g.contact_data()), _contact_map(o.contact_data())
    assert gm.keys() == om.keys()
    worst = (0.0, None)
    for key, gr in gm.items():
        orr = om[key]
        cnt = int(gr[:3].copy().view(np.int32)[2])
        assert cnt == int(orr[:3].copy().view(np.int32)[2]), key
        dn = float(np.abs(gr[4:7] - orr[4:7]).max())
        gp, op = gr[8:8 + 4 * cnt].reshape(cnt, 4), orr[8:8 + 4 * cnt].reshape(cnt, 4)
        dp = max(float(np.abs(op - r).max(1).min()) for r in gp) if cnt else 0.0
        if max(dn / tol_n, dp / tol_p) > worst[0]:
            worst = (max(dn / tol_n, dp / tol_p), key, dn, dp)
    assert worst[0] <= 1.0, worst
    return len(gm)


def _pair(sc, steps):
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    for _ in range(steps):
        g.step(DT)
        o.step(DT)
    return g, o


@pytest.mark.parametrize("kind", ["convex", "mixed"])
def test_convex_one_step_matches_oracle(cuda, kind):
    sc = S.convex_layers(n_bodies=256, per_side=8) if kind == "convex" else _mixed()
    _one_step_from_pile(sc, kind, min_constraints=100)


@pytest.mark.parametrize("kind", ["hulls", "mixed"])
def test_large_hulls_one_step_matches_oracle(cuda, kind):
    """Hulls past what one wavefront stages in LDS (more than 64 vertices or
    128 faces) take the single-lane separation search, against other large
    hulls and (mixed) against boxes and spheres; same one-step parity."""
    if kind == "hulls":
        sc = S.convex_layers(n_bodies=96, per_side=4, n_points=100, seed=8)
    else:
        sc = _mixed(n=144, per_side=4, seed=8, n_points=100)
    n0 = int(sc.hull_counts[0])
    h = P.cook_convex(sc.hull_points[:n0])
    assert len(h["verts"]) > 64 or len(h["planes"]) > 128, (len(h["verts"]), len(h["planes"]))
    _, rec = _one_step_from_pile(sc, kind, min_constraints=30)
    if kind == "mixed":
        box_hull = [(a, b) for a, b in _contact_map(rec).keys()
                    if b >= 0 and {int(sc.shape_type[a]), int(sc.shape_type[b])} == {S.BOX, S.CONVEX}]
        print("touching box-hull pairs:", len(box_hull), flush=True)
        assert len(box_hull) > 5, len(box_hull)


def test_dense_convex_cluster_overflows_pair_buffer(cuda):
    """300 hulls and boxes packed into a 2.4 m cube: every body overlaps many
    others, far more pairs than the initial buffer of 4 per body, so the
    narrowphase repeats with a larger buffer (the convex search and touching
    lists included). Pair set, constraints and every contact match the
    oracle; one step moves the bodies as the oracle does."""
    rng = np.random.default_rng(12)
    sc = _mixed(n=300, per_side=10, seed=12)
    pose = sc.pose.copy()
    pose[:, :3] = rng.uniform(-1.2, 1.2, (sc.n, 3)) + [0.0, 2.0, 0.0]
    st = S.Scene(sc.shape_type, sc.dims, pose, sc.lin_vel, sc.ang_vel, sc.mass, sc.inertia, sc.scene_id, sc.planes,
                 sc.plane_scene, 1, "dense_convex", {}, sc.hull_points, sc.hull_counts)
    g, o = _pair(st, 1)
    gs, os_ = g.stats(), o.stats()
    print(gs, flush=True)
    assert gs["bp_pairs"] > 4 * sc.n, gs
    for k in ("bp_pairs", "constraints", "contacts"):
        assert gs[k] == os_[k], (k, gs, os_)
    key = lambda c: set(map(tuple, c[:, [0, 1, 3]].tolist()))  # noqa: E731
    assert key(g.constraints()) == key(o.constraints())
    assert _compare_contacts(g, o) == gs["constraints"]
    _, gl, _ = g.bodies()
    _, ol, _ = o.bodies()
    assert np.isfinite(gl).all()
    assert np.abs(gl - ol).max() <= 1e-3 * max(1.0, np.abs(ol).max())


def test_full_size_convex_pile_step_matches_oracle(cuda):
    """The 65,536-hull pile (24-point ellipsoid hulls, TGS 4/1, dt 1/60) at
    step 200: one step from that state on the GPU and in the oracle gives the
    same pair and constraint counts, every contact within 1e-4 m / 1e-3 and
    velocities within 1e-4 of max(|v|, g*dt) for all but 0.1% of the hulls."""
    _one_step_from_pile(S.convex_layers(), "convex_65536", min_constraints=50000, warm=200, tol_v=1e-4,
                        max_off=0.001)


def _one_step_from_pile(sc, kind, min_constraints, warm=45, tol_v=1e-5, max_off=0.0):
    import time
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    g.step(DT, warm)   # the first layers have landed and pile up
    pose, lv, av = g.bodies()
    # both pipelines start from that state with no history (friction anchors
    # and colours are carried from step to step, so they start empty alike)
    st = S.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes,
                 sc.plane_scene, 1, sc.name, {}, sc.hull_points, sc.hull_counts)
    g = P.Scene(st, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(st, pos_iters=4, vel_iters=1)
    g.step(DT)
    t0 = time.time()
    o.step(DT)
    print(f"{kind}: {sc.n} bodies, oracle step {time.time() - t0:.1f} s", flush=True)
    gs, os_ = g.stats(), o.stats()
    print(kind, gs, flush=True)
    assert gs["constraints"] > min_constraints
    for k in ("bp_pairs", "constraints", "contacts"):
        assert gs[k] == os_[k], (k, gs, os_)
    key = lambda c: set(map(tuple, c[:, [0, 1, 3]].tolist()))  # noqa: E731
    assert key(g.constraints()) == key(o.constraints())
    grec = g.contact_data()
    assert _compare_contacts(g, o) == gs["constraints"]
    gp, gl, ga = g.bodies()
    op, ol, oa = o.bodies()
    tol = tol_v * np.maximum(np.linalg.norm(ol, axis=1), 9.81 * DT)
    off = np.abs(gl - ol).max(1) > tol
    print("bodies off:", int(off.sum()), "max", float(np.abs(gl - ol).max()), flush=True)
    assert off.sum() <= max_off * len(gl), np.abs(gl - ol).max()
    if max_off == 0.0:
        assert np.abs(gp[:, :3] - op[:, :3]).max() < 1e-6
    return gs, grec


def test_convex_multi_step_tracks_oracle(cuda):
    sc = _mixed(n=128)
    g, o = _pair(sc, 60)
    gp, _, _ = g.bodies()
    op, _, _ = o.bodies()
    assert np.isfinite(gp).all()
    assert np.abs(gp[:, :3] - op[:, :3]).max() < 5e-3


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_convex_long_run_statistics_match_reference(cuda):
    """256-hull piles settling for 300 steps on the GPU against the reference
    CPU SDK stepping the same point clouds (cooked by its own cooker), means
    over four seeds (one pile's statistics vary by a few percent from seed to
    seed in the reference itself): height and potential energy within 5%,
    resting penetration within 0.5 mm, resting jitter within 20%."""
    import pile_stats as PS
    a, b = [], []
    for seed in (5, 6, 7, 8):
        sc = S.convex_layers(n_bodies=256, per_side=8, seed=seed)
        a.append(PS.run_stats(P.Scene(sc, position_iterations=4, velocity_iterations=1), sc, 300))
        b.append(PS.run_stats(L.RefScene(sc, "tgs", 4, 1, threads=8), sc, 300))
        print(seed, "gpu", a[-1], "ref", b[-1], flush=True)
    mean = lambda rs: {k: float(np.mean([r[k] for r in rs])) for k in rs[0]}  # noqa: E731
    PS.assert_stats_close(mean(a), mean(b), rel=0.05, pen_abs=5e-4, jitter_rel=0.2)
