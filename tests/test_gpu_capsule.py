"""Capsules on the GPU: piles of spheres, capsules, boxes and convex hulls
(every pair class with a capsule: sphere-, capsule-, box-, convex-capsule
and plane-capsule) through the HIP path against the oracle, and against the
reference CPU SDK by pile statistics. Tolerances as for convex meshes."""
import numpy as np
import pytest

import oracle_libs as L
import physx_amd as P
from physx_amd import scenes as S
from test_gpu_convex import _compare_contacts, _pair

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def test_capsule_pile_one_step_matches_oracle(cuda):
    sc = S.mixed_pile(n_bodies=512, per_side=8)
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    g.step(DT, 50)   # the first layers have landed and pile up
    pose, lv, av = g.bodies()
    st = S.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes,
                 sc.plane_scene, 1, sc.name, {}, sc.hull_points, sc.hull_counts)
    g, o = _pair(st, 1)
    gs, os_ = g.stats(), o.stats()
    print(gs, flush=True)
    assert gs["constraints"] > 150
    for k in ("bp_pairs", "constraints", "contacts"):
        assert gs[k] == os_[k], (k, gs, os_)
    con = g.constraints()
    key = lambda c: set(map(tuple, c[:, [0, 1, 3]].tolist()))  # noqa: E731
    assert key(con) == key(o.constraints())
    # every capsule pair class is exercised
    t = sc.shape_type
    a, b = con[:, 0].astype(np.int64), con[:, 1].astype(np.int64)
    kinds = {(int(t[i]), int(t[j]) if j >= 0 else 1) for i, j in zip(a, b)}
    for pair in ((2, 1), (0, 2), (2, 2), (2, 3), (2, 5)):
        assert pair in kinds or pair[::-1] in kinds, (pair, kinds)
    assert _compare_contacts(g, o) == gs["constraints"]
    gp, gl, ga = g.bodies()
    op, ol, oa = o.bodies()
    tol_v = 1e-5 * np.maximum(np.linalg.norm(ol, axis=1), 9.81 * DT)
    assert np.all(np.abs(gl - ol).max(1) <= tol_v), np.abs(gl - ol).max()
    assert np.abs(gp[:, :3] - op[:, :3]).max() < 1e-6


def test_capsule_pile_multi_step_tracks_oracle(cuda):
    """The first 30 steps: the bottom layer lands and rolls or tumbles against
    its neighbours. A hull landing on an edge passes over its face twice (as
    the reference does), and which copy of a repeated vertex the reduction to
    four keeps, and so the order the solver visits the points in, turns on the
    last bit of the pose (a 1e-7 m shift flips it half the time in the
    oracle itself), so a few tumbling bodies part ways; the rest track."""
    sc = S.mixed_pile(n_bodies=256, per_side=8, seed=10)
    g, o = _pair(sc, 30)
    gp, _, _ = g.bodies()
    op, _, _ = o.bodies()
    assert np.isfinite(gp).all()
    d = np.abs(gp[:, :3] - op[:, :3]).max(1)
    print("max", d.max(), "share within 1e-3", float((d < 1e-3).mean()), flush=True)
    assert np.median(d) < 1e-4
    assert (d < 1e-3).mean() >= 0.95
    assert d.max() < 0.1


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_capsule_pile_long_run_statistics_match_reference(cuda):
    """512-body mixed piles settling for 300 steps on the GPU against the
    reference CPU SDK, means over four seeds: height and potential energy
    within 5%, resting penetration within 0.5 mm, resting jitter within 20%."""
    import pile_stats as PS
    a, b = [], []
    for seed in (9, 10, 11, 12):
        sc = S.mixed_pile(n_bodies=512, per_side=8, seed=seed)
        a.append(PS.run_stats(P.Scene(sc, position_iterations=4, velocity_iterations=1), sc, 300))
        b.append(PS.run_stats(L.RefScene(sc, "tgs", 4, 1, threads=8), sc, 300))
        print(seed, "gpu", a[-1], "ref", b[-1], flush=True)
    mean = lambda rs: {k: float(np.mean([r[k] for r in rs])) for k in rs[0]}  # noqa: E731
    PS.assert_stats_close(mean(a), mean(b), rel=0.05, pen_abs=5e-4, jitter_rel=0.2)


def _rand_quat(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


@pytest.mark.parametrize("kind", ["sphere-capsule", "capsule-capsule", "parallel", "plane-capsule", "capsule-box"])
def test_capsule_immediate_contacts_match_oracle(cuda, kind):
    """Batched immediate-mode generation (PxGenerateContacts) of capsule
    pairs on the device against the oracle, pair by pair: the same counts,
    points and separations within 1e-4 m, normals within 1e-3 (capsule-box
    pairs compared as point sets, near-coincident points merged)."""
    from test_cpu_capsule import _same
    rng = np.random.default_rng(["sphere-capsule", "capsule-capsule", "parallel", "plane-capsule",
                                 "capsule-box"].index(kind) + 200)
    n = 3000
    g0 = np.zeros((n, 4), np.float32)
    g1 = np.zeros((n, 4), np.float32)
    p0 = np.zeros((n, 7), np.float32)
    p1 = np.zeros((n, 7), np.float32)
    u = rng.normal(size=(n, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    p0[:, :3] = u * rng.uniform(0.0, 0.9, (n, 1))
    p0[:, 3:] = _rand_quat(rng, n)
    p1[:, 3:] = _rand_quat(rng, n)
    g0[:, 0] = rng.uniform(0.05, 0.3, n)
    g0[:, 1] = rng.uniform(0.0, 0.5, n)
    g1[:, 0] = rng.uniform(0.05, 0.3, n)
    g1[:, 1] = rng.uniform(0.0, 0.5, n)
    t0, t1 = {"sphere-capsule": (0, 2), "capsule-capsule": (2, 2), "parallel": (2, 2), "plane-capsule": (1, 2),
              "capsule-box": (2, 3)}[kind]
    if kind == "parallel":
        p0[:, 3:] = p1[:, 3:]
    elif kind == "plane-capsule":
        p0[:] = [0.0, -0.3, 0.0, 0.0, 0.0, 0.7071068, 0.7071068]
        p1[:, 1] = rng.uniform(-0.3, 0.6, n)
    elif kind == "capsule-box":
        g1[:, :3] = rng.uniform(0.1, 0.4, (n, 3))
    cnt, out = P.generate_contacts(np.full(n, t0, np.int32), g0, p0, np.full(n, t1, np.int32), g1, p1, 0.04, 1.0)
    bad = touching = 0
    for k in range(n):
        ref = L.contacts("pxo", t0, g0[k], p0[k], t1, g1[k], p1[k])
        got = out[k, : max(int(cnt[k]), 0)]
        touching += len(ref) > 0
        if not _same(got, ref, dedupe=kind == "capsule-box"):
            bad += 1
    print(kind, "touching", touching, "differ", bad, flush=True)
    assert touching > n // 10
    assert bad <= max(2, touching // 500), (bad, touching)
