"""Batched and served scenes on the GPU: a scene's result does not depend on
the other scenes stepped with it.

Every stage keys its work by world (per-world broadphase rows, world-local
colouring keys, world-relative friction-patch keys), so one world stepped
inside a batch, stepped after other worlds were removed from the batch, or
admitted into a running serving loop must give the same bits as the same
world stepped alone in its own scene."""
import numpy as np
import pytest

import physx_amd as P
from physx_amd import scenes

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
ITERS = dict(position_iterations=4, velocity_iterations=1)


def _mix(n_scenes=12):
    # worlds of 8..96 bodies: box stacks (even worlds) and sphere/box piles (odd)
    return scenes.batched_piles(n_scenes=n_scenes, nmin=8, nmax=96, seed=7)


def _alone(sc, world, steps, solver=P.PXH_SOLVER_TGS):
    s = P.Scene(sc.subset(world), solver_type=solver, **ITERS)
    s.step(DT, steps)
    out = s.bodies()
    s.close()
    return out


def _same(got, want, label):
    for g, w, name in zip(got, want, ("pose", "linear velocity", "angular velocity")):
        assert g.shape == w.shape, f"{label}: {name} shape {g.shape} vs {w.shape}"
        if not np.array_equal(g, w):
            d = np.abs(g.astype(np.float64) - w)
            raise AssertionError(f"{label}: {name} differs (max {d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)})")


@pytest.mark.parametrize("solver", [P.PXH_SOLVER_TGS, P.PXH_SOLVER_PGS], ids=["tgs", "pgs"])
def test_world_in_batch_matches_world_alone(cuda, solver):
    sc = _mix()
    full = P.Scene(sc, solver_type=solver, **ITERS)
    full.step(DT, 30)
    pose, lv, av = full.bodies()
    assert full.world_count() == sc.n_scenes
    for w in (0, 3, 7, 10, 11):
        f, c = full.world_range(w)
        assert c == int((sc.scene_id == w).sum())
        print(f"world {w}: bodies {f}..{f + c}", flush=True)
        _same((pose[f:f + c], lv[f:f + c], av[f:f + c]), _alone(sc, w, 30, solver), f"world {w}")


def test_removed_worlds_leave_the_others_unchanged(cuda):
    sc = _mix()
    s = P.Scene(sc, **ITERS)
    s.step(DT, 12)
    gone = [1, 4, 5, 10]
    s.remove_worlds(gone)
    keep = [w for w in range(sc.n_scenes) if w not in gone]
    assert s.world_count() == len(keep)
    assert s.n == int(np.isin(sc.scene_id, keep).sum())
    s.step(DT, 18)
    pose, lv, av = s.bodies()
    for i, w in enumerate(keep):
        f, c = s.world_range(i)
        print(f"kept world {w} -> {i}: bodies {f}..{f + c}", flush=True)
        _same((pose[f:f + c], lv[f:f + c], av[f:f + c]), _alone(sc, w, 30), f"world {w}")


def test_remove_worlds_rejects_bad_ids(cuda):
    sc = _mix(3)
    s = P.Scene(sc, **ITERS)
    with pytest.raises(P.PxhError):
        s.remove_worlds([3])
    with pytest.raises(P.PxhError):
        s.remove_worlds([-1])
    s.remove_worlds([0, 1, 2])
    assert s.world_count() == 0 and s.n == 0
    s.step(DT, 1)  # an empty scene steps


def test_served_scene_matches_scene_alone(cuda):
    """Scenes arrive over time with different step counts into a loop that
    holds at most four at once; each result equals the scene stepped alone."""
    sc = _mix()
    sv = P.Server(hbm_budget_bytes=1 << 30, max_worlds=4, **ITERS)
    steps = {w: 6 + 5 * (w % 4) for w in range(sc.n_scenes)}
    ids, results = {}, {}
    t = 0
    while len(results) < sc.n_scenes:
        for w in range(sc.n_scenes):
            if 2 * w == t:
                ids[sv.submit(sc.subset(w), steps[w])] = w
        st = sv.stats()
        assert st["active"] <= 4
        sv.step(DT)
        for rid in sv.finished():
            results[ids[rid]] = sv.result(rid)
        t += 1
        assert t < 500
    st = sv.stats()
    assert st["queued"] == 0 and st["active"] == 0 and st["completed"] == sc.n_scenes
    assert st["body_steps"] == sum(int((sc.scene_id == w).sum()) * steps[w] for w in range(sc.n_scenes))
    for w in range(sc.n_scenes):
        pose, lv, av, tm = results[w]
        assert tm[0] <= tm[1] <= tm[2]
        print(f"served world {w}: {steps[w]} steps", flush=True)
        _same((pose, lv, av), _alone(sc, w, steps[w]), f"served world {w}")
    sv.close()


def test_server_admission_respects_hbm_budget(cuda):
    sc = _mix()
    per_body = 8192
    sizes = [int((sc.scene_id == w).sum()) for w in range(sc.n_scenes)]
    budget = per_body * 2 * max(sizes)
    sv = P.Server(hbm_budget_bytes=budget, max_worlds=64, bytes_per_body=per_body, **ITERS)
    for w in range(sc.n_scenes):
        sv.submit(sc.subset(w), 3)
    done = 0
    for _ in range(200):
        sv.step(DT)
        st = sv.stats()
        assert st["active_bytes"] <= budget
        assert st["active_bytes"] == per_body * st["active_bodies"]
        for rid in sv.finished():
            sv.result(rid)
            done += 1
        if done == sc.n_scenes:
            break
    assert done == sc.n_scenes
    with pytest.raises(P.PxhError):
        big = scenes.batched_piles(n_scenes=1, nmin=400, nmax=400, seed=3)
        sv.submit(big, 1)  # larger than the whole budget
    sv.close()


def test_served_stream_of_16_to_4096_body_scenes(cuda):
    """The served workload's shape: 24 scenes of log-uniform [16, 4096]
    bodies (box stacks and sphere/box piles, the same generator as the served
    benchmark's pool) arriving at seeded exponential intervals while others
    run, each asking for k ~ U{1..32} steps. Every result equals the scene
    stepped alone for its k steps, bit for bit. For three of them (the
    largest among them) one step from the served result, with no history, on
    the GPU and in the oracle: same touching pairs and contact counts,
    velocities within 1e-4 of max(|v|, g*dt)."""
    import oracle_libs as L
    pool = scenes.batched_piles(n_scenes=24, nmin=16, nmax=4096, seed=11)
    sizes = [int((pool.scene_id == w).sum()) for w in range(pool.n_scenes)]
    assert max(sizes) > 2000 and min(sizes) < 64, sizes
    rng = np.random.default_rng(17)
    k = rng.integers(1, 33, pool.n_scenes)
    arrive = np.cumsum(rng.exponential(1.5, pool.n_scenes)).astype(int)   # serving step of each arrival
    sv = P.Server(hbm_budget_bytes=200 << 30, max_worlds=65536, **ITERS)
    ids, results, t = {}, {}, 0
    while len(results) < pool.n_scenes:
        for w in np.nonzero(arrive == t)[0]:
            ids[sv.submit(pool.subset(int(w)), int(k[w]))] = int(w)
        sv.step(DT)
        for rid in sv.finished():
            results[ids[rid]] = sv.result(rid)
        t += 1
        assert t < 2000
    st = sv.stats()
    assert st["completed"] == pool.n_scenes and st["active"] == 0
    assert st["device_bytes"] > 0
    sv.close()
    for w in range(pool.n_scenes):
        pose, lv, av, tm = results[w]
        assert tm[0] <= tm[1] <= tm[2]
        _same((pose, lv, av), _alone(pool, w, int(k[w])), f"served scene {w} ({sizes[w]} bodies, {k[w]} steps)")
    sample = sorted(range(pool.n_scenes), key=lambda w: -sizes[w])[:1] + [3, 10]
    for w in sample:
        pose, lv, av, _ = results[w]
        sub = pool.subset(w)
        st = scenes.Scene(sub.shape_type, sub.dims, pose, lv, av, sub.mass, sub.inertia, sub.scene_id, sub.planes,
                          sub.plane_scene, 1, sub.name, {})
        g = P.Scene(st, **ITERS)
        o = L.OracleWorld(st, pos_iters=4, vel_iters=1)
        g.step(DT)
        o.step(DT)
        gs, os_ = g.stats(), o.stats()
        print(f"scene {w} ({sizes[w]} bodies) after {k[w]} served steps:", gs, flush=True)
        for key in ("bp_pairs", "constraints", "contacts"):
            assert gs[key] == os_[key], (w, key, gs, os_)
        ck = lambda c: set(map(tuple, c.tolist()))  # noqa: E731
        assert ck(g.constraints()) == ck(o.constraints())
        _, gl, _ = g.bodies()
        _, ol, _ = o.bodies()
        tol = 1e-4 * np.maximum(np.linalg.norm(ol, axis=1), 9.81 * DT)
        assert np.all(np.abs(gl - ol).max(1) <= tol), (w, float(np.abs(gl - ol).max()))
        g.close()


@pytest.mark.parametrize("lds", [512, 16])
def test_world_solve_matches_colour_launches(cuda, lds):
    """The per-world persistent solve (each world's iterations in one
    workgroup, velocity records in LDS, or in HBM for worlds past the LDS
    limit: lds=16 sends most worlds there) gives the same bits as one launch
    per colour and iteration, over 30 steps of 64 worlds of 8..512 bodies."""
    sc = scenes.batched_piles(n_scenes=64, nmin=8, nmax=512, seed=9)
    a = P.Scene(sc, **ITERS)
    a.set_tuning(layout=P.PXH_SOLVE_LAYOUT_COLORS)
    b = P.Scene(sc, **ITERS)
    b.set_tuning(layout=P.PXH_SOLVE_LAYOUT_WORLDS, lds_bodies=lds, world_workgroups=16)
    for s in range(30):
        a.step(DT)
        b.step(DT)
        if s in (0, 29):
            sa, sb = a.stats(), b.stats()
            for k in ("bp_pairs", "constraints", "contacts", "colors"):
                assert sa[k] == sb[k], (s, k, sa, sb)
    _same(b.bodies(), a.bodies(), f"world solve (lds {lds}) vs colour launches")


def test_world_solve_single_large_world_and_free_fall(cuda):
    """The per-world layout forced on one 2,000-body world (past the LDS
    limit: one workgroup works in HBM) and on bodies in free fall (no
    constraints: the kernel only integrates) matches the colour launches."""
    big = scenes.batched_piles(n_scenes=1, nmin=2000, nmax=2000, seed=13)
    high = scenes.batched_piles(n_scenes=3, nmin=20, nmax=40, seed=14)
    high.pose[:, 0] = 2.0 * np.arange(high.n)   # apart, far above the floor
    high.pose[:, 1] += 50.0
    for sc in (big, high):
        a = P.Scene(sc, **ITERS)
        a.set_tuning(layout=P.PXH_SOLVE_LAYOUT_COLORS)
        b = P.Scene(sc, **ITERS)
        b.set_tuning(layout=P.PXH_SOLVE_LAYOUT_WORLDS)
        a.step(DT, 12)
        b.step(DT, 12)
        assert a.stats()["constraints"] == b.stats()["constraints"]
        _same(b.bodies(), a.bodies(), f"{sc.name}: world layout vs colour launches")
    assert a.stats()["constraints"] == 0


def test_autotuned_solve_keeps_every_bit(cuda):
    """Autotuning the solve's launch shape (each of the 9 candidates for 2
    steps, or for one when that step is 25% slower than the best mean so far,
    then the fastest mean) changes no bit of the simulation: 30 steps of
    48 worlds equal the same steps with fixed colour launches, and the log
    holds every candidate with its measured time."""
    sc = scenes.batched_piles(n_scenes=48, nmin=8, nmax=256, seed=15)
    a = P.Scene(sc, **ITERS)
    a.set_tuning(layout=P.PXH_SOLVE_LAYOUT_COLORS)
    b = P.Scene(sc, **ITERS)
    b.set_tuning(autotune_steps=2)
    a.step(DT, 30)
    b.step(DT, 30)
    _same(b.bodies(), a.bodies(), "autotuned vs colour launches")
    t = b.tuning()
    log = b.tuning_log()
    print("chosen", t, "log", log, flush=True)
    assert t[3] == 0 and len(log) == 9
    assert all(m > 0 for *_, m in log)
    assert tuple(t[:3]) == min(log, key=lambda c: c[3])[:3]
