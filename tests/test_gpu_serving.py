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
