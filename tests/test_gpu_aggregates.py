"""GPU aggregates (pxh_scene_add_aggregate) against the oracle, which
tests/test_cpu_aggregates.py pins to the reference SDK: the same broadphase
pairs and constraint table, one step's velocities within 1e-5; aggregates
with self-collisions change no bit; bodies added after an aggregate and
world removal keep the membership."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L
from physx_amd import scenes as S

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def test_aggregates_one_step_match_oracle(cuda):
    sc = S.aggregate_pile(n_groups=64, per_group=4, seed=7)
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    for s in range(3):
        g.step(DT)
        o.step(DT)
        gs, os_ = g.stats(), o.stats()
        assert gs["bp_pairs"] == os_["bp_pairs"], (s, gs, os_)
        assert gs["constraints"] == os_["constraints"], (s, gs, os_)
        gc = sorted(map(tuple, g.constraints()[:, :2].tolist()))
        oc = sorted(map(tuple, o.constraints()[:, :2].tolist()))
        assert gc == oc, s
    _, gl, ga = g.bodies()
    _, ol, oa = o.bodies()
    tol = 1e-5 * np.maximum(np.abs(ol).max(1), 9.81 * DT) * 3
    assert np.all(np.abs(gl - ol).max(1) <= tol), float(np.abs(gl - ol).max())
    assert np.abs(ga - oa).max() <= 3e-5 * max(1.0, float(np.abs(oa).max()))


def test_self_colliding_aggregates_change_nothing(cuda):
    sc = S.aggregate_pile(n_groups=36, per_group=4, seed=9)
    sc.aggregates = [(m, True) for m, _ in sc.aggregates]
    free = S.aggregate_pile(n_groups=36, per_group=4, seed=9)
    free.aggregates = None
    a = P.Scene(sc)
    b = P.Scene(free)
    for _ in range(10):
        a.step(DT)
        b.step(DT)
    for x, y in zip(a.bodies(), b.bodies()):
        assert np.array_equal(x, y)


def test_aggregate_membership_survives_adds_and_removal(cuda):
    """Two worlds of aggregated groups: bodies added later (a third world)
    and removing world 0 leave world 1 stepping bit for bit as it does alone."""
    w = S.aggregate_pile(n_groups=16, per_group=4, seed=11)
    both = S.aggregate_pile(n_groups=16, per_group=4, seed=11)
    # two copies of the pile as worlds 0 and 1 of one scene
    g = P.Scene(position_iterations=4, velocity_iterations=1)
    g.add(both)
    second = S.aggregate_pile(n_groups=16, per_group=4, seed=11)
    second.scene_id = np.ones(second.n, np.int32)
    second.plane_scene = np.ones(1, np.int32)
    g.add(second)
    alone = P.Scene(w, position_iterations=4, velocity_iterations=1)
    for _ in range(5):
        g.step(DT)
        alone.step(DT)
    third = S.aggregate_pile(n_groups=4, per_group=4, seed=12)
    third.scene_id = np.full(third.n, 2, np.int32)
    third.plane_scene = np.full(1, 2, np.int32)
    g.add(third)
    g.remove_worlds([0])
    for _ in range(5):
        g.step(DT)
        alone.step(DT)
    f, c = g.world_range(0)
    for x, y in zip(alone.bodies(), g.bodies()):
        assert np.array_equal(x, y[f:f + c])


def test_aggregate_rejects_bad_arguments(cuda):
    g = P.Scene(S.aggregate_pile(n_groups=4, per_group=4, seed=7))
    with pytest.raises(P.PxhError, match="already in an aggregate"):
        g.add_aggregate([0, 1])
    with pytest.raises(P.PxhError, match="out of range"):
        g.add_aggregate([999])
