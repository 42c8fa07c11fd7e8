"""GPU joint limits and drives (pxh_scene_set_joint_params: hard limits of
revolute and prismatic joints, locked limits, restitution, revolute velocity
drives with force limits, freespin and gear ratios) against the oracle,
which tests/test_cpu_joint_limits.py pins to the reference SDK: one step at a
time along the GPU's own trajectory, the single-joint closed-form cases, and
the argument checks."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L
from physx_amd import scenes as S
from test_cpu_joint_limits import _drive_cases, _one_joint, _q

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def _close(g, o, rel, tag):
    gp, gl, ga = g
    op, ol, oa = o
    lt = rel * np.maximum(np.abs(ol).max(1), 9.81 * DT)
    at = rel * np.maximum(np.abs(oa).max(1), 1.0)
    bl = np.where(np.abs(gl - ol).max(1) > lt)[0]
    ba = np.where(np.abs(ga - oa).max(1) > at)[0]
    assert bl.size == 0 and ba.size == 0, (tag, bl[:4], ba[:4], gl[bl[:2]], ol[bl[:2]])
    assert np.abs(gp[:, :3] - op[:, :3]).max() <= 1e-6, tag


def _track(sc, steps, rel, gravity=(0.0, -9.81, 0.0)):
    """each step from the GPU's state: the GPU's next state and the oracle's"""
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1, gravity=gravity)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1, gravity=gravity)
    for s in range(steps):
        o.set_bodies(*g.bodies())
        g.step(DT)
        o.step(DT)
        _close(g.bodies(), o.bodies(), rel, s)
    return g


@pytest.mark.parametrize("seed", [7, 8])
def test_distance_joints_track_oracle(cuda, seed):
    """48 distance joints (max, min and max, exact distance, springs): 60
    steps, each within 3e-5 of the oracle's step from the same state."""
    _track(S.distance_joints(48, seed), 60, 3e-5)


@pytest.mark.parametrize("seed", [7, 8])
def test_cone_limited_joints_track_oracle(cuda, seed):
    """48 spherical joints with elliptic cone limits (hard and soft): 60
    steps, each within 3e-5 of the oracle's step from the same state (the
    swing's arc tangents and tangents are the device's)."""
    _track(S.cone_limited_joints(48, seed), 60, 3e-5)


@pytest.mark.parametrize("seed", [7, 8])
def test_limited_joints_track_oracle(cuda, seed):
    """48 revolute and prismatic joints with hard limits (some locked, some
    with restitution), force-limited / unbounded / freespin drives and gear
    ratios: 60 steps, each within 3e-5 of max(|v|, g dt) of the oracle's step
    from the same state (measured up to 1.4e-5: the twist angle's arc cosine
    is the device's, not the host library's, and a limit resting on its stop
    sits where rounding decides whether it keeps its bias)."""
    _track(S.limited_joints(48, seed), 60, 3e-5)


def test_single_joint_cases_track_oracle(cuda):
    """The oracle's closed-form cases (a hinge resting on its stop, a slider
    on its stop, locked limits, the three drives): 120 steps each within 1e-5
    of the oracle's step from the same state."""
    f1 = np.concatenate([[0.0, 2.0, 0.0], _q([0, 1, 0], -np.pi / 2)])
    f0 = np.concatenate([[-0.5, 0.0, 0.0], f1[3:]])
    up = _q([0, 0, 1], np.pi / 2)
    cases = [(_one_joint(S.JOINT_REVOLUTE, S.BOX, [0.3, 0.1, 0.1, 0], [0.5, 2.0, 0.0, 0, 0, 0, 1], f0, f1,
                         [1, -0.3, 0.2, 0, 0, 0, 0, 1, 0, 0]), (0.0, -9.81, 0.0)),
             (_one_joint(S.JOINT_PRISMATIC, S.BOX, [0.1, 0.1, 0.1, 0], [0, 1, 0, 0, 0, 0, 1],
                         np.concatenate([[0, 0, 0], up]), np.concatenate([[0, 1, 0], up]),
                         [1, -0.2, 0.3, 0, 0, 0, 0, 1, 0, 0]), (0.0, -9.81, 0.0))]
    cases += [(c, (0.0, 0.0, 0.0)) for c in _drive_cases()]
    for k, (sc, grav) in enumerate(cases):
        _track(sc, 120, 1e-5, grav)


def test_joint_params_argument_checks(cuda):
    sc = S.joint_singles(8, 3)
    g = P.Scene(sc)
    kinds = sc.joints["type"]
    fixed = int(np.where(kinds == S.JOINT_FIXED)[0][0])
    prism = int(np.where(kinds == S.JOINT_PRISMATIC)[0][0])
    with pytest.raises(P.PxhError, match="fixed joints take no limits"):
        g.set_joint_params([fixed], [[1, -0.1, 0.1, 0, 0, 0, 0, 1, 0, 0]])
    with pytest.raises(P.PxhError, match="only revolute joints take a drive"):
        g.set_joint_params([prism], [[2, 0, 0, 0, 0, 1, 1, 1, 0, 0]])
    with pytest.raises(P.PxhError, match="out of range"):
        g.set_joint_params([len(kinds)], [[1, -0.1, 0.1, 0, 0, 0, 0, 1, 0, 0]])
    with pytest.raises(P.PxhError, match="unknown flags"):
        g.set_joint_params([prism], [[33, -0.1, 0.1, 0, 0, 0, 0, 1, 0, 0]])
    g.set_joint_params([prism], [[1, -0.1, 0.1, 0, 0, 0, 0, 1, 0, 0]])
    g.step(DT)


def _limited_chains():
    """jointed_chains with limits on every revolute (±0.3 rad, soft on every
    third) and prismatic (±0.05 m) joint and a drive on every fourth revolute"""
    sc = S.jointed_chains(n_scenes=8, links=12, drops=6, seed=3)
    k = len(sc.joints["type"])
    p = np.zeros((k, 10), np.float32)
    p[:, 7] = 1.0
    rev = sc.joints["type"] == S.JOINT_REVOLUTE
    pri = sc.joints["type"] == S.JOINT_PRISMATIC
    ids = np.arange(k)
    p[rev, 0] = 1
    p[rev, 1], p[rev, 2] = -0.3, 0.3
    p[rev & (ids % 3 == 0), 8] = 50.0
    p[rev & (ids % 3 == 0), 9] = 1.0
    p[rev & (ids % 4 == 1), 0] = 3
    p[rev & (ids % 4 == 1), 5] = 1.5
    p[rev & (ids % 4 == 1), 6] = 2.0
    p[pri, 0] = 1
    p[pri, 1], p[pri, 2] = -0.05, 0.05
    sc.joints["params"] = p
    return sc


def test_limited_world_alone_in_batch_and_after_removal(cuda):
    """Jointed chains with limits (hard and soft) and drives: a world's bits
    stepped alone equal its slice of a batch of eight, also after another
    world is removed (joint parameters renumbered with the joints)."""
    sc = _limited_chains()
    assert (sc.joints["params"][:, 0] > 0).any()
    alone = P.Scene(sc.subset(5), position_iterations=4, velocity_iterations=1)
    batch = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    batch.set_tuning(layout=P.PXH_SOLVE_LAYOUT_COLORS)
    for _ in range(20):
        alone.step(DT)
        batch.step(DT)
    f, c = batch.world_range(5)
    for x, y in zip(alone.bodies(), batch.bodies()):
        assert np.array_equal(x, y[f:f + c])
    batch.remove_worlds([2])
    f2, c2 = batch.world_range(4)
    for _ in range(20):
        alone.step(DT)
        batch.step(DT)
    for x, y in zip(alone.bodies(), batch.bodies()):
        assert np.array_equal(x, y[f2:f2 + c2])


def test_jointed_bodies_collide_when_enabled(cuda):
    """The overlapping jointed boxes: no constraint by default, one with
    collision enabled (parameter flag 8), the velocities within 1e-5 of the
    oracle's either way."""
    from test_cpu_joint_limits import _overlapping_pair
    for flags, touching in ((0, 0), (8, 1)):
        sc = _overlapping_pair(flags)
        g = P.Scene(sc)
        o = L.OracleWorld(sc)
        g.step(DT)
        o.step(DT)
        assert g.stats()["constraints"] == touching == o.stats()["constraints"]
        _close(g.bodies(), o.bodies(), 1e-5, flags)
