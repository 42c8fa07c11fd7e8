"""Edge cases of the GPU step against the oracle: an empty scene, planes
without bodies, bodies in free flight (no pairs, closed-form velocities), a
dense cluster whose broadphase pairs overflow the initial pair buffer (the
step enlarges it and runs the narrowphase again), and bodies added to a
scene that has already stepped."""
import numpy as np
import pytest

import oracle_libs as L
import physx_amd as P
from physx_amd import scenes as S

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


def _scene(t, dims, pos, planes=None, lin=None):
    n = len(t)
    t = np.asarray(t, np.int32)
    dims = np.asarray(dims, np.float32).reshape(n, 4)
    quat = np.tile([0.0, 0.0, 0.0, 1.0], (n, 1))
    pose = np.concatenate([np.asarray(pos, np.float64).reshape(n, 3), quat], 1).astype(np.float32)
    m, inn = S.mass_props(t, dims, 1000.0)
    planes = np.zeros((0, 4), np.float32) if planes is None else np.asarray(planes, np.float32).reshape(-1, 4)
    lv = np.zeros((n, 3), np.float32) if lin is None else np.asarray(lin, np.float32).reshape(n, 3)
    return S.Scene(t, dims, pose, lv, np.zeros((n, 3), np.float32), m, inn, np.zeros(n, np.int32), planes,
                   np.zeros(len(planes), np.int32), 1, "edge", {})


def test_empty_scene_and_planes_only(cuda):
    g = P.Scene(_scene([], np.zeros((0, 4)), np.zeros((0, 3))), position_iterations=4, velocity_iterations=1)
    g.step(DT, 3)
    st = g.stats()
    assert st["bodies"] == 0 and st["constraints"] == 0 and st["bp_pairs"] == 0
    g = P.Scene(_scene([], np.zeros((0, 4)), np.zeros((0, 3)), planes=[[0, 1, 0, 0]]), position_iterations=4,
                velocity_iterations=1)
    g.step(DT, 3)
    assert g.stats()["constraints"] == 0


def test_free_flight_matches_closed_form(cuda):
    """Bodies far apart and above the floor: no pairs, no constraints; after k
    steps v = v0 + g k dt (no damping), positions by the same substeps as the
    oracle."""
    rng = np.random.default_rng(3)
    n = 300
    t = np.where(np.arange(n) % 3 == 0, S.SPHERE, np.where(np.arange(n) % 3 == 1, S.BOX, S.CAPSULE))
    dims = np.zeros((n, 4))
    dims[:, 0] = 0.2
    dims[:, 1] = 0.15
    dims[t == S.BOX, 2] = 0.1
    pos = np.stack([np.arange(n) * 2.0, np.full(n, 50.0), np.zeros(n)], 1)
    lin = rng.uniform(-1, 1, (n, 3))
    sc = _scene(t, dims, pos, planes=[[0, 1, 0, 0]], lin=lin)
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    k = 10
    g.step(DT, k)
    o.step(DT, k)
    gs = g.stats()
    assert gs["bp_pairs"] == 0 and gs["constraints"] == 0
    gp, gl, _ = g.bodies()
    op, ol, _ = o.bodies()
    want = lin + np.array([0.0, -9.81, 0.0]) * DT * k
    np.testing.assert_allclose(gl, want, atol=2e-5)
    np.testing.assert_allclose(gl, ol, atol=1e-6)
    np.testing.assert_allclose(gp[:, :3], op[:, :3], atol=1e-5)


def test_dense_cluster_overflows_pair_buffer(cuda):
    """400 small spheres and boxes packed into a 1.2 m cube: every body
    overlaps dozens of others, far more pairs than the initial buffer of 4 per
    body. The pair set and the contacts still match the oracle, and one step
    moves the bodies as the oracle does."""
    rng = np.random.default_rng(5)
    n = 400
    t = np.where(rng.uniform(size=n) < 0.5, S.SPHERE, S.BOX)
    dims = np.zeros((n, 4))
    dims[:, 0] = rng.uniform(0.1, 0.2, n)
    dims[t == S.BOX, 1] = rng.uniform(0.1, 0.2, int((t == S.BOX).sum()))
    dims[t == S.BOX, 2] = rng.uniform(0.1, 0.2, int((t == S.BOX).sum()))
    pos = rng.uniform(-0.6, 0.6, (n, 3)) + [0.0, 2.0, 0.0]
    sc = _scene(t, dims, pos, planes=[[0, 1, 0, 0]])
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    g.step(DT)
    o.step(DT)
    gs, os_ = g.stats(), o.stats()
    assert gs["bp_pairs"] > 8 * n, gs
    for key in ("bp_pairs", "constraints", "contacts"):
        assert gs[key] == os_[key], (key, gs, os_)
    ckey = lambda c: set(map(tuple, c[:, [0, 1, 3]].tolist()))  # noqa: E731
    assert ckey(g.constraints()) == ckey(o.constraints())
    gp, gl, _ = g.bodies()
    op, ol, _ = o.bodies()
    assert np.isfinite(gl).all()
    # heavily interpenetrating bodies: the colour order decides much of the
    # result, and both sides colour alike, so the velocities still agree closely
    assert np.abs(gl - ol).max() <= 1e-3 * max(1.0, np.abs(ol).max())


def test_bodies_added_after_stepping(cuda):
    """A scene steps, then more bodies join (Scene.add): the step after that
    carries the earlier bodies' friction history and colours and integrates
    everything; the new bodies fall freely until they land."""
    sc1 = S.batched_piles(n_scenes=1, nmin=64, nmax=64, seed=31)
    g = P.Scene(sc1, position_iterations=4, velocity_iterations=1)
    g.step(DT, 20)
    n1 = g.n
    extra = _scene([S.SPHERE] * 4, [[0.2, 0, 0, 0]] * 4, [[40.0 + 3 * i, 30.0, 0.0] for i in range(4)])
    g.add(extra)
    assert g.n == n1 + 4
    g.step(DT, 5)
    _, gl, _ = g.bodies()
    np.testing.assert_allclose(gl[n1:, 1], -9.81 * DT * 5, atol=1e-4)
    assert np.isfinite(gl).all()
