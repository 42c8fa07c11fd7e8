"""GPU joints (pxh_scene_add_joints: fixed, spherical, revolute, prismatic;
TGS) against the oracle, which tests/test_cpu_joints.py pins to the
reference SDK's joints: one step of independent joints; jointed chains with
boxes dropped on them, one step from a mid-run state and many steps side by
side; jointed bodies never collide; the per-world bits of a world stepped
alone and in a batch; world removal renumbers the joints; PGS with joints is
refused."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L
from physx_amd import scenes as S

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def _close(g, o, rel, tag):
    gp, gl, ga = g
    op, ol, oa = o
    lin_tol = rel * np.maximum(np.abs(ol).max(1), 9.81 * DT)
    ang_tol = rel * np.maximum(np.abs(oa).max(1), 1.0)
    bl = np.where(np.abs(gl - ol).max(1) > lin_tol)[0]
    ba = np.where(np.abs(ga - oa).max(1) > ang_tol)[0]
    assert bl.size == 0, (tag, bl[:5], gl[bl[:3]], ol[bl[:3]])
    assert ba.size == 0, (tag, ba[:5], ga[ba[:3]], oa[ba[:3]])


@pytest.mark.parametrize("iters", [(4, 1), (8, 1), (1, 0)])
def test_joint_singles_one_step_match_oracle_and_reference(cuda, iters):
    """Every joint kind, to the world as either body and between two bodies,
    each starting 5 cm / 0.1 rad off: one step within 1e-5 of the oracle and
    within 1e-4 of the reference SDK's golden step (tests/golden/joints.npz)."""
    g = np.load(L.os.path.join(L.ROOT, "tests", "golden", "joints.npz"))
    n = g["type"].shape[0]
    sc = S.Scene(g["type"], g["dims"], g["pose"], g["lv"], g["av"], g["mass"], g["inertia"],
                 np.zeros(n, np.int32), np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "joints", {},
                 None, None, {"type": g["jtype"], "body0": g["jbody0"], "body1": g["jbody1"],
                              "frame0": g["jframe0"], "frame1": g["jframe1"]})
    k = [tuple(x) for x in g["iters"].tolist()].index(iters)
    gpu = P.Scene(sc, position_iterations=iters[0], velocity_iterations=iters[1])
    ora = L.OracleWorld(sc, pos_iters=iters[0], vel_iters=iters[1])
    gpu.step(DT)
    ora.step(DT)
    _close(gpu.bodies(), ora.bodies(), 1e-5, ("oracle", iters))
    _close(gpu.bodies(), (g[f"pose1_{k}"], g[f"lv1_{k}"], g[f"av1_{k}"]), 1e-4, ("reference", iters))


def test_jointed_chains_one_step_from_mid_run_state(cuda):
    """Chains of 12 boxes of every joint kind with boxes dropped on them
    (8 worlds): the GPU runs 40 steps, then the GPU and the oracle each take
    one step from that state (fresh scenes: no carried friction patches):
    the same constraint table, velocities within 1e-5 of max(|v|, g dt)."""
    sc = S.jointed_chains(n_scenes=8, links=12, drops=6, seed=3)
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    g.step(DT, 40)
    pose, lv, av = g.bodies()
    g.close()
    st = S.Scene(sc.shape_type, sc.dims, pose, lv, av, sc.mass, sc.inertia, sc.scene_id, sc.planes, sc.plane_scene,
                 sc.n_scenes, sc.name, {}, None, None, sc.joints)
    g = P.Scene(st, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(st, pos_iters=4, vel_iters=1)
    g.step(DT)
    o.step(DT)
    gs, os_ = g.stats(), o.stats()
    assert gs["constraints"] == os_["constraints"] > 20, (gs, os_)
    assert gs["contacts"] == os_["contacts"]
    gc = g.constraints()
    oc = o.constraints()
    assert sorted(map(tuple, gc[:, :2].tolist())) == sorted(map(tuple, oc[:, :2].tolist()))
    _close(g.bodies(), o.bodies(), 1e-5, "chains")


@pytest.mark.parametrize("drops,steps", [(6, 40), (0, 60)])
def test_jointed_chains_track_oracle(cuda, drops, steps):
    """The chains stepped side by side on the GPU and in the oracle: every
    step's constraint count equal and the link positions within 1 mm at the
    end. The GPU rounds a few operations apart from the C oracle (sincosf;
    the square-root inverse inertia kept as a symmetric matrix); with boxes
    landing on swinging chains the run turns chaotic after step 40 (in the
    oracle alone, moving every body by 1e-6 m grows to a 1.6 cm difference by
    step 50, when one pair touches on one side only), so that run is
    compared over its first 40 steps; without the boxes over 60 steps (the
    same perturbation stays below 4e-6 m)."""
    sc = S.jointed_chains(n_scenes=8, links=12, drops=drops, seed=3)
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    for s in range(steps):
        g.step(DT)
        o.step(DT)
        assert g.stats()["constraints"] == o.stats()["constraints"], s
    gp, _, _ = g.bodies()
    op, _, _ = o.bodies()
    links = np.concatenate([np.where(sc.scene_id == w)[0][:12] for w in range(8)])
    d = float(np.abs(gp[links, :3] - op[links, :3]).max())
    print(f"largest link position difference after {steps} steps: {d:.2e}")
    assert d < 1e-3
    assert np.isfinite(gp).all()


def test_jointed_bodies_do_not_collide(cuda):
    """Overlapping boxes: a contact when free, none once jointed (the pair
    reaches the broadphase but its contacts are dropped); the pair still
    counts among the broadphase pairs, as in the reference."""
    st = np.array([S.BOX, S.BOX], np.int32)
    dims = np.array([[0.3, 0.3, 0.3, 0], [0.3, 0.3, 0.3, 0]], np.float32)
    pose = np.array([[0, 5, 0, 0, 0, 0, 1], [0.5, 5, 0, 0, 0, 0, 1]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    base = (st, dims, pose, np.zeros((2, 3), np.float32), np.zeros((2, 3), np.float32), m, inn,
            np.zeros(2, np.int32), np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "pair", {}, None, None)
    free = P.Scene(S.Scene(*base))
    free.step(DT)
    assert free.stats()["constraints"] == 1
    j = {"type": np.array([S.JOINT_SPHERICAL], np.int32), "body0": np.array([0], np.int32),
         "body1": np.array([1], np.int32), "frame0": np.array([[0.25, 0, 0, 0, 0, 0, 1]], np.float32),
         "frame1": np.array([[-0.25, 0, 0, 0, 0, 0, 1]], np.float32)}
    joined = P.Scene(S.Scene(*base, j))
    joined.step(DT)
    assert joined.stats()["constraints"] == 0
    assert joined.stats()["bp_pairs"] == 1


def test_jointed_world_alone_and_in_batch_bit_identical(cuda):
    """A jointed world gives the same bits stepped alone and inside a batch
    of eight, and after another world is removed (joints renumbered)."""
    sc = S.jointed_chains(n_scenes=8, links=12, drops=6, seed=3)
    alone = P.Scene(sc.subset(5), position_iterations=4, velocity_iterations=1)
    batch = P.Scene(sc, position_iterations=4, velocity_iterations=1)
    for _ in range(20):
        alone.step(DT)
        batch.step(DT)
    f, c = batch.world_range(5)
    a = alone.bodies()
    b = batch.bodies()
    for x, y in zip(a, b):
        assert np.array_equal(x, y[f:f + c])
    batch.remove_worlds([2])
    f2, c2 = batch.world_range(4)   # world 5 is now world 4
    assert c2 == c
    for _ in range(10):
        alone.step(DT)
        batch.step(DT)
    a = alone.bodies()
    b = batch.bodies()
    for x, y in zip(a, b):
        assert np.array_equal(x, y[f2:f2 + c2])


def test_joints_refuse_pgs_and_bad_arguments(cuda):
    sc = S.jointed_chains(n_scenes=1, links=3, drops=0, seed=3)
    g = P.Scene(sc, position_iterations=4, velocity_iterations=1, solver_type=P.PXH_SOLVER_PGS)
    with pytest.raises(P.PxhError, match="TGS"):
        g.simulate(DT)
    g2 = P.Scene(position_iterations=4, velocity_iterations=1)
    g2.add(S.jointed_chains(n_scenes=1, links=2, drops=0, seed=3))
    with pytest.raises(P.PxhError, match="out of range"):
        g2.add_joints([0], [0], [99], np.zeros((1, 7)), np.zeros((1, 7)))
    with pytest.raises(P.PxhError, match="different actors"):
        g2.add_joints([0], [-1], [-1], np.zeros((1, 7)), np.zeros((1, 7)))
