"""Pins the oracle's joints (oracle/pxo.c: joint_rows, prep_joint,
solve_joint) before they check the GPU: against golden vectors the reference
PhysX 5.6.1 CPU SDK wrote for independent joints (tests/golden/joints.npz),
against the live reference on other seeds and on contact-rich jointed
chains, and against closed forms."""
import numpy as np
import pytest

import oracle_libs as L
from physx_amd import scenes as S

GOLDEN = L.os.path.join(L.ROOT, "tests", "golden")
DT = 1.0 / 60.0


def _golden_scene(g):
    n = g["type"].shape[0]
    return S.Scene(g["type"], g["dims"], g["pose"], g["lv"], g["av"], g["mass"], g["inertia"],
                   np.zeros(n, np.int32), np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "joints", {},
                   None, None, {"type": g["jtype"], "body0": g["jbody0"], "body1": g["jbody1"],
                                "frame0": g["jframe0"], "frame1": g["jframe1"]})


def _check_one_step(o_bodies, r_bodies, tag):
    """velocities within 1e-4 of max(|v|, g*dt) (linear) and of max(|w|,
    1 rad/s) (angular); positions within 1e-5 m"""
    op, ol, oa = o_bodies
    rp, rl, ra = r_bodies
    lin_tol = 1e-4 * np.maximum(np.abs(rl).max(1), 9.81 * DT)
    ang_tol = 1e-4 * np.maximum(np.abs(ra).max(1), 1.0)
    bad_l = np.where(np.abs(ol - rl).max(1) > lin_tol)[0]
    bad_a = np.where(np.abs(oa - ra).max(1) > ang_tol)[0]
    assert bad_l.size == 0, (tag, bad_l[:5], ol[bad_l[:3]], rl[bad_l[:3]])
    assert bad_a.size == 0, (tag, bad_a[:5], oa[bad_a[:3]], ra[bad_a[:3]])
    assert np.abs(op[:, :3] - rp[:, :3]).max() <= 1e-5, tag


def test_joints_one_step_vs_reference_golden():
    """Every joint kind, to the world (as body 0 or body 1) and between two
    bodies, each starting 5 cm / 0.1 rad off: one TGS step at 4/1, 8/1 and 1/0
    iterations against the reference SDK's PxFixedJoint / PxSphericalJoint /
    PxRevoluteJoint / PxPrismaticJoint. Each body is in one joint, so the
    result does not depend on the solver's order."""
    g = np.load(L.os.path.join(GOLDEN, "joints.npz"))
    sc = _golden_scene(g)
    dt = float(g["dt"])
    assert set(g["jtype"].tolist()) == {0, 1, 2, 3}
    for k, (p_it, v_it) in enumerate(g["iters"]):
        o = L.OracleWorld(sc, pos_iters=int(p_it), vel_iters=int(v_it))
        o.step(dt)
        _check_one_step(o.bodies(), (g[f"pose1_{k}"], g[f"lv1_{k}"], g[f"av1_{k}"]), (int(p_it), int(v_it)))


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
@pytest.mark.parametrize("seed", [11, 12])
def test_joints_one_step_vs_live_reference(seed):
    sc = S.joint_singles(64, seed)
    for p_it, v_it in ((4, 1), (2, 2)):
        o = L.OracleWorld(sc, pos_iters=p_it, vel_iters=v_it)
        r = L.RefScene(sc, "tgs", p_it, v_it, threads=1)
        o.step(DT)
        r.step(DT)
        _check_one_step(o.bodies(), r.bodies(), (seed, p_it, v_it))


def test_joint_at_rest_stays_at_rest():
    """Closed form: a sphere hung by a spherical joint at its own centre,
    and a box held by a fixed joint exactly where it is, both at rest: the
    joint cancels gravity, so velocities stay zero and the bodies stay put
    (to rounding) over 30 steps."""
    st = np.array([S.SPHERE, S.BOX], np.int32)
    dims = np.array([[0.3, 0, 0, 0], [0.2, 0.3, 0.4, 0]], np.float32)
    q = np.array([0.1, 0.2, -0.3, 0.9], np.float64)
    q /= np.linalg.norm(q)
    pose = np.array([[0, 1, 0, 0, 0, 0, 1], [3, 2, 0, *q]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    ident = [0, 0, 0, 0, 0, 0, 1]
    joints = {"type": np.array([S.JOINT_SPHERICAL, S.JOINT_FIXED], np.int32),
              "body0": np.array([-1, 1], np.int32), "body1": np.array([0, -1], np.int32),
              "frame0": np.array([[0, 1, 0, 0, 0, 0, 1], ident], np.float32),
              "frame1": np.array([ident, pose[1]], np.float32)}
    sc = S.Scene(st, dims, pose, np.zeros((2, 3), np.float32), np.zeros((2, 3), np.float32), m, inn,
                 np.zeros(2, np.int32), np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "rest", {}, None,
                 None, joints)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    for _ in range(30):
        o.step(DT)
    p, lv, av = o.bodies()
    assert np.abs(lv).max() < 1e-4 and np.abs(av).max() < 1e-4, (lv, av)
    assert np.abs(p[:, :3] - pose[:, :3]).max() < 1e-5, p


def test_revolute_pendulum_conserves_hinge_axis():
    """Invariant: a box on a revolute joint about world z, started swinging
    with an angular velocity about z plus an off-axis component: after the
    first step the angular velocity is about z only (the two locked rotation
    axes remove the rest), and the hinge point stays on the anchor (within
    1 mm) over 60 steps."""
    st = np.array([S.BOX], np.int32)
    dims = np.array([[0.5, 0.1, 0.1, 0]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    qz = [0.0, -np.sqrt(0.5), 0.0, np.sqrt(0.5)]      # joint x axis -> world z
    joints = {"type": np.array([S.JOINT_REVOLUTE], np.int32), "body0": np.array([-1], np.int32),
              "body1": np.array([0], np.int32), "frame0": np.array([[0, 2, 0, *qz]], np.float32),
              "frame1": np.array([[-0.5, 0, 0, *qz]], np.float32)}
    sc = S.Scene(st, dims, np.array([[0.5, 2, 0, 0, 0, 0, 1]], np.float32), np.zeros((1, 3), np.float32),
                 np.array([[0.5, 0.3, 2.0]], np.float32), m, inn, np.zeros(1, np.int32), np.zeros((0, 4), np.float32),
                 np.zeros(0, np.int32), 1, "pendulum", {}, None, None, joints)
    o = L.OracleWorld(sc, pos_iters=8, vel_iters=1)
    o.step(DT)
    _, _, av = o.bodies()
    assert abs(av[0, 0]) < 0.02 * abs(av[0, 2]) and abs(av[0, 1]) < 0.02 * abs(av[0, 2]), av
    for _ in range(59):
        o.step(DT)
    p, _, _ = o.bodies()
    x, y, z, w = p[0, 3:]
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                  [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                  [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
    hinge = p[0, :3] + R @ np.array([-0.5, 0, 0])
    assert np.abs(hinge - np.array([0, 2, 0])).max() < 1e-3, hinge


def test_jointed_bodies_do_not_collide():
    """Two overlapping boxes: touching (one constraint) when free, no
    constraint when a joint connects them (PxConstraintFlag::eCOLLISION_ENABLED
    off)."""
    st = np.array([S.BOX, S.BOX], np.int32)
    dims = np.array([[0.3, 0.3, 0.3, 0], [0.3, 0.3, 0.3, 0]], np.float32)
    pose = np.array([[0, 5, 0, 0, 0, 0, 1], [0.5, 5, 0, 0, 0, 0, 1]], np.float32)
    m, inn = S.mass_props(st, dims, 1000.0)
    base = (st, dims, pose, np.zeros((2, 3), np.float32), np.zeros((2, 3), np.float32), m, inn,
            np.zeros(2, np.int32), np.zeros((0, 4), np.float32), np.zeros(0, np.int32), 1, "pair", {}, None, None)
    free = L.OracleWorld(S.Scene(*base))
    free.step(DT)
    assert free.stats()["constraints"] == 1
    j = {"type": np.array([S.JOINT_SPHERICAL], np.int32), "body0": np.array([0], np.int32),
         "body1": np.array([1], np.int32), "frame0": np.array([[0.25, 0, 0, 0, 0, 0, 1]], np.float32),
         "frame1": np.array([[-0.25, 0, 0, 0, 0, 0, 1]], np.float32)}
    joined = L.OracleWorld(S.Scene(*base, j))
    joined.step(DT)
    assert joined.stats()["constraints"] == 0


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_jointed_chains_long_run_vs_live_reference():
    """Chains of 12 boxes (fixed, spherical, revolute and prismatic joints,
    two worlds each) with boxes dropped on them, against the reference SDK
    stepping each world as its own scene. The joints take the reference's
    partition order, so the chains track the reference closely while the
    contacts keep to order-free cases: after 45 steps (the drops have landed)
    every world's mean link height within 5 mm and its largest joint-end gap
    (fixed, spherical, revolute) within 1 mm of the reference's; after 90
    steps, when contact ordering has had time to diverge, within 5 cm and
    1 cm."""
    sc = S.jointed_chains(n_scenes=8, links=12, drops=6, seed=3)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    refs = [L.RefScene(sc.subset(w), "tgs", 4, 1, threads=1) for w in range(8)]
    links = 12
    for steps, h_tol, gap_tol in ((45, 0.005, 0.001), (45, 0.05, 0.01)):
        o.step(DT, steps)
        for r in refs:
            r.step(DT, steps)
        op, _, _ = o.bodies()
        for w in range(8):
            idx = np.where(sc.scene_id == w)[0][:links]
            rp, _, _ = refs[w].bodies()
            ho, hr = op[idx, 1].mean(), rp[:links, 1].mean()
            assert abs(ho - hr) < h_tol, (w, steps, ho, hr)
            if w % 4 == S.JOINT_PRISMATIC:
                continue   # the ends slide apart along the free axis
            go, gr = _max_anchor_gap(op, idx), _max_anchor_gap(rp, np.arange(links))
            assert abs(go - gr) < gap_tol, (w, steps, go, gr)


def _max_anchor_gap(p, idx):
    """largest distance between the joined ends of consecutive links"""
    def rot(q, v):
        u, w = q[:3].astype(np.float64), q[3]
        return v + 2 * w * np.cross(u, v) + 2 * np.cross(u, np.cross(u, v))
    gaps = []
    for a, b in zip(idx[:-1], idx[1:]):
        ea = p[a, :3] + rot(p[a, 3:], np.array([0.25, 0, 0]))
        eb = p[b, :3] + rot(p[b, 3:], np.array([-0.25, 0, 0]))
        gaps.append(np.linalg.norm(ea - eb))
    return max(gaps)
