"""Pins the oracle's sleeping (oracle/pxo.c: sleep_check, sleep_islands,
plane_counted) before it checks the GPU: against the live reference PhysX
5.6.1 CPU SDK (oracle/_ref; PxRigidDynamic::setSleepThreshold 0.005,
getWakeCounter, isSleeping) step by step on scenes whose dynamics the oracle
follows to rounding, and against what the sleep rule fixes in closed form."""
import numpy as np
import pytest

import oracle_libs as L
from physx_amd import scenes as S

DT = 1.0 / 60.0
THR = 0.005
RESET = np.float32(20.0) * np.float32(0.02)   # the wake counter reset value, 20 * 0.02 s in float (0.39999998)
need_ref = pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")


def _scene(types, dims, pose, lv=None, planes=((0, 1, 0, 0),), joints=None):
    n = len(types)
    st = np.array(types, np.int32)
    dm = np.array(dims, np.float32)
    m, inn = S.mass_props(st, dm, 1000.0)
    lv = np.zeros((n, 3), np.float32) if lv is None else np.array(lv, np.float32)
    pl = np.array(planes, np.float32).reshape(-1, 4)
    return S.Scene(st, dm, np.array(pose, np.float32), lv, np.zeros((n, 3), np.float32), m, inn,
                   np.zeros(n, np.int32), pl, np.zeros(len(pl), np.int32), 1, "sleep", {}, None, None, joints)


def stack_and_ball(drop_height):
    """two stacked unit boxes, a third box apart, a sphere dropped on the stack"""
    return _scene([S.BOX, S.BOX, S.BOX, S.SPHERE], [[0.5, 0.5, 0.5, 0]] * 3 + [[0.3, 0, 0, 0]],
                  [[0, 0.5, 0, 0, 0, 0, 1], [0, 1.5, 0, 0, 0, 0, 1], [3, 0.5, 0, 0, 0, 0, 1],
                   [0.2, drop_height, 0, 0, 0, 0, 1]])


def resting_box():
    return _scene([S.BOX], [[0.5, 0.5, 0.5, 0]], [[0, 0.5, 0, 0, 0, 0, 1]])


def dropped_box():
    return _scene([S.BOX], [[0.5, 0.5, 0.5, 0]], [[0, 2.0, 0, 0, 0, 0, 1]])


def drifting_box(speed):
    """zero gravity, no plane: a box drifting at a constant speed"""
    return _scene([S.BOX], [[0.5, 0.5, 0.5, 0]], [[0, 5, 0, 0, 0, 0, 1]], lv=[[speed, 0, 0]], planes=())


def held_box():
    """a box held where it is by a fixed joint to the world"""
    pose = [3, 2, 0, 0, 0, 0, 1]
    j = {"type": np.array([S.JOINT_FIXED], np.int32), "body0": np.array([0], np.int32),
         "body1": np.array([-1], np.int32), "frame0": np.array([[0, 0, 0, 0, 0, 0, 1]], np.float32),
         "frame1": np.array([pose], np.float32)}
    return _scene([S.BOX], [[0.2, 0.3, 0.4, 0]], [pose], planes=(), joints=j)


def _trace(world, steps, gravity=True):
    out = []
    for _ in range(steps):
        world.step(DT)
        wc, asleep = world.sleep_state()
        out.append((wc.copy(), asleep.copy(), world.bodies()))
    return out


def _same_traces(a, b, tag, wc_tol=1e-6):
    for s, ((wa, sa, _), (wb, sb, _)) in enumerate(zip(a, b)):
        assert np.array_equal(sa, sb), (tag, s, sa, sb)
        assert np.abs(wa - wb).max() <= wc_tol, (tag, s, wa, wb)


CASES = {
    "resting box": (resting_box, 40, (0.0, -9.81, 0.0)),
    "dropped box": (dropped_box, 120, (0.0, -9.81, 0.0)),
    "drifting box 5 mm/s": (lambda: drifting_box(0.005), 40, (0.0, 0.0, 0.0)),
    "drifting box 5 cm/s": (lambda: drifting_box(0.05), 60, (0.0, 0.0, 0.0)),
    "stack hit by a ball": (lambda: stack_and_ball(3.9), 150, (0.0, -9.81, 0.0)),
    "box held by a joint": (held_box, 40, (0.0, -9.81, 0.0)),
}


@need_ref
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("solver", ["tgs", "pgs"])
def test_sleep_trace_matches_reference(case, solver):
    """Every step's wake counters (within 1e-6 s) and sleeping flags equal the
    reference's: the 0.4 s counter, its countdown, the restarts while the
    accumulated motion stays above the threshold, sleeping one step after the
    counters reach zero, and the wake-up of the ball's island only."""
    make, steps, g = CASES[case]
    if solver == "pgs" and case == "box held by a joint":
        pytest.skip("joints run with TGS only")
    if solver == "pgs" and case == "stack hit by a ball":
        pytest.skip("PGS stacks settle 1e-4 m from the reference's (the Gauss-Seidel order of the stack's two "
                    "contacts differs), so the impact's energy check can fall on the other side")
    sc = make()
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1, gravity=g, solver=solver, sleep_threshold=THR)
    r = L.RefScene(sc, solver, 4, 1, threads=1, gravity=g, sleep_threshold=THR)
    _same_traces(_trace(o, steps), _trace(r, steps), (case, solver))


@need_ref
def test_force_wakes_a_sleeping_box_like_reference():
    """A sleeping box pushed up by a one-step force: it wakes with a 0.4 s
    counter, leaves the ground and lands again, the counter trace and the
    heights equal the reference's (losing the ground contact restarts
    nothing)."""
    sc = resting_box()
    o = L.OracleWorld(sc, sleep_threshold=THR)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1, sleep_threshold=THR)
    F = np.array([[0, sc.mass[0] * 2.0 / DT, 0]], np.float32)
    for s in range(90):
        if s == 30:
            o.set_forces(F)
            r.add_forces(F)
        o.step(DT)
        r.step(DT)
        ow, oa = o.sleep_state()
        rw, ra = r.sleep_state()
        assert np.array_equal(oa, ra) and abs(float(ow[0] - rw[0])) <= 1e-6, (s, ow, rw)
        assert abs(float(o.bodies()[0][0, 1] - r.bodies()[0][0, 1])) < 1e-5, s
    assert oa[0]   # asleep again


def test_sleeping_bodies_keep_their_bits():
    """Closed form: once asleep a body's pose keeps its bits and its velocities
    are zero; the wake counter of a resting body is 20 * 0.02 - k dt rounded
    step by step and reaches 0 after 24 steps (0.4 s); it sleeps at step 25."""
    o = L.OracleWorld(resting_box(), sleep_threshold=THR)
    wc = RESET
    for k in range(24):
        o.step(DT)
        w, a = o.sleep_state()
        wc = max(np.float32(wc - np.float32(DT)), np.float32(0.0))
        assert w[0] == wc and not a[0], (k, w[0], wc)
    assert wc == 0.0
    o.step(DT)
    p0, lv, av = o.bodies()
    assert o.sleep_state()[1][0] and not lv.any() and not av.any()
    for _ in range(30):
        o.step(DT)
    p1, lv, av = o.bodies()
    assert np.array_equal(p0, p1) and not lv.any() and not av.any()


def test_wake_up_names_the_island():
    """PxRigidDynamic::wakeUp on one box of a sleeping stack: the box wakes at
    once with a 0.4 s counter, its stack-mate at the next step (same island);
    the box standing apart stays asleep."""
    o = L.OracleWorld(stack_and_ball(30.0), sleep_threshold=THR)   # the ball is far above
    o.step(DT, 30)
    _, a = o.sleep_state()
    assert a[:3].all()
    o.wake([1])
    w, a = o.sleep_state()
    assert not a[1] and a[0] and a[2] and w[1] == RESET
    o.step(DT)
    w, a = o.sleep_state()
    assert not a[0] and not a[1] and a[2]


def test_worlds_sleep_independently():
    """Two worlds stepped in one oracle world: each world's trace equals that
    world stepped alone (islands never cross worlds)."""
    a, b = stack_and_ball(3.9), dropped_box()
    both = S.Scene(np.concatenate([a.shape_type, b.shape_type]), np.concatenate([a.dims, b.dims]),
                   np.concatenate([a.pose, b.pose]), np.concatenate([a.lin_vel, b.lin_vel]),
                   np.concatenate([a.ang_vel, b.ang_vel]), np.concatenate([a.mass, b.mass]),
                   np.concatenate([a.inertia, b.inertia]), np.array([0, 0, 0, 0, 1], np.int32),
                   np.concatenate([a.planes, b.planes]), np.array([0, 1], np.int32), 2, "two", {})
    o2 = L.OracleWorld(both, sleep_threshold=THR)
    oa = L.OracleWorld(a, sleep_threshold=THR)
    ob = L.OracleWorld(b, sleep_threshold=THR)
    for s in range(120):
        o2.step(DT)
        oa.step(DT)
        ob.step(DT)
        w2, s2 = o2.sleep_state()
        wa, sa = oa.sleep_state()
        wb, sb = ob.sleep_state()
        assert np.array_equal(w2, np.concatenate([wa, wb])) and np.array_equal(s2, np.concatenate([sa, sb])), s


def test_counted_interactions_set_the_threshold():
    """Closed form of the restart value: a box resting on the plane restarts
    with 0.4 + dt * (counted interactions) when its accumulated energy crosses
    the threshold. A box drifting at 5 cm/s in zero gravity with no neighbour
    (0 interactions) restarts at exactly factor * 0.2 = 0.2 s (its energy
    equals the threshold once 0.1 m/s has accumulated: factor 1)."""
    o = L.OracleWorld(drifting_box(0.05), gravity=(0.0, 0.0, 0.0), sleep_threshold=THR)
    seen = []
    for _ in range(40):
        o.step(DT)
        seen.append(float(o.sleep_state()[0][0]))
    assert max(seen[15:]) == pytest.approx(0.2, abs=1e-6)
    assert not o.sleep_state()[1][0]


@need_ref
def test_box_pile_all_asleep_like_reference():
    """A pile of 48 boxes dropped on the ground (chaotic: the piles part by
    centimetres): both end with every box asleep after 400 steps, and half the
    boxes sleep within 30 steps of when half the reference's do."""
    sc = S.box_grid_drop(4, 3, 4, seed=2)
    o = L.OracleWorld(sc, sleep_threshold=THR)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1, sleep_threshold=THR)
    half = [None, None]
    for s in range(400):
        o.step(DT)
        r.step(DT)
        for k, w in enumerate((o, r)):
            if half[k] is None and w.sleep_state()[1].sum() >= sc.n // 2:
                half[k] = s
    assert o.sleep_state()[1].all() and r.sleep_state()[1].all()
    assert half[0] is not None and half[1] is not None and abs(half[0] - half[1]) <= 30, half
