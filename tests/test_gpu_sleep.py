"""GPU sleeping (enable_sleeping: wake counters, islands, wake-up and
deactivation) against the oracle, which tests/test_cpu_sleep.py pins to the
reference SDK: every step's sleeping flags and wake counters on single
bodies, stacks hit by balls, jointed bodies and batches of worlds; sleeping
bodies keep their bits; wake-ups through pxh_scene_wake_bodies and
pxh_scene_set_rigid_dynamic_data; world removal; a world's bits alone and in
a batch; argument checks."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L
from physx_amd import scenes as S
from test_cpu_sleep import CASES, RESET, THR, stack_and_ball, resting_box

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def gpu_scene(sc, solver="tgs", gravity=(0.0, -9.81, 0.0), **kw):
    return P.Scene(sc, position_iterations=4, velocity_iterations=1, gravity=gravity,
                   solver_type=P.PXH_SOLVER_TGS if solver == "tgs" else P.PXH_SOLVER_PGS, enable_sleeping=1,
                   sleep_threshold=THR, **kw)


def _compare(g, o, tag, wc_tol=1e-5, pos_tol=1e-5):
    """flags equal; wake counters within 1e-5 s (a restart sets the counter in
    proportion to the accumulated energy, which carries the velocities'
    relative rounding of about 1e-6 between the GPU and the C oracle);
    positions within 1e-5 m"""
    gw, ga = g.sleep_state()
    ow, oa = o.sleep_state()
    assert np.array_equal(ga, oa), (tag, np.where(ga != oa)[0][:8])
    assert np.abs(gw - ow).max() <= wc_tol, (tag, np.abs(gw - ow).max())
    gp, gl, gv = g.bodies()
    op, ol, ov = o.bodies()
    assert np.abs(gp[:, :3] - op[:, :3]).max() <= pos_tol, (tag, np.abs(gp[:, :3] - op[:, :3]).max())
    # sleeping bodies: zero velocities on both sides
    assert not gl[ga].any() and not gv[ga].any(), tag


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("solver", ["tgs", "pgs"])
def test_sleep_trace_matches_oracle(cuda, case, solver):
    make, steps, g = CASES[case]
    if solver == "pgs" and case == "box held by a joint":
        pytest.skip("joints run with TGS only")
    sc = make()
    gs = gpu_scene(sc, solver, g)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1, gravity=g, solver=solver, sleep_threshold=THR)
    for s in range(steps):
        gs.step(DT)
        o.step(DT)
        _compare(gs, o, (case, solver, s))


def _batch(heights):
    """one stack-and-ball world per drop height"""
    worlds = [stack_and_ball(h) for h in heights]
    cat = lambda f: np.concatenate([getattr(w, f) for w in worlds])  # noqa: E731
    n = worlds[0].n
    return S.Scene(cat("shape_type"), cat("dims"), cat("pose"), cat("lin_vel"), cat("ang_vel"), cat("mass"),
                   cat("inertia"), np.repeat(np.arange(len(worlds), dtype=np.int32), n), cat("planes"),
                   np.arange(len(worlds), dtype=np.int32), len(worlds), "stacks", {})


def test_batched_stacks_track_oracle(cuda):
    """64 worlds, a ball dropped on a sleeping stack from a different height in
    each: every step's flags and wake counters equal the oracle's; each ball
    wakes only its own stack."""
    heights = 3.0 + 0.05 * np.arange(64)
    sc = _batch(heights)
    g = gpu_scene(sc)
    o = L.OracleWorld(sc, sleep_threshold=THR)
    woke_apart = False
    for s in range(160):
        g.step(DT)
        o.step(DT)
        _compare(g, o, s)
        _, a = g.sleep_state()
        woke_apart |= not a.reshape(64, 4)[:, 2].all() and s > 30
    assert not woke_apart


def test_sleeping_bits_frozen_and_wake_bodies(cuda):
    """Once asleep the poses keep their bits over 40 steps; wake_bodies on one
    box of a stack wakes it at once (wake counter 20 * 0.02 s) and its
    stack-mate at the next step, like the oracle; the box apart sleeps on."""
    sc = stack_and_ball(30.0)
    g = gpu_scene(sc)
    o = L.OracleWorld(sc, sleep_threshold=THR)
    g.step(DT, 30)
    o.step(DT, 30)
    p0 = g.bodies()[0]
    assert g.sleep_state()[1][:3].all()
    for _ in range(10):
        g.step(DT)
        o.step(DT)
    assert np.array_equal(g.bodies()[0][:3], p0[:3])
    g.wake_bodies([1])
    o.wake([1])
    w, a = g.sleep_state()
    assert not a[1] and a[0] and a[2] and w[1] == RESET
    for s in range(60):
        g.step(DT)
        o.step(DT)
        _compare(g, o, s)


def test_set_velocity_wakes(cuda):
    """Writing a velocity through the Direct GPU API wakes a sleeping box (the
    oracle's set_bodies wakes every written body the same way)."""
    import torch
    sc = resting_box()
    g = gpu_scene(sc)
    o = L.OracleWorld(sc, sleep_threshold=THR)
    g.step(DT, 30)
    o.step(DT, 30)
    assert g.sleep_state()[1][0] and o.sleep_state()[1][0]
    v = torch.tensor([[0.0, 2.0, 0.0]], dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    g.set_rigid_dynamic_data(v, P.PXH_DATA_LINEAR_VELOCITY)
    o.set_bodies(None, np.array([[0.0, 2.0, 0.0]], np.float32), None)
    assert not g.sleep_state()[1][0] and g.sleep_state()[0][0] == RESET
    for s in range(90):
        g.step(DT)
        o.step(DT)
        _compare(g, o, s)
    assert g.sleep_state()[1][0]


def test_world_alone_and_in_batch_and_after_removal(cuda):
    """A world's bits and sleep state stepped alone equal its slice of a batch
    of eight, and stay equal after another world is removed (island labels
    renumbered with the bodies)."""
    sc = _batch(3.0 + 0.1 * np.arange(8))
    alone = gpu_scene(sc.subset(5))
    batch = gpu_scene(sc)
    for _ in range(40):
        alone.step(DT)
        batch.step(DT)
    f, c = batch.world_range(5)
    for x, y in zip(alone.bodies(), batch.bodies()):
        assert np.array_equal(x, y[f:f + c])
    batch.remove_worlds([2])
    f2, c2 = batch.world_range(4)
    for _ in range(80):
        alone.step(DT)
        batch.step(DT)
    for x, y in zip(alone.bodies(), batch.bodies()):
        assert np.array_equal(x, y[f2:f2 + c2])
    wa, sa = alone.sleep_state()
    wb, sb = batch.sleep_state()
    assert np.array_equal(wa, wb[f2:f2 + c2]) and np.array_equal(sa, sb[f2:f2 + c2])


def test_box_pile_falls_asleep(cuda):
    """A two-layer pile of 128 boxes: every box asleep by step 300 (the oracle
    and the reference SDK both sleep it whole by step 200) and the GPU's
    sleeping count within 20% of the oracle's at steps 60, 100 and 150 (the
    piles part chaotically)."""
    sc = S.box_grid_drop(8, 2, 8, seed=2)
    g = gpu_scene(sc)
    o = L.OracleWorld(sc, sleep_threshold=THR)
    for s in range(300):
        g.step(DT)
        o.step(DT)
        if s in (59, 99, 149):
            ng, no = int(g.sleep_state()[1].sum()), int(o.sleep_state()[1].sum())
            assert abs(ng - no) <= 0.2 * sc.n, (s, ng, no)
    assert g.sleep_state()[1].all() and o.sleep_state()[1].all()


def test_sleep_arguments(cuda):
    sc = resting_box()
    with pytest.raises(P.PxhError, match="positive sleep threshold"):
        P.Scene(sc, enable_sleeping=1, sleep_threshold=0.0)
    with pytest.raises(P.PxhError, match="touch events"):
        P.Scene(sc, enable_sleeping=1, report_touch_events=1)
    g = P.Scene(sc)
    with pytest.raises(P.PxhError, match="sleeping is off"):
        g.sleep_state()
    g2 = gpu_scene(sc)
    with pytest.raises(P.PxhError, match="out of range"):
        g2.wake_bodies([5])
