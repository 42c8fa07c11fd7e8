"""Summary statistics of a settling pile (floor plane y = 0), for comparing
long runs whose trajectories diverge chaotically but whose statistics agree."""
import numpy as np


def lowest_points(shape_type, dims, pose, scene=None):
    """Height of the lowest point of each sphere (type 0), capsule (type 2,
    radius and half height along the body x axis), box (type 3) or convex
    mesh (type 5, vertices from scene.hull_points)."""
    x, y, z, w = pose[:, 3], pose[:, 4], pose[:, 5], pose[:, 6]
    # second row of the rotation matrix: world-y component of each body axis
    ry = np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1)
    reach = (np.abs(ry) * dims[:, :3]).sum(1)
    reach = np.where(shape_type == 0, dims[:, 0], reach)
    reach = np.where(shape_type == 2, np.abs(ry[:, 0]) * dims[:, 1] + dims[:, 0], reach)
    cvx = np.nonzero(shape_type == 5)[0]
    if len(cvx):
        off = np.concatenate([[0], np.cumsum(scene.hull_counts)])
        for i in cvx:
            h = int(dims[i, 0])
            v = scene.hull_points[off[h]:off[h + 1]]
            reach[i] = -(v @ ry[i]).min()
    return pose[:, 1] - reach


def pile_stats(scene, pose, lin_vel, rest_speed=0.1):
    """height: mean body height; potential: total potential energy;
    penetration: worst floor penetration among resting bodies (speed below
    `rest_speed`); jitter: mean speed of the resting bodies."""
    low = lowest_points(scene.shape_type, scene.dims, pose, scene)
    speed = np.linalg.norm(lin_vel, axis=1)
    rest = speed < rest_speed
    return {
        "height": float(pose[:, 1].mean()),
        "potential": float((scene.mass * 9.81 * pose[:, 1]).sum()),
        "penetration": float(max(0.0, -low[rest].min())) if rest.any() else 0.0,
        "jitter": float(speed[rest].mean()) if rest.any() else 0.0,
    }


def run_stats(world, scene, steps, dt=1.0 / 60.0, tail=60):
    """Steps `world` (anything with step(dt) and bodies()) and returns the final
    height and potential energy, with the resting penetration and jitter
    averaged over the last `tail` steps (transient impacts of bodies still
    tumbling off the pile are left out: they are single chaotic events)."""
    pen, jit = [], []
    for s in range(steps):
        world.step(dt)
        if s >= steps - tail:
            pose, lv, _ = world.bodies()
            st = pile_stats(scene, pose, lv)
            pen.append(st["penetration"])
            jit.append(st["jitter"])
    st["penetration"] = float(np.mean(pen))
    st["jitter"] = float(np.mean(jit))
    return st


def assert_stats_close(a, b, rel=0.05, pen_abs=3e-4, jitter_rel=0.1):
    """Height and potential energy within `rel`; resting penetration within
    `pen_abs` metres (about a third of its typical 0.8 mm; per-seed spread in
    the reference itself is of that order); resting jitter within
    `jitter_rel`."""
    assert abs(a["height"] - b["height"]) <= rel * abs(b["height"]), (a, b)
    assert abs(a["potential"] - b["potential"]) <= rel * abs(b["potential"]), (a, b)
    assert abs(a["penetration"] - b["penetration"]) <= pen_abs, (a, b)
    assert abs(a["jitter"] - b["jitter"]) <= jitter_rel * abs(b["jitter"]), (a, b)
