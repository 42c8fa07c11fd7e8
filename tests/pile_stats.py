"""Summary statistics of a settling pile (floor plane y = 0), for comparing
long runs whose trajectories diverge chaotically but whose statistics agree."""
import numpy as np


def lowest_points(shape_type, dims, pose):
    """Height of the lowest point of each sphere (type 0) or box (type 3)."""
    x, y, z, w = pose[:, 3], pose[:, 4], pose[:, 5], pose[:, 6]
    # second row of the rotation matrix: world-y component of each body axis
    ry = np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], 1)
    reach = (np.abs(ry) * dims[:, :3]).sum(1)
    reach = np.where(shape_type == 0, dims[:, 0], reach)
    return pose[:, 1] - reach


def pile_stats(scene, pose, lin_vel):
    low = lowest_points(scene.shape_type, scene.dims, pose)
    return {
        "height": float(pose[:, 1].mean()),
        "potential": float((scene.mass * 9.81 * pose[:, 1]).sum()),
        "penetration": float(max(0.0, -low.min())),
        "speed": float(np.linalg.norm(lin_vel, axis=1).mean()),
    }


def run_stats(world, scene, steps, dt=1.0 / 60.0, tail=60):
    """Steps `world` (anything with step(dt) and bodies()) and returns the final
    height and potential energy with the worst penetration and the mean speed
    over the last `tail` steps."""
    pen, speed = 0.0, []
    for s in range(steps):
        world.step(dt)
        if s >= steps - tail:
            pose, lv, _ = world.bodies()
            st = pile_stats(scene, pose, lv)
            pen = max(pen, st["penetration"])
            speed.append(st["speed"])
    st["penetration"] = pen
    st["speed"] = float(np.mean(speed))
    return st


def assert_stats_close(a, b, rel=0.05, pen_abs=1e-3, speed_abs=0.02):
    """Height and potential energy within `rel`; worst penetration within
    `pen_abs` metres; mean speed within `speed_abs` m/s."""
    assert abs(a["height"] - b["height"]) <= rel * abs(b["height"]), (a, b)
    assert abs(a["potential"] - b["potential"]) <= rel * abs(b["potential"]), (a, b)
    assert abs(a["penetration"] - b["penetration"]) <= pen_abs, (a, b)
    assert abs(a["speed"] - b["speed"]) <= speed_abs, (a, b)
