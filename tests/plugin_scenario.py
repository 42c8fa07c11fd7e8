"""Incremental broadphase updates in the SDK's form (PxBroadPhaseUpdateData):
persistent bounds / groups / distances arrays indexed by handle, and per
update the created, updated and removed handles. Statics, kinematics and
dynamics (some dynamics sharing an actor's group, so they never pair), with
removals whose pairs must not be reported as deleted and handles freed by
removal reused later."""
import numpy as np


def updates(lib, seed=3, cap=6000, n0=4000, rounds=4):
    """Yields (created, updated, removed, bounds, groups, distances) per update;
    lib: the harness library (for the SDK's filter-group encodings)."""
    rng = np.random.default_rng(seed)
    bounds = np.zeros((cap, 6), np.float32)
    groups = np.zeros(cap, np.uint32)
    dist = np.zeros(cap, np.float32)
    live = np.zeros(cap, bool)
    actor = [0]

    def place(h):
        c = rng.uniform(0.0, 22.0, (len(h), 3))
        e = rng.uniform(0.1, 0.6, (len(h), 3))
        bounds[h, :3] = c - e
        bounds[h, 3:] = c + e
        dist[h] = rng.uniform(0.0, 0.05, len(h))
        for i in h:
            kind = rng.random()
            if kind < 0.1:
                groups[i] = lib.pxrefgpu_static_group()
            elif kind < 0.2:
                actor[0] += 1
                groups[i] = lib.pxrefgpu_kinematic_group(actor[0])
            elif kind < 0.3 and actor[0] > 0:
                groups[i] = groups[i - 1] if i > 0 and live[i - 1] else lib.pxrefgpu_dynamic_group(actor[0])
            else:
                actor[0] += 1
                groups[i] = lib.pxrefgpu_dynamic_group(actor[0])

    created = np.arange(n0, dtype=np.uint32)
    place(created)
    live[created] = True
    yield created, np.zeros(0, np.uint32), np.zeros(0, np.uint32), bounds.copy(), groups.copy(), dist.copy()
    for r in range(rounds - 1):
        alive = np.nonzero(live)[0]
        removed = rng.choice(alive, size=len(alive) // 12, replace=False).astype(np.uint32)
        live[removed] = False
        alive = np.nonzero(live)[0]
        moved = rng.choice(alive, size=len(alive) // 3, replace=False).astype(np.uint32)
        bounds[moved] += rng.uniform(-0.4, 0.4, (len(moved), 1)).astype(np.float32)
        # new objects in free handles, not those removed in this same update
        free = np.setdiff1d(np.nonzero(~live)[0], removed)
        created = np.sort(rng.choice(free, size=min(len(free), 400), replace=False)).astype(np.uint32)
        place(created)
        live[created] = True
        yield created, np.sort(moved), np.sort(removed), bounds.copy(), groups.copy(), dist.copy()
