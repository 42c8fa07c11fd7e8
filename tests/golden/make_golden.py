"""Writes the golden vectors under tests/golden/ from the reference PhysX
5.6.1 CPU SDK, compiled from the reference sources into
oracle/_ref/libpxref.so by oracle/ref.mk (run `make -f oracle/ref.mk -j8`).

The reference ships no test vectors for this path, so these fixtures are the
reference's own outputs on seeded inputs:
  contacts.npz  PxGenerateContacts (fresh PxCache) for sphere-sphere,
                sphere-plane, sphere-box, plane-box and box-box pairs
  broadphase.npz  PxCreateBroadPhase(eABP) created pairs for seeded boxes
  step.npz      one PxScene TGS step (4/1 iterations, dt 1/60) of single bodies
                on a plane: positions, orientations, velocities afterwards
  convex_pairs.npz  PxGenerateContacts (fresh PxCache) on the touching pairs
                of two piles formed in the reference SDK (4,096 24-point
                hulls, and a 12,288-body mix of spheres, capsules, boxes and
                16-point hulls, 100 steps each), 2,000 pairs per class
                (convex-convex, box-convex, sphere-convex, plane-convex,
                capsule-convex, capsule-box) with the reference-cooked hulls
  joints.npz    one PxScene TGS step of independent joints (fixed,
                spherical, revolute, prismatic; one body to the world or two
                bodies to each other, each starting with an error) at 4/1,
                8/1 and 1/0 iterations: the inputs and every body afterwards

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle_libs as L  # noqa: E402
from physx_amd import scenes as S  # noqa: E402

KINDS = ("ss", "sp", "sb", "pb", "bb")
PLANE_POSE = np.array([0, 0, 0, 0, 0, np.sqrt(0.5), np.sqrt(0.5)], np.float32)


def rq(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def qm(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def box_sat_sep(e0, R0, c0, e1, R1, c1):
    A = [R0[:, i] for i in range(3)]
    B = [R1[:, i] for i in range(3)]
    best = -1e9
    for L_ in A + B + [np.cross(a, b) for a in A for b in B]:
        nl = np.linalg.norm(L_)
        if nl < 1e-6:
            continue
        L_ = L_ / nl
        ra = sum(e0[i] * abs(A[i] @ L_) for i in range(3))
        rb = sum(e1[i] * abs(B[i] @ L_) for i in range(3))
        best = max(best, abs((c0 - c1) @ L_) - ra - rb)
    return best


def pair(rng, kind):
    """Seeded pair near contact: (type0, geom0, pose0, type1, geom1, pose1)."""
    g0 = np.zeros(4, np.float32)
    g1 = np.zeros(4, np.float32)
    if kind == "ss":
        g0[0], g1[0] = rng.uniform(.1, .3, 2)
        d = rng.normal(size=3); d /= np.linalg.norm(d)
        p0 = np.r_[d * (g0[0] + g1[0] + rng.uniform(-.05, .035)), 0, 0, 0, 1]
        p1 = np.r_[0, 0, 0, 0, 0, 0, 1]
        return 0, g0, p0, 0, g1, p1
    if kind == "sp":
        g0[0] = rng.uniform(.1, .3)
        p0 = np.r_[rng.uniform(-1, 1), g0[0] + rng.uniform(-.05, .035), rng.uniform(-1, 1), 0, 0, 0, 1]
        return 0, g0, p0, 1, g1, PLANE_POSE
    if kind == "sb":
        g0[0] = rng.uniform(.1, .3)
        g1[:3] = rng.uniform(.1, .3, 3)
        p0 = np.r_[rng.uniform(-.5, .5, 3), 0, 0, 0, 1]
        p1 = np.r_[rng.uniform(-.05, .05, 3), rq(rng)]
        return 0, g0, p0, 3, g1, p1
    if kind == "pb":
        # at most 4 corners within the contact distance (no reduction needed)
        g1[:3] = rng.uniform(.1, .5, 3)
        while True:
            q = rq(rng)
            R = qm(q)
            s = np.array([[a, b, c] for a in (1, -1) for b in (1, -1) for c in (1, -1)], np.float64)
            low = -((R @ (s * g1[:3]).T).T[:, 1]).max()
            y = -low + rng.uniform(-.03, .035)
            h = (R @ (s * g1[:3]).T).T[:, 1] + y
            if (h < 0.04).sum() <= 4:
                break
        return 1, g1 * 0, PLANE_POSE, 3, g1, np.r_[rng.uniform(-1, 1), y, rng.uniform(-1, 1), q]
    g0[:3] = rng.uniform(.1, .3, 3)
    g1[:3] = rng.uniform(.1, .3, 3)
    q0, q1 = rq(rng), rq(rng)
    d = rng.normal(size=3); d /= np.linalg.norm(d)
    target = rng.uniform(-0.03, 0.035)
    lo, hi = 0.0, 2.0
    for _ in range(40):
        mid = 0.5 * (lo + hi)
        if box_sat_sep(g0, qm(q0), d * mid, g1, qm(q1), np.zeros(3)) < target:
            lo = mid
        else:
            hi = mid
    return 3, g0, np.r_[d * lo, q0], 3, g1, np.r_[0, 0, 0, q1]


def make_contacts(n_per_kind=400, seed=1234):
    rng = np.random.default_rng(seed)
    rows = {k: [] for k in ("kind", "t0", "g0", "p0", "t1", "g1", "p1", "count", "contacts")}
    for ki, kind in enumerate(KINDS):
        for _ in range(n_per_kind):
            t0, g0, p0, t1, g1, p1 = pair(rng, kind)
            c = L.contacts("ref", t0, g0, p0, t1, g1, p1, 0.04, 1.0)
            out = np.zeros((4, 8), np.float32)
            out[: c.shape[0]] = c[:4]
            rows["kind"].append(ki); rows["t0"].append(t0); rows["g0"].append(g0); rows["p0"].append(p0)
            rows["t1"].append(t1); rows["g1"].append(g1); rows["p1"].append(p1)
            rows["count"].append(c.shape[0]); rows["contacts"].append(out)
    np.savez_compressed(os.path.join(HERE, "contacts.npz"), **{k: np.array(v) for k, v in rows.items()},
                        contact_distance=np.float32(0.04), tolerance_length=np.float32(1.0))


def make_broadphase(n=3000, seed=99):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0, 15, (n, 3)).astype(np.float32)
    h = rng.uniform(0.1, 0.4, (n, 3)).astype(np.float32)
    b = np.concatenate([c - h, c + h], 1).astype(np.float32)
    d = rng.uniform(0.0, 0.05, n).astype(np.float32)
    pairs = L.bp_pairs_ref(b, d)
    np.savez_compressed(os.path.join(HERE, "broadphase.npz"), bounds=b, distances=d, pairs=pairs.astype(np.uint32))


STEP_CASES = [
    # name, type, dims, pose, lin vel, ang vel
    ("free_fall", 0, [0.3, 0, 0, 0], [0, 5, 0, 0, 0, 0, 1], [0, 0, 0], [0, 0, 0]),
    ("sphere_rest", 0, [0.3, 0, 0, 0], [0, 0.29, 0, 0, 0, 0, 1], [0, 0, 0], [0, 0, 0]),
    ("sphere_roll", 0, [0.3, 0, 0, 0], [0, 0.299, 0, 0, 0, 0, 1], [2, 0, 0], [0, 0, 0]),
    ("box_rest", 3, [0.3, 0.2, 0.25, 0], [0, 0.19, 0, 0, 0, 0, 1], [0, 0, 0], [0, 0, 0]),
    ("box_slide", 3, [0.3, 0.2, 0.25, 0], [0, 0.199, 0, 0, 0, 0, 1], [2, 0, 0.5], [0, 0, 0]),
    ("box_spin", 3, [0.3, 0.2, 0.25, 0], [0, 0.199, 0, 0, 0, 0, 1], [0.5, 0, 0], [0, 3, 1]),
    ("box_tilted", 3, [0.3, 0.2, 0.25, 0], [0, 0.3, 0, 0, 0, np.sin(0.3), np.cos(0.3)], [0, 0, 0], [0, 0, 0]),
]


def make_step(dt=1.0 / 60.0):
    out = {"name": [], "type": [], "dims": [], "pose0": [], "lv0": [], "av0": [], "pose1": [], "lv1": [], "av1": []}
    for name, t, dims, pose, lv, av in STEP_CASES:
        dims = np.array([dims], np.float32)
        m, inertia = S.mass_props(np.array([t]), dims, 1000.0)
        sc = S.Scene(np.array([t], np.int32), dims, np.array([pose], np.float32), np.array([lv], np.float32),
                     np.array([av], np.float32), m, inertia, np.zeros(1, np.int32),
                     np.array([[0, 1, 0, 0]], np.float32), np.zeros(1, np.int32))
        r = L.RefScene(sc, "tgs", 4, 1, threads=1)
        r.step(dt)
        p1, l1, a1 = r.bodies()
        for k, v in (("name", name), ("type", t), ("dims", dims[0]), ("pose0", np.float32(pose)),
                     ("lv0", np.float32(lv)), ("av0", np.float32(av)), ("pose1", p1[0]), ("lv1", l1[0]),
                     ("av1", a1[0])):
            out[k].append(v)
    np.savez_compressed(os.path.join(HERE, "step.npz"), **{k: np.array(v) for k, v in out.items()},
                        dt=np.float32(dt), pos_iters=np.int32(4), vel_iters=np.int32(1))


CONVEX_CLASSES = ("convex-convex", "box-convex", "sphere-convex", "plane-convex", "capsule-convex", "capsule-box")


def _pair_class(ta, tb):
    k = {0: "sphere", 1: "plane", 2: "capsule", 3: "box", 5: "convex"}
    lo, hi = sorted((ta, tb))
    return f"{k[lo]}-{k[hi]}"


def make_convex_pairs(dt=1.0 / 60.0, steps=100, per_class=2000):
    """Touching pairs of piles formed in the reference SDK, regenerated with
    PxGenerateContacts on a fresh cache in the reference's shape order (lower
    geometry type first; two shapes of one type: the later-added body
    first)."""
    hulls, hull_of_ref = [], {}
    rows = {c: [] for c in CONVEX_CLASSES}
    for sc in (S.convex_layers(n_bodies=4096, per_side=32), S.mixed_pile(n_bodies=12288, per_side=32)):
        # reference-cooked meshes of the scene's point clouds
        counts = sc.hull_counts
        offs = np.concatenate([[0], np.cumsum(counts)])
        mesh = []
        for k in range(len(counts)):
            mid, h = L.ref_cook(sc.hull_points[offs[k]:offs[k + 1]])
            hull_of_ref[mid] = len(hulls)
            hulls.append(h)
            mesh.append(mid)
        r = L.RefScene(sc, "tgs", 4, 1, threads=8, report_contacts=True)
        r.step(dt, steps)
        pose, _, _ = r.bodies()
        pairs = {(int(a), int(b)) for a, b in r.contacts()[:, :2]}
        for a, b in sorted(pairs):
            i = a - 1
            if b < 0:
                ti, tj = int(sc.shape_type[i]), 1
            else:
                j = b - 1
                ti, tj = int(sc.shape_type[i]), int(sc.shape_type[j])
            cls = _pair_class(ti, tj)
            if cls not in rows or len(rows[cls]) >= per_class:
                continue

            def shape(body):
                t = int(sc.shape_type[body])
                g = sc.dims[body].copy()
                if t == 5:
                    g = np.array([mesh[int(g[0])], 0, 0, 0], np.float32)
                return t, g, pose[body]
            if b < 0:
                s0 = (1, np.zeros(4, np.float32), PLANE_POSE)
                s1 = shape(i)
            else:
                j = b - 1
                if ti == tj:
                    s0, s1 = shape(max(i, j)), shape(min(i, j))
                elif ti < tj:
                    s0, s1 = shape(i), shape(j)
                else:
                    s0, s1 = shape(j), shape(i)
            c = L.contacts("ref", s0[0], s0[1], s0[2], s1[0], s1[1], s1[2])
            if len(c) == 0:
                continue
            out = np.zeros((4, 8), np.float32)
            out[:len(c)] = c[:4]
            g0, g1 = s0[1].copy(), s1[1].copy()
            if s0[0] == 5:
                g0[0] = hull_of_ref[int(g0[0])]
            if s1[0] == 5:
                g1[0] = hull_of_ref[int(g1[0])]
            rows[cls].append((s0[0], g0, s0[2], s1[0], g1, s1[2], len(c), out))
    # keep the hulls the stored pairs use, renumbered in order of use
    used = {}
    for rr in rows.values():
        for x in rr:
            for t, g in ((x[0], x[1]), (x[3], x[4])):
                if t == 5:
                    g[0] = used.setdefault(int(g[0]), len(used))
    hulls = [h for _, h in sorted(((v, hulls[k]) for k, v in used.items()), key=lambda kv: kv[0])]
    out = {}
    for cls, rr in rows.items():
        key = cls.replace("-", "_")
        out[f"{key}_t"] = np.array([[x[0], x[3]] for x in rr], np.int32).reshape(-1, 2)
        out[f"{key}_g"] = np.array([np.concatenate([x[1], x[4]]) for x in rr], np.float32).reshape(-1, 8)
        out[f"{key}_p"] = np.array([np.concatenate([x[2], x[5]]) for x in rr], np.float32).reshape(-1, 14)
        out[f"{key}_n"] = np.array([x[6] for x in rr], np.int32)
        out[f"{key}_c"] = np.array([x[7] for x in rr], np.float32).reshape(-1, 4, 8)
    # the hulls, concatenated
    nv = np.array([h["verts"].shape[0] for h in hulls], np.int32)
    npl = np.array([h["planes"].shape[0] for h in hulls], np.int32)
    nr = np.array([h["vrefs"].shape[0] for h in hulls], np.int32)
    out.update(hull_nv=nv, hull_np=npl, hull_nr=nr,
               hull_verts=np.concatenate([h["verts"] for h in hulls]).astype(np.float32),
               hull_planes=np.concatenate([h["planes"] for h in hulls]).astype(np.float32),
               hull_polys=np.concatenate([h["polys"] for h in hulls]).astype(np.uint16),
               hull_vrefs=np.concatenate([h["vrefs"] for h in hulls]).astype(np.uint8),
               hull_ie=np.array([h["internal_extents"] for h in hulls], np.float32),
               hull_com=np.array([h["com"] for h in hulls], np.float32))
    np.savez_compressed(os.path.join(HERE, "convex_pairs.npz"), **out)
    return {c: len(r) for c, r in rows.items()}


JOINT_ITERS = ((4, 1), (8, 1), (1, 0))


def make_joints(dt=1.0 / 60.0, n_groups=64, seed=5):
    sc = S.joint_singles(n_groups, seed)
    out = {"type": sc.shape_type, "dims": sc.dims, "pose": sc.pose, "lv": sc.lin_vel, "av": sc.ang_vel,
           "mass": sc.mass, "inertia": sc.inertia, "jtype": sc.joints["type"], "jbody0": sc.joints["body0"],
           "jbody1": sc.joints["body1"], "jframe0": sc.joints["frame0"], "jframe1": sc.joints["frame1"],
           "iters": np.array(JOINT_ITERS, np.int32), "dt": np.float32(dt)}
    for k, (p_it, v_it) in enumerate(JOINT_ITERS):
        r = L.RefScene(sc, "tgs", p_it, v_it, threads=1)
        r.step(dt)
        p1, l1, a1 = r.bodies()
        out[f"pose1_{k}"], out[f"lv1_{k}"], out[f"av1_{k}"] = p1, l1, a1
    np.savez_compressed(os.path.join(HERE, "joints.npz"), **out)


if __name__ == "__main__":
    assert L.have_ref(), "build oracle/_ref/libpxref.so first (make -f oracle/ref.mk -j8)"
    make_contacts()
    make_broadphase()
    make_step()
    make_joints()
    print(make_convex_pairs())
    for f in ("contacts.npz", "broadphase.npz", "step.npz", "joints.npz", "convex_pairs.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)))
