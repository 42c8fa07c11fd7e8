"""Convex meshes on the CPU: the library's hull cooking against the
reference cooker, and the oracle's convex contact generation against the
reference's PxGenerateContacts (both built from source in oracle/_ref).

Contacts are compared for pairs placed at realistic separations (deepest
point no more than 5 cm inside): the reference's GJK/EPA stop at relative
tolerances, the oracle uses the exact closest features / minimum translation,
so for a small share of pairs the two pick a neighbouring witness face; the
tests bound that share."""
import numpy as np
import pytest

import oracle_libs as L
import physx_amd as P
from physx_amd import scenes

pytestmark = pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")

CD, TOL = 0.04, 1.0


def _cloud(rng, n=24):
    ax = rng.uniform(0.2, 0.6, 3)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * ax).astype(np.float32)


def _quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _poly_sets(h):
    return sorted(tuple(sorted(map(tuple, h["verts"][h["vrefs"][f:f + c]].tolist()))) for f, c in h["polys"])


def test_cooked_hull_matches_reference_cooker():
    """Same vertices bit for bit; the same polygons, adjacent triangles within
    3 degrees merged as the reference's quickhull post-merge does, for all but
    a few hulls (where one borderline merge goes the other way: the merged
    plane's normal is summed from a different leading edge); mass properties
    and internal objects as the reference computes them."""
    rng = np.random.default_rng(11)
    exact = same_polys = 0
    n = 80
    for t in range(n):
        pts = _cloud(rng)
        mine = P.cook_convex(pts)
        _, ref = L.ref_cook(pts)
        assert set(map(tuple, mine["verts"].tolist())) == set(map(tuple, ref["verts"].tolist()))
        if _poly_sets(mine) != _poly_sets(ref):
            continue
        same_polys += 1
        # mass properties: the library integrates the hull's triangles, the
        # reference its (merged, slightly non-planar) polygons, so hulls with
        # merged faces agree less tightly
        merged = ref["poly_counts"].max() > 3
        assert abs(mine["volume"] - ref["mass"]) <= (2e-3 if merged else 1e-5) * ref["mass"]
        np.testing.assert_allclose(mine["com"], ref["com"], atol=2e-3 if merged else 1e-5)
        # the reference reports the inertia about the mesh origin
        c = mine["com"]
        origin = mine["inertia"] + mine["volume"] * (np.dot(c, c) * np.eye(3) - np.outer(c, c))
        np.testing.assert_allclose(origin, ref["inertia_origin"],
                                   atol=(1e-2 if merged else 1e-5) * np.abs(ref["inertia_origin"]).max())
        np.testing.assert_array_equal(mine["bounds_center"], ref["bounds_center"])
        np.testing.assert_array_equal(mine["bounds_extents"], ref["bounds_extents"])
        if ref["poly_counts"].max() == 3:
            # no merged faces: planes agree to float rounding, and so do the
            # internal objects that set the margins
            exact += 1
            for pl, (f, cnt) in zip(ref["planes"], ref["polys"]):
                key = sorted(map(tuple, ref["verts"][ref["vrefs"][f:f + cnt]].tolist()))
                j = _poly_sets(mine).index(tuple(key))
                mp = [mine["planes"][k] for k, (ff, cc) in enumerate(mine["polys"])
                      if sorted(map(tuple, mine["verts"][mine["vrefs"][ff:ff + cc]].tolist())) == key][0]
                np.testing.assert_allclose(mp, pl, atol=3e-7)
                assert j >= 0
            np.testing.assert_allclose(mine["internal_extents"], ref["internal_extents"], atol=2e-5)
    assert exact >= 10
    assert same_polys >= 0.85 * n


def _gen(lib, args):
    t0, g0, p0, t1, g1, p1 = args
    return L.contacts(lib, t0, g0, p0, t1, g1, p1, cd=CD, tol=TOL)


def _same(a, b, dedupe=False):
    if dedupe:
        a = np.unique(np.round(a, 5), axis=0)
        b = np.unique(np.round(b, 5), axis=0)
    if len(a) != len(b):
        return False
    if len(a) == 0:
        return True
    if np.abs(a[0, 4:7] - b[0, 4:7]).max() > 1e-3:
        return False
    used = set()
    for r in a:
        d = np.abs(b[:, :4] - r[:4]).max(1)
        for j in used:
            d[j] = 9.0
        j = int(np.argmin(d))
        if d[j] > 1e-4:
            return False
        used.add(j)
    return True


@pytest.mark.parametrize("kind,min_share", [("convex-convex", 0.97), ("box-convex", 0.97), ("sphere-convex", 0.93),
                                            ("plane-convex", 0.97)])
def test_convex_contacts_match_reference(kind, min_share):
    rng = np.random.default_rng({"convex-convex": 7, "box-convex": 8, "sphere-convex": 9, "plane-convex": 10}[kind])
    L.pxo()
    ok = bad = 0
    for t in range(260):
        cloud_a, cloud_b = _cloud(rng), _cloud(rng)
        ra, ha = L.ref_cook(cloud_a)
        rb, hb = L.ref_cook(cloud_b)
        oa, ob = L.pxo_add_hull(ha), L.pxo_add_hull(hb)
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        pa = np.concatenate([u * rng.uniform(0.1, 0.9), _quat(rng)])
        pb = np.concatenate([[0.0, 0.0, 0.0], _quat(rng)])
        if kind == "convex-convex":
            ref_args, ora_args = (5, [ra], pa, 5, [rb], pb), (5, [oa], pa, 5, [ob], pb)
        elif kind == "box-convex":
            h = rng.uniform(0.1, 0.4, 3)
            ref_args, ora_args = (3, h, pa, 5, [rb], pb), (3, h, pa, 5, [ob], pb)
        elif kind == "sphere-convex":
            r = [rng.uniform(0.1, 0.4)]
            ref_args, ora_args = (0, r, pa, 5, [rb], pb), (0, r, pa, 5, [ob], pb)
        else:
            plane = np.array([0.0, -0.3, 0.0, 0.0, 0.0, 0.7071068, 0.7071068])   # normal +y through y=-0.3
            pb = np.concatenate([[0.0, rng.uniform(-0.1, 0.3), 0.0], _quat(rng)])
            ref_args, ora_args = (1, [0], plane, 5, [rb], pb), (1, [0], plane, 5, [ob], pb)
        a, b = _gen("ref", ref_args), _gen("pxo", ora_args)
        if len(a) and a[:, 3].min() < -0.05:
            continue
        if len(a) == 0 and len(b) == 0:
            continue
        # plane-convex repeats a face's vertices (the reference passes over the
        # chosen face twice); which duplicate survives the reduction depends on
        # last-bit rounding, so those are compared as point sets
        if _same(a, b, dedupe=kind == "plane-convex"):
            ok += 1
        else:
            bad += 1
    assert ok >= 25
    assert ok / (ok + bad) >= min_share, (ok, bad)


def test_plane_convex_closed_form():
    """A tetrahedron resting on its base on the plane y = 0: three contacts at
    the base corners, all at the same separation, normal along the plane's."""
    base = np.array([[0.5, 0.0, 0.0], [-0.25, 0.0, 0.433], [-0.25, 0.0, -0.433], [0.0, 0.6, 0.0]], np.float32)
    _, h = L.ref_cook(base)
    oid = L.pxo_add_hull(h)
    plane = np.array([0.0, 0.0, 0.0, 0.0, 0.0, 0.7071068, 0.7071068])
    pose = np.array([1.0, 0.01, -2.0, 0.0, 0.0, 0.0, 1.0])
    c = L.contacts("pxo", 1, [0], plane, 5, [oid], pose, cd=CD, tol=TOL)
    assert len(c) == 3
    np.testing.assert_allclose(c[:, 3], 0.01, atol=1e-6)
    np.testing.assert_allclose(c[:, 4:7], np.tile([0.0, -1.0, 0.0], (3, 1)), atol=1e-6)
    got = np.array(sorted(map(tuple, c[:, :3].astype(np.float64).tolist())))
    want = np.array(sorted(map(tuple, (base[:3].astype(np.float64) + [1.0, 0.0, -2.0]).tolist())))
    np.testing.assert_allclose(got, want, atol=1e-6)


def test_convex_scene_generator_principal_frames():
    """Hulls in their principal frames: centre of mass at the origin and the
    inertia tensor diagonal, equal to the body's stated inertia."""
    sc = scenes.convex_layers(n_bodies=32, per_side=4)
    for i in range(sc.n):
        h = P.cook_convex(sc.hull_points[24 * i:24 * (i + 1)])
        np.testing.assert_allclose(h["com"], 0.0, atol=1e-6)
        I = 1000.0 * h["inertia"]
        np.testing.assert_allclose(np.diag(I), sc.inertia[i], rtol=1e-4)
        assert np.abs(I - np.diag(np.diag(I))).max() <= 1e-4 * np.abs(I).max()
        assert abs(1000.0 * h["volume"] - sc.mass[i]) <= 1e-4 * sc.mass[i]
