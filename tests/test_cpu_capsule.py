"""Capsules on the CPU: the oracle's capsule contact generation against the
reference's PxGenerateContacts (built from source in oracle/_ref), closed
forms that pin it without the reference, and the capsule mass properties
against numerical integration.

The closed-form pairs (sphere-, capsule-capsule, plane-capsule) must agree
pair for pair. Capsule-box and capsule-convex run the reference's GJK/EPA on
the capsule's core segment, which stop at relative tolerances while the oracle
takes the exact closest points / minimum translation; where the segment's end
sits over a polygon edge the endpoint contact and the edge contact coincide
and rounding decides whether both are kept, so those pairs are compared as
point sets and the share of matching pairs is bounded."""
import math

import numpy as np
import pytest

import oracle_libs as L
import physx_amd as P
from physx_amd import scenes as S

CD, TOL = 0.04, 1.0
needs_ref = pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")


def _quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def _merge(c, tol=1e-4):
    """Contacts with points and separations within tol of an earlier one dropped."""
    keep = []
    for r in c:
        if all(np.abs(r[:4] - k[:4]).max() > tol for k in keep):
            keep.append(r)
    return np.array(keep).reshape(-1, c.shape[1]) if len(c) else c


def _same(a, b, dedupe=False):
    if dedupe:
        a, b = _merge(a), _merge(b)
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


def _cloud(rng, n=24):
    ax = rng.uniform(0.2, 0.6, 3)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return (d * ax).astype(np.float32)


def _cases(kind, rng, n):
    for _ in range(n):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        cap0 = [rng.uniform(0.05, 0.3), rng.uniform(0.0, 0.5)]
        cap1 = [rng.uniform(0.05, 0.3), rng.uniform(0.0, 0.5)]
        p0 = np.concatenate([u * rng.uniform(0.0, 0.9), _quat(rng)])
        p1 = np.concatenate([[0.0, 0.0, 0.0], _quat(rng)])
        if kind == "sphere-capsule":
            yield (0, [cap0[0]], p0, 2, cap1, p1), None
        elif kind == "capsule-capsule":
            yield (2, cap0, p0, 2, cap1, p1), None
        elif kind == "capsule-capsule-parallel":
            q = p1[3:] * (1.0 if rng.uniform() < 0.5 else -1.0)
            yield (2, cap0, np.concatenate([p0[:3], q]), 2, cap1, p1), None
        elif kind == "plane-capsule":
            plane = [0.0, -0.3, 0.0, 0.0, 0.0, 0.7071068, 0.7071068]   # normal +y through y=-0.3
            yield (1, [0.0], plane, 2, cap1, np.concatenate([[0.0, rng.uniform(-0.3, 0.6), 0.0], _quat(rng)])), None
        elif kind == "capsule-box":
            yield (2, cap0, p0, 3, rng.uniform(0.1, 0.4, 3), p1), None
        else:
            cloud = _cloud(rng)
            rid, h = L.ref_cook(cloud)
            yield (2, cap0, p0, 5, [rid], p1), (2, cap0, p0, 5, [L.pxo_add_hull(h)], p1)


@needs_ref
@pytest.mark.parametrize("kind", ["sphere-capsule", "capsule-capsule", "capsule-capsule-parallel", "plane-capsule"])
def test_closed_form_capsule_contacts_match_reference(kind):
    rng = np.random.default_rng({"sphere-capsule": 1, "capsule-capsule": 2, "capsule-capsule-parallel": 3,
                                 "plane-capsule": 4}[kind])
    L.pxo()
    n = 0
    for ref_args, _ in _cases(kind, rng, 300):
        a = L.contacts("ref", *ref_args, cd=CD, tol=TOL)
        b = L.contacts("pxo", *ref_args, cd=CD, tol=TOL)
        assert _same(a, b), (ref_args, a, b)
        n += len(a) > 0
    assert n >= 60


@needs_ref
@pytest.mark.parametrize("kind,seed", [("capsule-box", 5), ("capsule-convex", 6)])
def test_capsule_polytope_contacts_match_reference(kind, seed):
    rng = np.random.default_rng(seed)
    L.pxo()
    ok = bad = 0
    for ref_args, ora_args in _cases(kind, rng, 600):
        a = L.contacts("ref", *ref_args, cd=CD, tol=TOL)
        b = L.contacts("pxo", *(ora_args or ref_args), cd=CD, tol=TOL)
        if len(a) and a[:, 3].min() < -0.05:
            continue
        if len(a) == 0 and len(b) == 0:
            continue
        if _same(a, b, dedupe=True):
            ok += 1
        else:
            bad += 1
    assert ok >= 50
    assert ok / (ok + bad) >= 0.97, (ok, bad)


def _pose(p, q=(0.0, 0.0, 0.0, 1.0)):
    return np.array(list(p) + list(q), np.float64)


def _qaxis(axis, ang):
    a = np.asarray(axis, np.float64) / np.linalg.norm(axis)
    return tuple(np.concatenate([a * math.sin(ang / 2), [math.cos(ang / 2)]]))


def test_capsule_resting_on_plane():
    """A horizontal capsule 1 mm above the plane y = 0: contacts under both
    segment ends, on the capsule's surface, separation 1 mm, normal from the
    capsule down the plane's normal (the plane is shape 0)."""
    L.pxo()
    plane = _pose((0.0, 0.0, 0.0), (0.0, 0.0, 0.7071068, 0.7071068))
    r, hh = 0.2, 0.5
    c = L.contacts("pxo", 1, [0.0], plane, 2, [r, hh], _pose((0.3, r + 0.001, -1.0)), cd=CD, tol=TOL)
    assert len(c) == 2
    np.testing.assert_allclose(c[:, 3], 0.001, atol=1e-6)
    np.testing.assert_allclose(c[:, 4:7], np.tile([0.0, -1.0, 0.0], (2, 1)), atol=1e-6)
    got = np.array(sorted(map(tuple, c[:, :3].astype(np.float64).tolist())))
    np.testing.assert_allclose(got, [(-0.2, 0.001, -1.0), (0.8, 0.001, -1.0)], atol=1e-6)


def test_crossed_capsules():
    """Perpendicular capsules one above the other: one contact on the lower
    capsule's top, separation = gap, normal along the common perpendicular."""
    L.pxo()
    r0, r1, gap = 0.15, 0.1, 0.01
    top = _pose((0.0, r0 + r1 + gap, 0.0), _qaxis((0, 1, 0), math.pi / 2))   # axis along z
    bottom = _pose((0.0, 0.0, 0.0))                                          # axis along x
    c = L.contacts("pxo", 2, [r1, 0.4], top, 2, [r0, 0.6], bottom, cd=CD, tol=TOL)
    assert len(c) == 1
    np.testing.assert_allclose(c[0, 3], gap, atol=1e-6)
    np.testing.assert_allclose(c[0, 4:7], [0.0, 1.0, 0.0], atol=1e-6)
    np.testing.assert_allclose(c[0, :3], [0.0, r0 + gap, 0.0], atol=1e-6)


def test_parallel_capsules_overlap_ends():
    """Parallel capsules side by side, offset along their axis: contacts at the
    two ends of the overlap (the endpoints that project inside the other)."""
    L.pxo()
    r, hh, gap = 0.1, 0.5, 0.005
    a = _pose((0.3, 2 * r + gap, 0.0))
    b = _pose((0.0, 0.0, 0.0))
    c = L.contacts("pxo", 2, [r, hh], a, 2, [r, hh], b, cd=CD, tol=TOL)
    assert len(c) == 2
    np.testing.assert_allclose(c[:, 3], gap, atol=1e-6)
    np.testing.assert_allclose(c[:, 4:7], np.tile([0.0, 1.0, 0.0], (2, 1)), atol=1e-6)
    xs = sorted(c[:, 0].tolist())
    np.testing.assert_allclose(xs, [-0.2, 0.5], atol=1e-6)


@pytest.mark.parametrize("shape", ["box", "convex"])
def test_capsule_lying_on_flat_face(shape):
    """A capsule lying across the top face of a box, or of a tetrahedron's
    upward base: two contacts under the segment ends, separation = gap,
    normal = the face normal."""
    L.pxo()
    r, hh, gap = 0.1, 0.2, 0.002
    cap = _pose((0.05, 0.3 + r + gap, 0.02))
    if shape == "box":
        args = (3, [0.5, 0.3, 0.5], _pose((0.0, 0.0, 0.0)))
    else:
        # tetrahedron with base triangle in y = 0.3, apex below
        pts = np.array([[1.0, 0.3, 0.0], [-0.5, 0.3, 0.866], [-0.5, 0.3, -0.866], [0.0, -0.7, 0.0]], np.float32)
        args = (5, [L.pxo_add_hull(P.cook_convex(pts))], _pose((0.0, 0.0, 0.0)))
    c = L.contacts("pxo", 2, [r, hh], cap, *args, cd=CD, tol=TOL)
    assert len(c) == 2
    np.testing.assert_allclose(c[:, 3], gap, atol=1e-6)
    np.testing.assert_allclose(c[:, 4:7], np.tile([0.0, 1.0, 0.0], (2, 1)), atol=1e-6)
    np.testing.assert_allclose(sorted(c[:, 0].tolist()), [0.05 - hh, 0.05 + hh], atol=1e-6)
    np.testing.assert_allclose(c[:, 1], 0.3 + gap, atol=1e-6)


def test_capsule_mass_properties_match_integration():
    """Volume and the principal moments of a capsule against a midpoint-rule
    integration over a fine grid of its bounding box."""
    r, hh = 0.2, 0.35
    m, inn = S.mass_props(np.array([S.CAPSULE]), np.array([[r, hh, 0.0, 0.0]]), 1000.0)
    k = 240
    xs = (np.arange(k) + 0.5) / k
    x = -hh - r + xs * 2 * (hh + r)
    y = -r + xs * 2 * r
    X, Y, Z = np.meshgrid(x, y, y, indexing="ij")
    cx = np.clip(X, -hh, hh)
    inside = (X - cx) ** 2 + Y ** 2 + Z ** 2 <= r * r
    dv = (2 * (hh + r) / k) * (2 * r / k) ** 2
    vol = inside.sum() * dv
    ixx = 1000.0 * ((Y ** 2 + Z ** 2) * inside).sum() * dv
    iyy = 1000.0 * ((X ** 2 + Z ** 2) * inside).sum() * dv
    assert abs(m[0] - 1000.0 * vol) <= 2e-3 * m[0]
    assert abs(inn[0, 0] - ixx) <= 3e-3 * ixx
    assert abs(inn[0, 1] - iyy) <= 3e-3 * iyy
    assert inn[0, 1] == inn[0, 2]
