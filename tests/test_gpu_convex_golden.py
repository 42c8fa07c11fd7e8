"""The GPU step's convex-class contacts on the reference's own cooked hulls
and real pile states: the 2,000 touching pairs per class of
tests/golden/convex_pairs.npz (taken from piles formed in the reference SDK,
hulls registered with pxh_scene_add_cooked_convex exactly as the reference
cooked them), each pair its own world, one step on the GPU: its contacts
(copyContactData) against the oracle stepping the same worlds, and against
the reference's PxGenerateContacts output, within the bounds
tests/test_cpu_convex_golden.py holds the oracle to."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L
from physx_amd import scenes as S
from test_cpu_convex_golden import BOUNDS

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


@pytest.fixture(scope="module")
def golden():
    g = L.ConvexGolden()
    g.register_with_oracle()
    return g


def _worlds(golden, cls):
    """one world per pair: shape1's body then shape0's (the later-added body
    of two of one type comes first in a pair), or the body and a plane y=0"""
    T, G, Pp, N, C = golden.pairs(cls)
    plane = T[0, 0] == S.PLANE
    types, dims, pose, wid = [], [], [], []
    for i in range(len(N)):
        if plane:
            types.append(T[i, 1]); dims.append(G[i, 4:]); pose.append(Pp[i, 7:]); wid.append(i)
        else:
            types += [T[i, 1], T[i, 0]]
            dims += [G[i, 4:], G[i, :4]]
            pose += [Pp[i, 7:], Pp[i, :7]]
            wid += [i, i]
    n = len(types)
    planes = np.tile(np.array([[0, 1, 0, 0]], np.float32), (len(N), 1)) if plane else np.zeros((0, 4), np.float32)
    pscene = np.arange(len(N), dtype=np.int32) if plane else np.zeros(0, np.int32)
    return S.Scene(np.array(types, np.int32), np.array(dims, np.float32), np.array(pose, np.float32),
                   np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.ones(n, np.float32),
                   np.ones((n, 3), np.float32), np.array(wid, np.int32), planes, pscene, len(N), cls, {})


def _by_pair(rec, per_world):
    """contact sets of the records by pair (world): (count, (k,8) contacts
    as point xyz, separation, normal xyz, 0) with the normal from shape1
    towards shape0"""
    out = {}
    ids = rec[:, :3].copy().view(np.int32)
    for r, (a, b, k) in zip(rec, ids):
        w = int(a) // per_world
        c = np.zeros((k, 8), np.float32)
        c[:, :4] = r[8:8 + 4 * k].reshape(k, 4)
        c[:, 4:7] = r[4:7] if b >= 0 else -r[4:7]   # plane pairs: the body is shape1
        out[w] = c
    return out


@pytest.mark.parametrize("cls", L.CONVEX_CLASSES)
def test_gpu_convex_contacts_on_reference_hulls(cuda, golden, cls):
    sc = _worlds(golden, cls)
    g = P.Scene(position_iterations=4, velocity_iterations=1)
    for h in range(golden.n_hulls):
        d = golden.hull(h)
        assert g.add_cooked_convex(d["verts"], d["planes"], d["polys"], d["vrefs"], d["com"]) == h
    g.add(sc)
    g.step(DT)
    o = L.OracleWorld(sc, pos_iters=4, vel_iters=1)
    o.step(DT)
    per_world = 1 if cls == "plane-convex" else 2
    gpu = _by_pair(g.contact_data(), per_world)
    ora = _by_pair(o.contact_data(), per_world)
    T, G, Pp, N, C = golden.pairs(cls)
    off_oracle = [i for i in range(len(N))
                  if L.same_contacts(L.dedupe(gpu.get(i, np.zeros((0, 8)))), L.dedupe(ora.get(i, np.zeros((0, 8)))))]
    print(cls, "pairs off the oracle:", len(off_oracle), off_oracle[:8], flush=True)
    assert len(off_oracle) <= len(N) // 500, off_oracle[:10]
    share, max_ang, max_sep = BOUNDS[cls]
    bad = []
    for i in range(len(N)):
        a = gpu.get(i, np.zeros((0, 8), np.float32))
        b = C[i, :N[i]]
        if L.same_contacts(L.dedupe(a), L.dedupe(b)) is None:
            continue
        assert len(a) > 0, (cls, i, "reference touches, GPU does not")
        ang = float(np.arccos(np.clip(np.dot(a[0, 4:7], b[0, 4:7]), -1.0, 1.0)))
        bad.append((i, ang, float(abs(a[:, 3].min() - b[:, 3].min()))))
    print(cls, "pairs off the reference:", len(bad), flush=True)
    assert 1.0 - len(bad) / len(N) >= share, (cls, len(bad))
    for i, ang, sep in bad:
        assert ang <= max_ang and sep <= max_sep, (cls, i, ang, sep)
