"""GPU check of the allocation-free broadphase update (pxh_bp_update_noalloc)
against the allocating one (pxh_bp_update on a broadphase object): two
updates in a row, the second with fewer boxes, give the same created and
deleted pair sets bit for bit, and both agree with the C oracle's ABP pair
set. Sizes past the maxima are rejected on the host with a clear error."""
import numpy as np
import pytest

import physx_amd as P
import oracle_libs as L

pytestmark = pytest.mark.gpu


def _boxes(rng, n, extent=30.0):
    c = rng.uniform(0, extent, (n, 3)).astype(np.float32)
    h = rng.uniform(0.1, 0.4, (n, 3)).astype(np.float32)
    return np.concatenate([c - h, c + h], 1).astype(np.float32)


def test_noalloc_update_matches_allocating_update(cuda):
    import torch
    rng = np.random.default_rng(31)
    n1, n2 = 12000, 5000
    b1 = _boxes(rng, n1)
    d1 = rng.uniform(0.0, 0.05, n1).astype(np.float32)
    g1 = np.arange(n1, dtype=np.uint32)
    g1[rng.random(n1) < 0.1] = 7            # some boxes share a group: never paired
    # second update: the first n2 boxes, a third of them moved
    b2 = b1[:n2].copy()
    mv = rng.random(n2) < 0.33
    b2[mv] += rng.uniform(-0.3, 0.3, (int(mv.sum()), 1)).astype(np.float32)
    d2, g2 = d1[:n2].copy(), g1[:n2].copy()

    dev = lambda a: torch.from_numpy(a).to(cuda)  # noqa: E731
    # a fresh diff base on this GPU: an update of zero boxes
    P.bp_update_noalloc(dev(np.zeros((0, 6), np.float32)))
    bp = P.BroadPhase()
    seen = []
    for b, d, g in ((b1, d1, g1), (b2, d2, g2)):
        print(f"noalloc update: {b.shape[0]} boxes", flush=True)
        B, D, G = dev(b), dev(d), dev(g.view(np.int32))
        bp.update(B, D, G)
        c_ref, d_ref = bp.fetch_results()
        c_na, d_na = P.bp_update_noalloc(B, D, G)
        assert np.array_equal(c_na, c_ref)
        assert np.array_equal(d_na, d_ref)
        seen.append((c_na, d_na))
    bp.close()
    # the first update from empty creates the oracle's whole pair set
    grp = g1.astype(np.int64)
    expect = L.bp_pairs_pxo(b1, d1)
    expect = expect[grp[expect[:, 0]] != grp[expect[:, 1]]]
    assert seen[0][0].shape[0] > 1000
    assert np.array_equal(seen[0][0], expect)
    assert seen[0][1].shape[0] == 0
    # the second: set differences against the first
    e2 = L.bp_pairs_pxo(b2, d2)
    e2 = e2[grp[e2[:, 0]] != grp[e2[:, 1]]]
    s1 = set(map(tuple, expect.tolist()))
    s2 = set(map(tuple, e2.tolist()))
    assert set(map(tuple, seen[1][0].tolist())) == s2 - s1
    assert set(map(tuple, seen[1][1].tolist())) == s1 - s2
    assert len(s1 - s2) > 0


def test_noalloc_rejects_sizes_past_the_maxima(cuda):
    import torch
    n = P.PXH_BP_NOALLOC_MAX_BOXES + 1
    b = torch.zeros((n, 6), dtype=torch.float32, device=cuda)
    with pytest.raises(P.PxhError, match="PXH_BP_NOALLOC_MAX_BOXES"):
        P.bp_update_noalloc(b)
    # outputs too small: the update is dropped, the sizes needed reported
    rng = np.random.default_rng(5)
    bx = _boxes(rng, 4000, extent=10.0)
    B = torch.from_numpy(bx).to(cuda)
    P.bp_update_noalloc(torch.zeros((0, 6), dtype=torch.float32, device=cuda))
    with pytest.raises(P.PxhError, match="exceed the output capacities"):
        P.bp_update_noalloc(B, max_created=1)
    created, deleted = P.bp_update_noalloc(B)
    assert np.array_equal(created, L.bp_pairs_pxo(bx, np.zeros(4000, np.float32)))
    assert deleted.shape[0] == 0
