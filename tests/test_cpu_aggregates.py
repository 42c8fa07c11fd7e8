"""Aggregates (PxAggregate) in the oracle against the reference SDK: the
members of an aggregate without self-collisions never pair with each other,
with self-collisions they pair as free bodies do, and pairs with every other
body are unchanged. Overlapping groups of boxes, half of them aggregates
without self-collisions: the first step's touching pairs equal the pairs the
reference scene reports."""
import numpy as np
import pytest

import oracle_libs as L
from physx_amd import scenes as S

DT = 1.0 / 60.0


def _norm(pairs):
    return {(a, -1) if b < 0 else (min(a, b), max(a, b)) for a, b in pairs}


def test_aggregate_filter_semantics():
    sc = S.aggregate_pile(n_groups=64, per_group=4, seed=7)
    o = L.OracleWorld(sc)
    o.step(DT)
    pairs = _norm(map(tuple, o.constraints()[:, :2].tolist()))
    group = {int(i): g for g, (m, _) in enumerate(sc.aggregates) for i in m}
    self_coll = {g: sc_ for g, (_, sc_) in enumerate(sc.aggregates)}
    inside = [(a, b) for a, b in pairs if b >= 0 and group[a] == group[b]]
    assert inside, "self-colliding aggregates keep their internal pairs"
    assert all(self_coll[group[a]] for a, b in inside)
    # without the aggregates the same scene also pairs the members of the
    # other groups
    free = S.aggregate_pile(n_groups=64, per_group=4, seed=7)
    free.aggregates = None
    f = L.OracleWorld(free)
    f.step(DT)
    fpairs = _norm(map(tuple, f.constraints()[:, :2].tolist()))
    extra = fpairs - pairs
    assert extra and all(b >= 0 and group[a] == group[b] and not self_coll[group[a]] for a, b in extra)
    assert f.stats()["bp_pairs"] - o.stats()["bp_pairs"] >= len(extra)


@pytest.mark.skipif(not L.have_ref(), reason="reference CPU SDK not built (oracle/_ref)")
def test_aggregate_pairs_vs_live_reference():
    sc = S.aggregate_pile(n_groups=64, per_group=4, seed=7)
    o = L.OracleWorld(sc)
    r = L.RefScene(sc, "tgs", 4, 1, threads=1, report_contacts=True)
    o.step(DT)
    r.step(DT)
    ours = _norm(map(tuple, o.constraints()[:, :2].tolist()))
    ref = _norm((int(a) - 1, int(b) - 1 if b > 0 else -1) for a, b in r.contacts()[:, :2])
    assert ours == ref, (sorted(ours - ref)[:5], sorted(ref - ours)[:5])
