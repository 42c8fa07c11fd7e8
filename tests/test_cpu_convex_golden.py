"""The oracle's convex-class contact generation against the reference's
PxGenerateContacts on real pile states: 2,000 touching pairs per class taken
from piles formed in the reference SDK (tests/golden/convex_pairs.npz,
reference-cooked hulls). Pairs compare as contact sets (coincident points
once: where a capsule end lies over a polygon edge, its end contact and the
edge contact coincide, and rounding decides whether both are kept).

Where the two disagree the cause is named and the residual bounded: the
reference's GJK and EPA stop at relative tolerances (GJK when the support
distance comes within 0.0225% of the current distance,
geomutils/src/gjk/GuGJKPenetration.h:113; EPA at its own tolerance,
GuEPA.cpp), so near-degenerate closest features (a sphere or capsule core
over a hull edge or vertex, two nearly parallel hull faces) come out with a
slightly different normal, or a neighbouring witness face; this oracle takes
the exact closest features / minimum translation. The tests bound the share
of such pairs per class, and the normal angle and deepest-separation
difference of every one of them."""
import numpy as np
import pytest

import oracle_libs as L

# class: (least share agreeing, largest normal angle of a residual (rad),
#         largest deepest-separation difference of a residual (m))
BOUNDS = {"convex-convex": (0.995, 0.5, 0.006), "box-convex": (0.998, 0.01, 0.001),
          "sphere-convex": (0.975, 0.06, 0.001), "plane-convex": (1.0, 0.0, 0.0),
          "capsule-convex": (0.975, 0.08, 0.04), "capsule-box": (0.995, 0.01, 0.03)}


@pytest.fixture(scope="module")
def golden():
    g = L.ConvexGolden()
    g.register_with_oracle()
    return g


@pytest.mark.parametrize("cls", L.CONVEX_CLASSES)
def test_convex_pairs_vs_reference_golden(golden, cls):
    T, G, Pp, N, C = golden.pairs(cls)
    assert len(N) == 2000
    share, max_ang, max_sep = BOUNDS[cls]
    bad = []
    for i in range(len(N)):
        a = L.contacts("pxo", int(T[i, 0]), G[i, :4], Pp[i, :7], int(T[i, 1]), G[i, 4:], Pp[i, 7:])
        b = C[i, :N[i]]
        if L.same_contacts(L.dedupe(a), L.dedupe(b)) is None:
            continue
        assert len(a) > 0, (cls, i, "reference touches, oracle does not")
        ang = float(np.arccos(np.clip(np.dot(a[0, 4:7], b[0, 4:7]), -1.0, 1.0)))
        sep = float(abs(a[:, 3].min() - b[:, 3].min()))
        bad.append((i, ang, sep))
    print(cls, "differing pairs:", len(bad), [(i, round(a, 4), round(s, 5)) for i, a, s in bad[:8]])
    assert 1.0 - len(bad) / len(N) >= share, (cls, len(bad))
    for i, ang, sep in bad:
        assert ang <= max_ang and sep <= max_sep, (cls, i, ang, sep)
