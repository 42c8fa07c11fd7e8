// Convex contact generation on the device, one pair per lane: convex mesh
// (and the box seen as a polytope) against convex mesh, sphere and plane.
// Each generator follows the reference's PCM contact generation on a fresh
// manifold, so a pair gives the contacts of PxGenerateContacts with an empty
// PxCache: the separation is the exact closest-feature distance or minimum
// translation (what the reference's GJK / EPA converge to on polytopes), the
// witness polygons, polygon clipping and reduction to four points follow
// geomutils/src/pcm/GuPCMContactGenBoxConvex.cpp and GuPCMContactGenUtil.cpp.
#pragma once
#include "pxh_kernels.h"
#include "pxh_np.h"

namespace pxh {

constexpr int GEOM_CONVEXMESH = 5;
constexpr int HULL_MAX_POLY_VERTS = 16;   // polygon size the device path accepts
constexpr int HULL_MAX_CONTACTS = 32;     // clipped points before reduction

// PCMPolygonalBox (geomutils/src/pcm/GuPCMShapeConvex.cpp:40-136): the quads
// of gPCMBoxPolygonData, their planes, and the box's edges in polygon order
__constant__ const uint8_t kBoxQuads[24] = {0, 3, 2, 1, 1, 2, 6, 5, 5, 6, 7, 4, 4, 7, 3, 0, 3, 7, 6, 2, 4, 0, 1, 5};
__constant__ const uint32_t kBoxEdges[12] = {0x03000300u, 0x04000203u, 0x01000102u, 0x05000001u,
                                             0x04010602u, 0x02010506u, 0x05010105u, 0x04020706u,
                                             0x03020407u, 0x05020504u, 0x04030307u, 0x05030400u};

// One polytope: a hull (its arrays, in global memory or staged in LDS), or a
// box by its half extents.
struct Poly
{
	const float4* v;         // vertices
	const float4* pl;        // planes
	const uint32_t* poly;    // first vref | count << 16
	const uint8_t* vr;       // vertex loops
	const uint32_t* ed;      // edges
	int nv, np, ne;
	bool box;
	f3 h;
	float margin;

	PXH_D f3 vert(int i) const
	{
		if (box)
			return mk((i == 1 || i == 2 || i == 5 || i == 6) ? h.x : -h.x,
			          (i == 2 || i == 3 || i == 6 || i == 7) ? h.y : -h.y, i >= 4 ? h.z : -h.z);
		return mk(v[i]);
	}
	PXH_D f3 pn(int f) const
	{
		if (box)
		{
			const float s = (f == 1 || f == 2 || f == 4) ? 1.0f : -1.0f;
			return (f == 1 || f == 3) ? mk(s, 0.0f, 0.0f) : ((f == 4 || f == 5) ? mk(0.0f, s, 0.0f) : mk(0.0f, 0.0f, s));
		}
		return mk(pl[f]);
	}
	PXH_D float pd(int f) const
	{
		if (box)
			return (f == 1 || f == 3) ? -h.x : ((f == 4 || f == 5) ? -h.y : -h.z);
		return pl[f].w;
	}
	PXH_D int pcount(int f) const { return box ? 4 : int(poly[f] >> 16); }
	PXH_D int pvert(int f, int k) const
	{
		return box ? int(kBoxQuads[4 * f + k]) : int(vr[int(poly[f] & 0xFFFFu) + k]);
	}
	PXH_D uint32_t edge(int e) const { return box ? kBoxEdges[e] : ed[e]; }
};

PXH_D Poly hullPoly(const HullSet& hs, int id)
{
	Poly p;
	const int4 a = hs.info[id], b = hs.info2[id];
	p.v = hs.v + a.x;
	p.pl = hs.pl + a.z;
	p.poly = hs.poly + a.z;
	p.vr = hs.vr + b.x;
	p.ed = hs.edge + b.y;
	p.nv = a.y; p.np = a.w; p.ne = b.z;
	p.box = false;
	p.h = mk(0.0f, 0.0f, 0.0f);
	p.margin = hs.bc[id].w;
	return p;
}

// BoxV margin: 15% of the smallest half extent (geomutils/src/gjk/GuVecBox.h:46, :56-75)
PXH_D Poly boxPoly(f3 h)
{
	Poly p;
	p.v = p.pl = nullptr;
	p.poly = p.ed = nullptr;
	p.vr = nullptr;
	p.nv = 8; p.np = 6; p.ne = 12;
	p.box = true;
	p.h = h;
	p.margin = fminf(h.x, fminf(h.y, h.z)) * 0.15f;
	return p;
}

PXH_D f3 mxtransinv(const mxf& t, f3 v) { return mtmul(t.r, v - t.p); }
PXH_D mxf mxfOf(const xf& t) { mxf m; m.r = qmat(t.q); m.p = t.p; return m; }

// a point of polygon f's plane inside its counter-clockwise loop
PXH_D bool inPolygon(const Poly& P, int f, f3 q)
{
	const f3 n = P.pn(f);
	const int c = P.pcount(f);
	for (int k = 0; k < c; k++)
	{
		const f3 a = P.vert(P.pvert(f, k)), b = P.vert(P.pvert(f, (k + 1) % c));
		if (dot(q - a, cross(b - a, n)) > 0.0f)
			return false;
	}
	return true;
}

// closest points of segments p1q1, p2q2 (Ericson, Real-Time Collision
// Detection, 5.1.9); returns the squared distance
PXH_D float segClosestSq(f3 p1, f3 q1, f3 p2, f3 q2, f3& c1, f3& c2)
{
	const f3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
	const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
	float s, t;
	if (a <= 1e-20f && e <= 1e-20f) { s = t = 0.0f; }
	else if (a <= 1e-20f) { s = 0.0f; t = fminf(fmaxf(f / e, 0.0f), 1.0f); }
	else
	{
		const float c = dot(d1, r);
		if (e <= 1e-20f) { t = 0.0f; s = fminf(fmaxf(-c / a, 0.0f), 1.0f); }
		else
		{
			const float b = dot(d1, d2), den = a * e - b * b;
			s = den != 0.0f ? fminf(fmaxf((b * f - c * e) / den, 0.0f), 1.0f) : 0.0f;
			t = (b * s + f) / e;
			if (t < 0.0f) { t = 0.0f; s = fminf(fmaxf(-c / a, 0.0f), 1.0f); }
			else if (t > 1.0f) { t = 1.0f; s = fminf(fmaxf((b - c) / a, 0.0f), 1.0f); }
		}
	}
	c1 = p1 + d1 * s;
	c2 = p2 + d2 * t;
	const f3 w = c1 - c2;
	return dot(w, w);
}

// Separation of A (placed in B's frame by aToB) and B. A face axis that
// separates them by more than cd ends the search (no contact possible; dist
// is that separation). Separated: exact
// closest points, dist > 0, n = (cA - cB) / dist. Overlapping: the minimum
// translation over both polytopes' face normals and the edge pairs that form
// a face of the Minkowski difference (their arcs cross on the Gauss maps;
// Gregorius, "The Separating Axis Test between Convex Polyhedra", GDC 2013);
// dist = -depth, n from B towards A.
PXH_D void polySeparation(const Poly& A, const mxf& aToB, const Poly& B, float cd, float& dist, f3& n, f3& cA,
                           f3& cB)
{
	float best = FLT_BIG;
	bool sep = false;
	for (int f = 0; f < B.np && !sep; f++)
	{
		const f3 bn = B.pn(f);
		const float bd = B.pd(f);
		float mn = FLT_BIG;
		f3 vmin = mk(0.0f, 0.0f, 0.0f);
		for (int i = 0; i < A.nv; i++)
		{
			const f3 va = mxtrans(aToB, A.vert(i));
			const float h = dot(bn, va) + bd;
			if (h < mn) { mn = h; vmin = va; }
		}
		if (mn > cd)
		{
			dist = mn;   // farther apart than the contact distance along this axis
			return;
		}
		if (mn > 0.0f) sep = true;
		else if (-mn < best)
		{
			best = -mn;
			n = bn;
			cA = vmin;
			cB = vmin - bn * mn;
		}
	}
	for (int f = 0; f < A.np && !sep; f++)
	{
		const f3 an = mmul(aToB.r, A.pn(f));
		const float ad = A.pd(f) - dot(an, aToB.p);
		float mn = FLT_BIG;
		f3 vmin = mk(0.0f, 0.0f, 0.0f);
		for (int i = 0; i < B.nv; i++)
		{
			const f3 vb = B.vert(i);
			const float h = dot(an, vb) + ad;
			if (h < mn) { mn = h; vmin = vb; }
		}
		if (mn > cd)
		{
			dist = mn;
			return;
		}
		if (mn > 0.0f) sep = true;
		else if (-mn < best)
		{
			best = -mn;
			n = -an;
			cB = vmin;
			cA = vmin - an * mn;
		}
	}
	if (!sep)
	{
		f3 bmean = mk(0.0f, 0.0f, 0.0f);
		for (int i = 0; i < B.nv; i++)
			bmean = bmean + B.vert(i);
		bmean = bmean * (1.0f / float(B.nv));
		for (int ea = 0; ea < A.ne && !sep; ea++)
		{
			const uint32_t e = A.edge(ea);
			const f3 a0 = mxtrans(aToB, A.vert(int(e & 0xFF))), a1 = mxtrans(aToB, A.vert(int((e >> 8) & 0xFF)));
			const f3 na = mmul(aToB.r, A.pn(int((e >> 16) & 0xFF))), nb = mmul(aToB.r, A.pn(int(e >> 24)));
			const f3 bxa = cross(nb, na);
			for (int ebi = 0; ebi < B.ne && !sep; ebi++)
			{
				const uint32_t g = B.edge(ebi);
				const f3 nc = -B.pn(int((g >> 16) & 0xFF)), nd = -B.pn(int(g >> 24));
				const f3 dxc = cross(nd, nc);
				const float cba = dot(nc, bxa), dba = dot(nd, bxa), adc = dot(na, dxc), bdc = dot(nb, dxc);
				if (!(cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f))
					continue;
				const f3 b0 = B.vert(int(g & 0xFF)), b1 = B.vert(int((g >> 8) & 0xFF));
				f3 ax = cross(a1 - a0, b1 - b0);
				const float l2 = dot(ax, ax);
				if (l2 < 1e-12f)
					continue;
				ax = ax * (1.0f / sqrtf(l2));
				if (dot(ax, b0 - bmean) < 0.0f)
					ax = -ax;
				const float ov = dot(ax, b0) - dot(ax, a0);
				if (ov < 0.0f) sep = true;
				else if (ov < best)
				{
					f3 c1, c2;
					segClosestSq(a0, a1, b0, b1, c1, c2);
					best = ov;
					n = ax;
					cA = c1;
					cB = c2;
				}
			}
		}
	}
	if (!sep)
	{
		dist = -best;
		return;
	}
	float d2 = FLT_BIG;
	for (int f = 0; f < B.np; f++)
	{
		const f3 bn = B.pn(f);
		const float bd = B.pd(f);
		for (int i = 0; i < A.nv; i++)
		{
			const f3 va = mxtrans(aToB, A.vert(i));
			const float h = dot(bn, va) + bd;
			if (h < 0.0f || h * h >= d2)
				continue;
			const f3 q = va - bn * h;
			if (!inPolygon(B, f, q))
				continue;
			d2 = h * h; cA = va; cB = q;
		}
	}
	for (int f = 0; f < A.np; f++)
	{
		const f3 an = mmul(aToB.r, A.pn(f));
		const float ad = A.pd(f) - dot(an, aToB.p);
		for (int i = 0; i < B.nv; i++)
		{
			const f3 vb = B.vert(i);
			const float h = dot(an, vb) + ad;
			if (h < 0.0f || h * h >= d2)
				continue;
			const f3 q = vb - an * h;
			if (!inPolygon(A, f, mxtransinv(aToB, q)))
				continue;
			d2 = h * h; cB = vb; cA = q;
		}
	}
	for (int ea = 0; ea < A.ne; ea++)
	{
		const uint32_t e = A.edge(ea);
		const f3 a0 = mxtrans(aToB, A.vert(int(e & 0xFF))), a1 = mxtrans(aToB, A.vert(int((e >> 8) & 0xFF)));
		for (int ebi = 0; ebi < B.ne; ebi++)
		{
			const uint32_t g = B.edge(ebi);
			f3 c1, c2;
			const float s = segClosestSq(a0, a1, B.vert(int(g & 0xFF)), B.vert(int((g >> 8) & 0xFF)), c1, c2);
			if (s < d2) { d2 = s; cA = c1; cB = c2; }
		}
	}
	dist = sqrtf(d2);
	n = (cA - cB) * (1.0f / dist);
}

// Gu::getWitnessPolygonIndex (geomutils/src/pcm/GuPCMContactGenUtil.cpp:201-271);
// the per-polygon distances are evaluated again in the second pass
PXH_D int witnessPolygon(const Poly& P, f3 normal, f3 closest, float tolerance)
{
	const float eps = -tolerance;
	float dist = dot(P.pn(0), closest) + P.pd(0);
	float minDist = dist >= eps ? fabsf(dist) : FLT_BIG;
	float maxDist = dist;
	int maxIdx = 0, idx = 0;
	for (int i = 1; i < P.np; i++)
	{
		dist = dot(P.pn(i), closest) + P.pd(i);
		const float pdi = dist >= eps ? fabsf(dist) : FLT_BIG;
		if (minDist > pdi) { minDist = pdi; idx = i; }
		if (dist > maxDist) { maxDist = dist; maxIdx = i; }
	}
	if (minDist == FLT_BIG)
		return maxIdx;
	float bestProj = dot(normalize(P.pn(idx)), normal);
	const int first = idx;
	for (int i = 0; i < P.np; i++)
	{
		const float di = dot(P.pn(i), closest) + P.pd(i);
		const float pdi = di >= eps ? fabsf(di) : FLT_BIG;
		if ((tolerance > (pdi - minDist)) && (first != i))
		{
			const float proj = dot(normalize(P.pn(i)), normal);
			if (bestProj > proj) { idx = i; bestProj = proj; }
		}
	}
	return idx;
}

// Gu::findRotationMatrixFromZAxis (geomutils/src/pcm/GuPersistentContactManifold.cpp:289-354):
// a rotation whose third row is `to`
PXH_D m33 rotFromZ(f3 to)
{
	m33 m;
	const float e = to.z, f = fabsf(e);
	if (0.9999f > f)
	{
		const float vx = -to.y, vy = to.x;
		const float h = 1.0f / (1.0f + e);
		const float hvx = h * vx, hvxy = hvx * vy;
		m.c0 = mk(hvx * vx + e, hvxy, vy);
		m.c1 = mk(hvxy, h * (vy * vy) + e, -vx);
		m.c2 = mk(-vy, vx, e);
		return m;
	}
	const f3 from = mk(0.0f, 0.0f, 1.0f), absFrom = mk(0.0f, 1.0f, 0.0f);
	const f3 u = absFrom - from, v = absFrom - to;
	const float du = dot(u, u), dv = dot(v, v), duv = dot(u, v);
	const float c1 = -(2.0f / du), c2 = -(2.0f / dv), c3 = c1 * (c2 * duv);
	const f3 c1u = u * c1, c2v = v * c2, c3v = v * c3;
	m.c0 = u * c1u.x + (v * c2v.x + u * c3v.x);
	m.c0.x += 1.0f;
	m.c1 = u * c1u.y + (v * c2v.y + u * c3v.y);
	m.c1.y += 1.0f;
	m.c2 = u * c1u.z + (v * c2v.z + u * c3v.z);
	m.c2.z += 1.0f;
	return m;
}

PXH_D float triArea2(f3 a, f3 b, f3 c)
{
	const f3 ca = a - c, cb = b - c;
	return ca.x * cb.y - ca.y * cb.x;
}

// Gu::contains (geomutils/src/pcm/GuPCMContactGenUtil.cpp:35-103) for an n-gon
PXH_D bool polyContainsN(const f3* verts, int nv, f3 p, f3 mn, f3 mx)
{
	if ((mn.x > p.x || p.x > mx.x) || (mn.y > p.y || p.y > mx.y))
		return false;
	const float tx = p.x, ty = p.y, eps = 1.1920928955078125e-7f;
	int inter = 0;
	for (int i = 0, j = nv - 1; i < nv; j = i++)
	{
		const float jy = verts[j].y, iy = verts[i].y, jx = verts[j].x, ix = verts[i].x;
		if ((tx == jx && ty == jy) || (tx == ix && ty == iy))
			return true;
		if ((jy > ty) != (iy > ty))
		{
			const float jix = ix - jx, jiy = iy - jy, jty = ty - jy;
			const float part1 = jty * jix, part2 = (jx + eps) * jiy, part3 = tx * jiy;
			const bool cmp = jiy > 0.0f;
			const float tmp = part1 + part2;
			if ((cmp ? tmp : part3) >= (cmp ? part3 : tmp))
			{
				if (inter == 1)
					return false;
				inter++;
			}
		}
	}
	return inter > 0;
}

// generatedContacts (geomutils/src/pcm/GuPCMContactGenBoxConvex.cpp:331-530):
// reference polygon `ref` of P0, incident polygon `inc` of P1, t0To1 maps
// P0's frame to P1's. Points keep la on P1 (P1's frame), lb on P0's
// reference plane (P0's frame), ln the reference normal (P0's frame).
PXH_D void generatedContacts(const Poly& P0, const Poly& P1, int ref, int inc, const mxf& t0To1, PC* mc, int& nc,
                             float cd)
{
	const float feps = 1.1920928955078125e-7f;
	const f3 n = normalize(P0.pn(ref));
	const m33 rot = rotFromZ(n);
	const int nr = P0.pcount(ref), ni = P1.pcount(inc);
	f3 p0[HULL_MAX_POLY_VERTS], p1[HULL_MAX_POLY_VERTS];
	float t1[HULL_MAX_POLY_VERTS];
	bool pen1[HULL_MAX_POLY_VERTS];
	const f3 sPoint = P1.vert(P1.pvert(inc, 0));
	f3 rMin = mk(FLT_BIG, FLT_BIG, FLT_BIG), rMax = -rMin;
	for (int i = 0; i < nr; i++)
	{
		p0[i] = mmul(rot, P0.vert(P0.pvert(ref, i)));
		rMin = mk(fminf(rMin.x, p0[i].x), fminf(rMin.y, p0[i].y), fminf(rMin.z, p0[i].z));
		rMax = mk(fmaxf(rMax.x, p0[i].x), fmaxf(rMax.y, p0[i].y), fmaxf(rMax.z, p0[i].z));
	}
	rMin = rMin - mk(feps, feps, feps);
	rMax = rMax + mk(feps, feps, feps);
	const float d = p0[0].z, rd = d + cd;
	f3 iMin = mk(FLT_BIG, FLT_BIG, FLT_BIG), iMax = -iMin;
	int inside = 0;
	for (int i = 0; i < ni; i++)
	{
		const f3 vert1 = P1.vert(P1.pvert(inc, i));
		const f3 a = mxtransinv(t0To1, vert1);
		f3 q = mmul(rot, a);
		const float z = q.z;
		t1[i] = z - d;
		q.z = d;
		p1[i] = q;
		iMin = mk(fminf(iMin.x, q.x), fminf(iMin.y, q.y), fminf(iMin.z, q.z));
		iMax = mk(fmaxf(iMax.x, q.x), fmaxf(iMax.y, q.y), fmaxf(iMax.z, q.z));
		pen1[i] = rd > z;
		if (pen1[i] && polyContainsN(p0, nr, q, rMin, rMax))
		{
			inside++;
			if (nc == HULL_MAX_CONTACTS)
				return;
			PC& c = mc[nc++];
			c.la = vert1;
			c.lb = mtmul(rot, q);
			c.ln = n;
			c.pen = t1[i];
		}
	}
	if (inside == ni)
		return;
	inside = 0;
	iMin = iMin - mk(feps, feps, feps);
	iMax = iMax + mk(feps, feps, feps);
	const f3 incN = normalize(P1.pn(inc));
	const f3 nIn1 = mmul(t0To1.r, n);
	for (int i = 0; i < nr; i++)
	{
		if (!polyContainsN(p1, ni, p0[i], iMin, iMax))
			continue;
		const f3 vert0 = mtmul(rot, p0[i]);
		const f3 a = mxtrans(t0To1, vert0);
		const float t = dot(incN, sPoint - a) / dot(incN, nIn1);
		if (t > cd)
			continue;
		inside++;
		if (nc == HULL_MAX_CONTACTS)
			return;
		PC& c = mc[nc++];
		c.la = nIn1 * t + a;
		c.lb = vert0;
		c.ln = n;
		c.pen = t;
	}
	if (inside == nr)
		return;
	for (int iS = 0, iE = ni - 1; iS < ni; iE = iS++)
	{
		if (!pen1[iS] && !pen1[iE])
			continue;
		const f3 ipA = p1[iS], ipB = p1[iE];
		const f3 ipAO = mk(ipA.x, ipA.y, t1[iS] + d), ipBO = mk(ipB.x, ipB.y, t1[iE] + d);
		const float iMnx = fminf(ipA.x, ipB.x), iMny = fminf(ipA.y, ipB.y);
		const float iMxx = fmaxf(ipA.x, ipB.x), iMxy = fmaxf(ipA.y, ipB.y);
		for (int rS = 0, rE = nr - 1; rS < nr; rE = rS++)
		{
			const f3 rpA = p0[rS], rpB = p0[rE];
			const float rMnx = fminf(rpA.x, rpB.x), rMny = fminf(rpA.y, rpB.y);
			const float rMxx = fmaxf(rpA.x, rpB.x), rMxy = fmaxf(rpA.y, rpB.y);
			if (iMnx > rMxx || iMny > rMxy || rMnx > iMxx || rMny > iMxy)
				continue;
			const float a1 = triArea2(rpA, rpB, ipA), a2 = triArea2(rpA, rpB, ipB);
			if (!(0.0f > a1 * a2))
				continue;
			const float a3 = triArea2(ipA, ipB, rpA), a4 = triArea2(ipA, ipB, rpB);
			if (!(0.0f > a3 * a4))
				continue;
			const float t = a1 / (a2 - a1);
			const f3 pBB = ipAO - (ipBO - ipAO) * t;
			const f3 pAA = mk(pBB.x, pBB.y, d);
			const float pen = pBB.z - pAA.z;
			if (pen > cd)
				continue;
			if (nc == HULL_MAX_CONTACTS)
				return;
			PC& c = mc[nc++];
			c.la = mxtrans(t0To1, mtmul(rot, pBB));
			c.lb = mtmul(rot, pAA);
			c.ln = n;
			c.pen = pen;
		}
	}
}

// reduction to four (PersistentContactManifold::addBatchManifoldContacts,
// geomutils/src/pcm/GuPersistentContactManifold.cpp:805-830)
PXH_D int batchReduce(PC* mc, int nc, float tol)
{
	if (nc <= 4)
		return nc;
	int ch[4];
	reduceBatchT<HULL_MAX_CONTACTS>(mc, nc, tol, ch);
	PC out[4];
	for (int i = 0; i < 4; i++)
		out[i] = mc[ch[i]];
	for (int i = 0; i < 4; i++)
		mc[i] = out[i];
	return 4;
}

// manifold -> contacts: points on shape1 in the world, the summed normal,
// points within the contact distance (PersistentContactManifold::
// addManifoldContactsToContactBuffer / getWorldNormal,
// geomutils/src/pcm/GuPersistentContactManifold.cpp:739-762, GuPersistentContactManifold.h:405-416)
PXH_D int manifoldOut(const PC* m, int n, const xf& x1, float cd, Contacts& o)
{
	if (n == 0)
		return 0;
	f3 s = m[0].ln;
	for (int i = 1; i < n; i++)
		s = s + m[i].ln;
	o.n = normalize(qrot(x1.q, s));
	int k = 0;
	for (int i = 0; i < n && k < 4; i++)
		if (cd >= m[i].pen)
		{
			o.p[k] = xtrans(x1, m[i].lb);
			o.sep[k] = m[i].pen;
			k++;
		}
	o.count = k;
	return k;
}

// Contacts of a box- or convex-convex pair from its separation (dist, n from
// B towards A, closest points or minimum-translation witnesses cA, cB in B's
// frame): witness polygons, clipping and reduction to four.
PXH_D int convexContactsFrom(const Poly& A, const xf& x0, const Poly& B, const xf& x1, const mxf& aToB, float dist,
                             f3 n, f3 cA, f3 cB, float cd, float tol, Contacts& o)
{
	if (dist > cd)
		return 0;
	const float lo = tol * 1e-2f, hi = tol * 5e-2f;
	const float tA = fminf(fmaxf(A.margin, lo), hi), tB = fminf(fmaxf(B.margin, lo), hi);
	const f3 negN = -n;
	const f3 nIn0 = mtmul(aToB.r, n);
	const int f1 = witnessPolygon(B, negN, cB, tB);
	const int f0 = witnessPolygon(A, nIn0, mxtransinv(aToB, cA), tA);
	const f3 refN = normalize(B.pn(f1)), incN = normalize(A.pn(f0));
	PC mc[HULL_MAX_CONTACTS];
	int nc = 0;
	if (fabsf(dot(refN, negN)) >= fabsf(dot(incN, nIn0)))
	{
		const mxf bToA = mxfOf(xinvmul(x0, x1));
		generatedContacts(B, A, f1, f0, bToA, mc, nc, cd);
	}
	else
	{
		generatedContacts(A, B, f0, f1, aToB, mc, nc, cd);
		const f3 nn = -mmul(aToB.r, incN);
		for (int i = 0; i < nc; i++)
		{
			const f3 t = mc[i].lb;
			mc[i].lb = mc[i].la;
			mc[i].la = t;
			mc[i].ln = nn;
		}
	}
	if (nc == 0)
	{
		// the closest-feature / minimum-translation point stays the only point
		mc[0].la = mxtransinv(aToB, cA);
		mc[0].lb = cB;
		mc[0].ln = n;
		mc[0].pen = dist;
		nc = 1;
	}
	else
		nc = batchReduce(mc, nc, tol);
	return manifoldOut(mc, nc, x1, cd, o);
}

// Gu::pcmContactConvexConvex / pcmContactBoxConvex on a fresh manifold
// (geomutils/src/pcm/GuPCMContactConvexConvex.cpp:111-276, GuPCMContactBoxConvex.cpp;
// Gu::generateFullContactManifold and Gu::addGJKEPAContacts,
// GuPCMContactGenBoxConvex.cpp:532-720). A is shape0, B is shape1.
PXH_D int convexConvex(const Poly& A, const xf& x0, const Poly& B, const xf& x1, float cd, float tol, Contacts& o)
{
	const mxf aToB = mxfOf(xinvmul(x1, x0));
	float dist = 0.0f;
	f3 n = mk(1.0f, 0.0f, 0.0f), cA = mk(0.0f, 0.0f, 0.0f), cB = cA;
	polySeparation(A, aToB, B, cd, dist, n, cA, cB);
	return convexContactsFrom(A, x0, B, x1, aToB, dist, n, cA, cB, cd, tol, o);
}

// Gu::pcmContactSphereConvex on a fresh manifold
// (geomutils/src/pcm/GuPCMContactSphereConvex.cpp:110-246)
PXH_D int sphereConvex(float r, f3 p0, const Poly& B, const xf& x1, float cd, Contacts& o)
{
	const f3 c = xtransinv(x1, p0);
	float hmax = -FLT_BIG;
	int fm = 0;
	for (int f = 0; f < B.np; f++)
	{
		const float h = dot(B.pn(f), c) + B.pd(f);
		if (h > hmax) { hmax = h; fm = f; }
	}
	f3 n;
	float pen;
	if (hmax <= 0.0f)
	{
		n = B.pn(fm);
		pen = hmax - r;
	}
	else
	{
		float d2 = FLT_BIG;
		f3 q = c;
		for (int f = 0; f < B.np; f++)
		{
			const f3 bn = B.pn(f);
			const float h = dot(bn, c) + B.pd(f);
			if (h < 0.0f || h * h >= d2)
				continue;
			const f3 pq = c - bn * h;
			if (inPolygon(B, f, pq)) { d2 = h * h; q = pq; }
		}
		for (int e = 0; e < B.ne; e++)
		{
			const uint32_t g = B.edge(e);
			f3 c1, c2;
			const float s = segClosestSq(c, c, B.vert(int(g & 0xFF)), B.vert(int((g >> 8) & 0xFF)), c1, c2);
			if (s < d2) { d2 = s; q = c2; }
		}
		const float dist = sqrtf(d2);
		if (dist > r + cd)
			return 0;
		n = (c - q) * (1.0f / dist);
		pen = dist - r;
	}
	const f3 wn = qrot(x1.q, n);
	o.n = wn;
	o.p[0] = p0 - wn * r;
	o.sep[0] = pen;
	o.count = 1;
	return 1;
}

// Capsule core segment pq against polytope B (B's frame), standing in for
// GJK / EPA on the segment (geomutils/src/gjk/GuGJKPenetration.h:95-306 with
// takeCoreShape: points on the segment, distance without the radius). Apart:
// the exact closest points, returns dist > 0 with n = (cA - cB)/dist.
// Entering B: the minimum translation over B's face normals and the
// directions normal to the segment and each edge of B (the face normals of
// their Minkowski difference), returns -depth, cB = cA + n depth.
PXH_D float segPoly(const Poly& B, f3 p, f3 q, f3& n, f3& cA, f3& cB)
{
	const f3 d = q - p;
	float t0 = 0.0f, t1 = 1.0f;
	bool enters = true;
	cA = p;
	cB = p;
	for (int f = 0; f < B.np && enters; f++)
	{
		const f3 fn = B.pn(f);
		const float hp = dot(fn, p) + B.pd(f), dn = dot(fn, d);
		if (dn == 0.0f)
		{
			enters = !(hp > 0.0f);
			continue;
		}
		const float t = -hp / dn;
		if (dn < 0.0f)
			t0 = fmaxf(t0, t);
		else
			t1 = fminf(t1, t);
		enters = !(t0 > t1);
	}
	if (!enters)
	{
		float best = FLT_BIG;
		for (int k = 0; k < 2; k++)
		{
			const f3 c = k ? q : p;
			for (int f = 0; f < B.np; f++)
			{
				const f3 fn = B.pn(f);
				const float h = dot(fn, c) + B.pd(f);
				if (h < 0.0f || h * h >= best)
					continue;
				const f3 pc = c - fn * h;
				if (inPolygon(B, f, pc)) { best = h * h; cA = c; cB = pc; }
			}
		}
		for (int e = 0; e < B.ne; e++)
		{
			const uint32_t ed = B.edge(e);
			f3 c1, c2;
			const float s = segClosestSq(p, q, B.vert(ed & 0xFF), B.vert((ed >> 8) & 0xFF), c1, c2);
			if (s < best) { best = s; cA = c1; cB = c2; }
		}
		const float dist = sqrtf(best);
		n = (cA - cB) * (1.0f / dist);
		return dist;
	}
	float best = FLT_BIG;
	int bestEdge = -1;
	f3 axis = mk(1.0f, 0.0f, 0.0f);
	for (int f = 0; f < B.np; f++)
	{
		const f3 fn = B.pn(f);
		const float depth = -B.pd(f) - fminf(dot(fn, p), dot(fn, q));
		if (depth < best) { best = depth; axis = fn; }
	}
	for (int e = 0; e < B.ne; e++)
	{
		const uint32_t ed = B.edge(e);
		const f3 a = B.vert(ed & 0xFF), b = B.vert((ed >> 8) & 0xFF);
		f3 c = cross(d, b - a);
		const float l2 = dot(c, c);
		if (!(l2 > 1e-12f))
			continue;
		c = c * (1.0f / sqrtf(l2));
		float mnB = FLT_BIG, mxB = -FLT_BIG;
		for (int i = 0; i < B.nv; i++)
		{
			const float s = dot(c, B.vert(i));
			mnB = fminf(mnB, s);
			mxB = fmaxf(mxB, s);
		}
		const float sp = dot(c, p), sq = dot(c, q);
		const float dp = mxB - fminf(sp, sq), dm = fmaxf(sp, sq) - mnB;
		if (dp < best) { best = dp; axis = c; bestEdge = e; }
		if (dm < best) { best = dm; axis = -c; bestEdge = e; }
	}
	n = axis;
	if (bestEdge >= 0)
	{
		const uint32_t ed = B.edge(bestEdge);
		f3 c2;
		segClosestSq(p, q, B.vert(ed & 0xFF), B.vert((ed >> 8) & 0xFF), cA, c2);
	}
	else
		cA = dot(axis, q) < dot(axis, p) ? q : p;
	cB = cA + axis * best;
	return -best;
}

constexpr int CAPSULE_MAX_CONTACTS = HULL_MAX_POLY_VERTS + 2;

// generatedCapsuleBoxFaceContacts (geomutils/src/pcm/GuPCMContactGenSphereCapsule.cpp:221-284):
// each endpoint moved along -n onto the reference polygon's plane, kept when
// that distance is within radius + contact distance and the foot lies in the
// polygon
PXH_D void capsuleFaceContacts(const Poly& B, int f, f3 p0, f3 p1, f3 n, float infl, const mxf& aToB, PC* mc, int& nc)
{
	const f3 pn = normalize(B.pn(f));
	const int nv = B.pcount(f);
	const f3 a = B.vert(B.pvert(f, 0));
	const float proj = dot(pn, n);
	const float numer = proj > 0.0f ? 1.0f / proj : 0.0f;
	const float t[2] = {dot(pn, p0 - a) * numer, dot(pn, p1 - a) * numer};
	if (!(infl >= t[0] || infl >= t[1]))
		return;
	const m33 rot = rotFromZ(pn);
	f3 pts[HULL_MAX_POLY_VERTS];
	f3 mn = mk(FLT_BIG, FLT_BIG, FLT_BIG), mx = -mn;
	for (int i = 0; i < nv; i++)
	{
		pts[i] = mmul(rot, B.vert(B.pvert(f, i)));
		mn = mk(fminf(mn.x, pts[i].x), fminf(mn.y, pts[i].y), fminf(mn.z, pts[i].z));
		mx = mk(fmaxf(mx.x, pts[i].x), fmaxf(mx.y, pts[i].y), fmaxf(mx.z, pts[i].z));
	}
	for (int k = 0; k < 2; k++)
	{
		if (!(infl >= t[k]))
			continue;
		const f3 e = k ? p1 : p0;
		const f3 foot = e - n * t[k];
		if (polyContainsN(pts, nv, mmul(rot, foot), mn, mx))
		{
			mc[nc].la = mxtransinv(aToB, e);
			mc[nc].lb = foot;
			mc[nc].ln = n;
			mc[nc].pen = t[k];
			nc++;
		}
	}
}

// intersectRayPolyhedron (GuPCMContactGenSphereCapsule.cpp:98-141): the entry
// parameter of the ray a + t dir, t >= 0, into B
PXH_D bool rayPolyhedron(const Poly& B, f3 a, f3 dir, float& tEnter)
{
	float tFirst = 0.0f, tLast = FLT_BIG;
	for (int k = 0; k < B.np; k++)
	{
		const f3 fn = B.pn(k);
		const float den = dot(fn, dir);
		const float dist = dot(fn, a) + B.pd(k);
		if (1e-7f > fabsf(den))
		{
			if (dist > 0.0f)
				return false;
		}
		else
		{
			const float t = -(dist / den);
			if (0.0f > den && t > tFirst)
				tFirst = t;
			if (!(0.0f > den) && tLast > t)
				tLast = t;
		}
		if (tFirst > tLast)
			return false;
	}
	tEnter = tFirst;
	return true;
}

// generatedFaceContacts (GuPCMContactGenSphereCapsule.cpp:286-310): each
// endpoint's ray along -n into the hull, entered within radius + contact
// distance
PXH_D void capsuleRayContacts(const Poly& B, f3 p0, f3 p1, f3 n, float infl, const mxf& aToB, PC* mc, int& nc)
{
	for (int k = 0; k < 2; k++)
	{
		const f3 e = k ? p1 : p0;
		float tE;
		if (rayPolyhedron(B, e, -n, tE) && infl >= tE)
		{
			mc[nc].la = mxtransinv(aToB, e);
			mc[nc].lb = (-n) * tE + e;
			mc[nc].ln = n;
			mc[nc].pen = tE;
			nc++;
		}
	}
}

// generateEE / generatedContactsEEContacts (GuPCMContactGenSphereCapsule.cpp:312-377):
// where the segment crosses the plane through a reference polygon edge
// spanned with n, a contact against that edge (its parameter within
// [-0.005, 1.005]) when the signed distance along n is within the radius
PXH_D void capsuleEEContacts(const Poly& B, int f, f3 p, f3 q, f3 normal, float infl, const mxf& aToB, PC* mc, int& nc)
{
	const int nv = B.pcount(f);
	for (int rs = 0, re = nv - 1; rs < nv; re = rs++)
	{
		const f3 a = B.vert(B.pvert(f, rs)), b = B.vert(B.pvert(f, re));
		const f3 ab = b - a;
		const f3 n = cross(ab, normal);
		const float d = dot(n, a);
		const float sp = dot(n, p) - d, sq = dot(n, q) - d;
		if (sp * sq > 0.0f)
			continue;
		const f3 pq = q - p;
		const float npq = dot(n, pq);
		if (npq == 0.0f)
			continue;
		const f3 lpa = pq * ((d - dot(n, p)) / npq) + p;
		const f3 perN = cross(normal, pq);
		const f3 ap = lpa - a;
		const float tv = dot(perN, ap) / dot(perN, ab);
		if (tv > 1.005f || -0.005f > tv)
			continue;
		const f3 v = ap - ab * tv;
		const float sd = dot(v, normal);
		if (infl >= sd)
		{
			mc[nc].la = mxtransinv(aToB, lpa);
			mc[nc].lb = lpa - v;
			mc[nc].ln = normal;
			mc[nc].pen = sd;
			nc++;
		}
	}
}

// PersistentContactManifold::reduceBatchContacts2
// (geomutils/src/pcm/GuPersistentContactManifold.cpp:1177-1243): the deepest
// point, the point farthest from it, then the last point (in order) of those
// nearer the second than the first that is shallower than the second
PXH_D void reduceBatch2(const PC* m, int n, int& i0, int& i1)
{
	int idx = 0;
	float maxD = m[0].pen;
	for (int i = 1; i < n; i++)
		if (maxD > m[i].pen) { maxD = m[i].pen; idx = i; }
	i0 = idx;
	const f3 b0 = m[i0].lb;
	f3 v = m[0].lb - b0;
	float far = dot(v, v);
	idx = 0;
	for (int i = 1; i < n; i++)
	{
		v = m[i].lb - b0;
		const float d = dot(v, v);
		if (d > far) { far = d; idx = i; }
	}
	const int first = idx;
	const float maxDepth = m[first].pen;
	const f3 b1 = m[first].lb;
	int second = first;
	for (int i = 0; i < n; i++)
	{
		if (i == i0 || i == first)
			continue;
		const f3 d0 = b0 - m[i].lb, d1 = b1 - m[i].lb;
		if (dot(d0, d0) > dot(d1, d1) && maxDepth > m[i].pen)
			second = i;
	}
	i1 = second;
}

// Gu::pcmContactCapsuleBox / pcmContactCapsuleConvex on a fresh manifold
// (geomutils/src/pcm/GuPCMContactCapsuleBox.cpp:42-200,
// GuPCMContactCapsuleConvex.cpp:43-236, full contact generation
// GuPCMContactGenSphereCapsule.cpp:379-468): the core segment against the
// polytope; box: endpoint contacts on the witness polygon, convex: endpoint
// rays into the hull; fewer than two: edge contacts on the witness polygon;
// none: the closest-point contact; reduced to two. Points on the capsule's
// surface, separation less the radius, normal from the polytope.
PXH_D int capsulePoly(float r, float hh, const xf& x0, const Poly& B, const xf& x1, float cd, float tol, Contacts& o)
{
	const mxf aToB = mxfOf(xinvmul(x1, x0));
	const f3 ext = mmul(aToB.r, mk(hh, 0.0f, 0.0f));
	const f3 p0 = aToB.p + ext, p1 = aToB.p - ext;
	f3 n, cA, cB;
	const float dist = segPoly(B, p0, p1, n, cA, cB);
	if (dist > r + cd)
		return 0;
	const float infl = r + cd;
	const float wtol = fminf(fmaxf(B.margin, tol * 1e-2f), tol * 5e-2f);
	PC mc[CAPSULE_MAX_CONTACTS];
	int nc = 0;
	if (B.box)
	{
		const int f = witnessPolygon(B, -n, cB, wtol);
		capsuleFaceContacts(B, f, p0, p1, n, infl, aToB, mc, nc);
		if (nc < 2)
			capsuleEEContacts(B, f, p0, p1, n, infl, aToB, mc, nc);
	}
	else
	{
		capsuleRayContacts(B, p0, p1, n, infl, aToB, mc, nc);
		if (nc < 2)
			capsuleEEContacts(B, witnessPolygon(B, -n, cB, wtol), p0, p1, n, infl, aToB, mc, nc);
	}
	int i0 = 0, i1 = 1;
	if (nc == 0)
	{
		mc[0].la = mxtransinv(aToB, cA);
		mc[0].lb = cB;
		mc[0].ln = n;
		mc[0].pen = dist;
		nc = 1;
	}
	else if (nc > 2)
	{
		reduceBatch2(mc, nc, i0, i1);
		nc = 2;
	}
	const f3 wn = qrot(x1.q, n);
	o.n = wn;
	o.count = 0;
	for (int i = 0; i < nc; i++)
	{
		const PC& c = mc[i == 0 ? i0 : i1];
		const float sep = c.pen - r;
		if (cd >= sep)
		{
			o.p[o.count] = xtrans(x0, c.la) - wn * r;
			o.sep[o.count] = sep;
			o.count++;
		}
	}
	return o.count;
}

// Gu::pcmContactPlaneConvex on a fresh manifold
// (geomutils/src/pcm/GuPCMContactPlaneConvex.cpp:36-227): the face most
// opposed to the plane normal, or the better face of the edge whose summed
// face normals are most opposed to it (5e-4 bias towards edges); the chosen
// face's vertices below the contact distance, taken once more when an edge
// was chosen (the reference passes over the same face twice); reduced to
// four. Normal from the hull towards the plane.
PXH_D int planeConvex(f3 pn, float pd, const Poly& B, const xf& tb, float cd, float tol, Contacts& o)
{
	const f3 n = normalize(qrotinv(tb.q, pn));
	const f3 nn = -n;
	float minProj = FLT_BIG;
	int face = 0;
	for (int i = 0; i < B.np; i++)
	{
		const float proj = dot(n, B.pn(i));
		if (minProj > proj) { minProj = proj; face = i; }
	}
	int edgeSel = -1;
	minProj = minProj - 5e-4f;
	float maxDpSq = minProj * minProj;
	for (int e = 0; e < B.ne; e++)
	{
		const uint32_t g = B.edge(e);
		const f3 en = B.pn(int((g >> 16) & 0xFF)) + B.pn(int(g >> 24));
		const float enMagSq = dot(en, en);
		const float dp = dot(en, nn);
		const float sqDp = dp * dp;
		if (dp >= 0.0f && sqDp > maxDpSq * enMagSq)
		{
			maxDpSq = sqDp / enMagSq;
			edgeSel = e;
		}
	}
	int passes = 1;
	if (edgeSel >= 0)
	{
		const uint32_t g = B.edge(edgeSel);
		const int f0 = int((g >> 16) & 0xFF), f1 = int(g >> 24);
		face = dot(B.pn(f0), nn) > dot(B.pn(f1), nn) ? f0 : f1;
		passes = 2;
	}
	PC mc[HULL_MAX_CONTACTS];
	int nc = 0;
	float deepest = FLT_BIG;
	int dv = 0;
	for (int i = 0; i < B.nv; i++)
	{
		const float s = dot(pn, xtrans(tb, B.vert(i))) + pd;
		if (s < deepest) { deepest = s; dv = i; }
	}
	bool deepestTaken = false;
	const int cnt = B.pcount(face);
	for (int pass = 0; pass < passes; pass++)
		for (int k = 0; k < cnt; k++)
		{
			const int vi = B.pvert(face, k);
			const f3 lv = B.vert(vi);
			const f3 pw = xtrans(tb, lv);
			const float s = dot(pn, pw) + pd;
			if (cd > s && nc < HULL_MAX_CONTACTS)
			{
				mc[nc].la = lv;
				mc[nc].lb = pw - pn * s;
				mc[nc].ln = pn;
				mc[nc].pen = s;
				nc++;
				deepestTaken |= vi == dv;
			}
		}
	// Deliberate deviation: when the chosen face does not hold the hull's
	// deepest vertex (an irregular hull can rest on a vertex that the face most
	// opposed to the plane does not contain), that vertex joins the batch. The
	// reference scene's persistent manifold keeps such a vertex from earlier
	// steps; a step that generates contacts afresh has to add it.
	if (!deepestTaken && cd > deepest && nc < HULL_MAX_CONTACTS)
	{
		const f3 lv = B.vert(dv);
		const f3 pw = xtrans(tb, lv);
		mc[nc].la = lv;
		mc[nc].lb = pw - pn * deepest;
		mc[nc].ln = pn;
		mc[nc].pen = deepest;
		nc++;
	}
	nc = batchReduce(mc, nc, tol);
	o.n = -pn;
	int k = 0;
	for (int i = 0; i < nc; i++)
		if (cd >= mc[i].pen)
		{
			o.p[k] = mc[i].lb;
			o.sep[k] = mc[i].pen;
			k++;
		}
	o.count = k;
	return k;
}

}  // namespace pxh

namespace pxh {

// ---------------------------------------------------------------------------
// One wavefront per convex pair: the 64 lanes split the face-axis, edge-pair
// and closest-feature searches of polySeparation over the pair's vertices,
// planes and edges staged in LDS; lane 0 then picks the witness polygons,
// clips and reduces. Every search keeps polySeparation's order (the first
// minimum in loop order wins), so the result is the same as the per-lane
// path's.
// ---------------------------------------------------------------------------
constexpr int CVX_WAVE_V = 64, CVX_WAVE_P = 128, CVX_WAVE_E = 192;

struct CvxLds
{
	float4 va[CVX_WAVE_V];   // A's vertices in B's frame
	float4 vb[CVX_WAVE_V];   // B's vertices
	float4 pa[CVX_WAVE_P];   // A's planes in B's frame
	float4 pb[CVX_WAVE_P];   // B's planes
	uint32_t polyA[CVX_WAVE_P], polyB[CVX_WAVE_P];
	uint32_t ea[CVX_WAVE_E];
	uint32_t eb[CVX_WAVE_E];
	uint8_t vrA[2 * CVX_WAVE_E], vrB[2 * CVX_WAVE_E];   // a closed polytope has 2 vrefs per edge
	uint8_t elA[CVX_WAVE_E], elB[CVX_WAVE_E];           // edges kept for the closest-feature search
};

PXH_D bool fitsWave(const Poly& A, const Poly& B)
{
	return A.nv <= CVX_WAVE_V && B.nv <= CVX_WAVE_V && A.np <= CVX_WAVE_P && B.np <= CVX_WAVE_P &&
	       A.ne <= CVX_WAVE_E && B.ne <= CVX_WAVE_E;
}

// LDS written by some lanes of the wave becomes visible to the others
PXH_D void waveSync()
{
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
	__builtin_amdgcn_wave_barrier();
}

// smallest value over the wave, the lowest order index among equal values
PXH_D void waveArgMin(float& v, uint32_t& idx)
{
	for (int off = 32; off >= 1; off >>= 1)
	{
		const float ov = __shfl_xor(v, off);
		const uint32_t oi = __shfl_xor(idx, off);
		if (ov < v || (ov == v && oi < idx))
		{
			v = ov;
			idx = oi;
		}
	}
}

PXH_D bool waveAny(bool b) { return __ballot(b) != 0ull; }

PXH_D f3 ldsV(const float4* a, int i) { return mk(a[i].x, a[i].y, a[i].z); }

// (i, j) of the flat index k = i * n + j as a lane walks k = lane, lane + 64,
// ...: one division up front, then a step of 64 costs an add and one compare
// (a runtime integer division is a long instruction sequence on this target).
struct FlatPair
{
	int n, q, r, i, j;
	PXH_D FlatPair(int lane, int n_) : n(n_ > 0 ? n_ : 1)
	{
		q = 64 / n;
		r = 64 - q * n;
		i = lane / n;
		j = lane - i * n;
	}
	PXH_D void next()
	{
		i += q;
		j += r;
		if (j >= n)
		{
			j -= n;
			i++;
		}
	}
};

// The edges of each polytope whose bounding sphere comes within lim of the
// other polytope's bounding box (both in B's frame, from the staged vertices),
// listed in index order in L.elA / L.elB; every edge when lim is FLT_BIG.
PXH_D void nearEdges(const Poly& A, const Poly& B, const CvxLds& L, int lane, float lim, int& nA, int& nB)
{
	const float BIG = FLT_BIG;
	f3 aMin = mk(BIG, BIG, BIG), aMax = -aMin, bMin = aMin, bMax = aMax;
	if (lane < A.nv)
	{
		aMin = ldsV(L.va, lane);
		aMax = aMin;
	}
	if (lane < B.nv)
	{
		bMin = ldsV(L.vb, lane);
		bMax = bMin;
	}
	for (int off = 32; off >= 1; off >>= 1)
	{
		aMin = mk(fminf(aMin.x, __shfl_xor(aMin.x, off)), fminf(aMin.y, __shfl_xor(aMin.y, off)),
		          fminf(aMin.z, __shfl_xor(aMin.z, off)));
		aMax = mk(fmaxf(aMax.x, __shfl_xor(aMax.x, off)), fmaxf(aMax.y, __shfl_xor(aMax.y, off)),
		          fmaxf(aMax.z, __shfl_xor(aMax.z, off)));
		bMin = mk(fminf(bMin.x, __shfl_xor(bMin.x, off)), fminf(bMin.y, __shfl_xor(bMin.y, off)),
		          fminf(bMin.z, __shfl_xor(bMin.z, off)));
		bMax = mk(fmaxf(bMax.x, __shfl_xor(bMax.x, off)), fmaxf(bMax.y, __shfl_xor(bMax.y, off)),
		          fmaxf(bMax.z, __shfl_xor(bMax.z, off)));
	}
	CvxLds& W = const_cast<CvxLds&>(L);
	nA = 0;
	nB = 0;
	for (int side = 0; side < 2; side++)
	{
		const int ne = side ? B.ne : A.ne;
		const f3 lo = side ? aMin : bMin, hi = side ? aMax : bMax;
		int cnt = 0;
		for (int base = 0; base < ne; base += 64)
		{
			const int e = base + lane;
			bool keep = false;
			if (e < ne)
			{
				const uint32_t g = side ? L.eb[e] : L.ea[e];
				const float4* vv = side ? L.vb : L.va;
				const f3 p0 = ldsV(vv, int(g & 0xFF)), p1 = ldsV(vv, int((g >> 8) & 0xFF));
				const f3 mid = (p0 + p1) * 0.5f, h = (p1 - p0) * 0.5f;
				const f3 q = mk(fminf(fmaxf(mid.x, lo.x), hi.x), fminf(fmaxf(mid.y, lo.y), hi.y),
				                fminf(fmaxf(mid.z, lo.z), hi.z));
				const f3 dq = mid - q;
				keep = !(sqrtf(dot(dq, dq)) - sqrtf(dot(h, h)) > lim);
			}
			const uint64_t m = __ballot(keep);
			if (keep)
			{
				const int at = cnt + int(__popcll(m & ((1ull << lane) - 1ull)));
				(side ? W.elB : W.elA)[at] = uint8_t(e);
			}
			cnt += int(__popcll(m));
		}
		(side ? nB : nA) = cnt;
	}
	waveSync();
}

// The face axes of polySeparation over the wave, from the staged vertices and
// planes only: B's (order 0 | f << 8 | vertex), then A's (order 1 | ...). The
// lane's least penetrating axis goes to best, bi; sep is set on every lane
// when some axis shows no penetration. True when some axis separates the
// polytopes by more than cd.
PXH_D bool faceAxesWave(const Poly& A, const Poly& B, float cd, const CvxLds& L, int lane, float& best, uint32_t& bi,
                        bool& sep)
{
	const float BIG = FLT_BIG;
	best = BIG;
	bi = 0xFFFFFFFFu;
	bool far = false;
	sep = false;
	for (int f = lane; f < B.np; f += 64)
	{
		const float4 pl = L.pb[f];
		float mn = BIG;
		int vi = 0;
		for (int i = 0; i < A.nv; i++)
		{
			const float4 v = L.va[i];
			const float h = pl.x * v.x + pl.y * v.y + pl.z * v.z + pl.w;
			if (h < mn) { mn = h; vi = i; }
		}
		far |= mn > cd;
		sep |= mn > 0.0f;
		if (!(mn > 0.0f) && -mn < best)
		{
			best = -mn;
			bi = (uint32_t(f) << 8) | uint32_t(vi);
		}
	}
	for (int f = lane; f < A.np; f += 64)
	{
		const float4 pl = L.pa[f];
		float mn = BIG;
		int vi = 0;
		for (int i = 0; i < B.nv; i++)
		{
			const float4 v = L.vb[i];
			const float h = pl.x * v.x + pl.y * v.y + pl.z * v.z + pl.w;
			if (h < mn) { mn = h; vi = i; }
		}
		far |= mn > cd;
		sep |= mn > 0.0f;
		if (!(mn > 0.0f) && -mn < best)
		{
			best = -mn;
			bi = (1u << 24) | (uint32_t(f) << 8) | uint32_t(vi);
		}
	}
	sep = waveAny(sep);
	return waveAny(far);
}

// The rest of polySeparation over the wave (same results), after
// faceAxesWave found no axis farther than cd; A, B in LDS as staged
PXH_D void polySeparationWave(const Poly& A, const mxf& aToB, const Poly& B, const CvxLds& L, int lane, float best,
                              uint32_t bi, bool sep, float& dist, f3& n, f3& cA, f3& cB)
{
	const float BIG = FLT_BIG;
	if (!sep)
	{
		f3 bmean = mk(0.0f, 0.0f, 0.0f);
		for (int i = 0; i < B.nv; i++)
			bmean = bmean + ldsV(L.vb, i);
		bmean = bmean * (1.0f / float(B.nv));
		// Edge pairs: each lane holds one of A's edges (its Gauss-map arc in
		// registers) and walks B's edges, so B's edge data are read as LDS
		// broadcasts. A block of R < 64 of A's edges gives each edge S = 64 / R
		// lanes, which split B's edges between them. A lane visits its pairs in
		// ascending index order, so with the lowest-index rule of waveArgMin
		// the pair chosen is the first minimum of the flat order.
		for (int base = 0; base < A.ne; base += 64)
		{
			const int R = min(64, A.ne - base), S = 64 / R, part = lane / R;
			if (part >= S)
				continue;
			const int ea = base + lane - part * R;
			const uint32_t e = L.ea[ea];
			const f3 na = ldsV(L.pa, int((e >> 16) & 0xFF)), nb = ldsV(L.pa, int(e >> 24));
			const f3 bxa = cross(nb, na);
			for (int ebi = part; ebi < B.ne; ebi += S)
			{
				const uint32_t g = L.eb[ebi];
				const f3 nc = -ldsV(L.pb, int((g >> 16) & 0xFF)), nd = -ldsV(L.pb, int(g >> 24));
				const f3 dxc = cross(nd, nc);
				const float cba = dot(nc, bxa), dba = dot(nd, bxa), adc = dot(na, dxc), bdc = dot(nb, dxc);
				if (!(cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f))
					continue;
				const f3 a0 = ldsV(L.va, int(e & 0xFF)), a1 = ldsV(L.va, int((e >> 8) & 0xFF));
				const f3 b0 = ldsV(L.vb, int(g & 0xFF)), b1 = ldsV(L.vb, int((g >> 8) & 0xFF));
				f3 ax = cross(a1 - a0, b1 - b0);
				const float l2 = dot(ax, ax);
				if (l2 < 1e-12f)
					continue;
				ax = ax * (1.0f / sqrtf(l2));
				if (dot(ax, b0 - bmean) < 0.0f)
					ax = -ax;
				const float ov = dot(ax, b0) - dot(ax, a0);
				if (ov < 0.0f)
					sep = true;
				else if (ov < best)
				{
					best = ov;
					bi = (2u << 24) | uint32_t(ea * CVX_WAVE_E + ebi);
				}
			}
		}
		sep = waveAny(sep);
	}
	if (!sep)
	{
		waveArgMin(best, bi);
		if (bi == 0xFFFFFFFFu)
		{
			dist = FLT_BIG;
			return;
		}
		const uint32_t kind = bi >> 24;
		if (kind == 2u)
		{
			const int k = int(bi & 0xFFFFFFu), ea = k / CVX_WAVE_E, ebi = k - ea * CVX_WAVE_E;
			const uint32_t e = L.ea[ea], g = L.eb[ebi];
			const f3 a0 = ldsV(L.va, int(e & 0xFF)), a1 = ldsV(L.va, int((e >> 8) & 0xFF));
			const f3 b0 = ldsV(L.vb, int(g & 0xFF)), b1 = ldsV(L.vb, int((g >> 8) & 0xFF));
			f3 bmean = mk(0.0f, 0.0f, 0.0f);
			for (int i = 0; i < B.nv; i++)
				bmean = bmean + ldsV(L.vb, i);
			bmean = bmean * (1.0f / float(B.nv));
			f3 ax = cross(a1 - a0, b1 - b0);
			ax = ax * (1.0f / sqrtf(dot(ax, ax)));
			if (dot(ax, b0 - bmean) < 0.0f)
				ax = -ax;
			f3 c1, c2;
			segClosestSq(a0, a1, b0, b1, c1, c2);
			n = ax;
			cA = c1;
			cB = c2;
		}
		else
		{
			const int f = int((bi >> 8) & 0xFFFFu), vi = int(bi & 0xFFu);
			if (kind == 0u)
			{
				const float4 pl = L.pb[f];
				const f3 bn = mk(pl.x, pl.y, pl.z), va = ldsV(L.va, vi);
				const float mn = dot(bn, va) + pl.w;
				n = bn;
				cA = va;
				cB = va - bn * mn;
			}
			else
			{
				const float4 pl = L.pa[f];
				const f3 an = mk(pl.x, pl.y, pl.z), vb = ldsV(L.vb, vi);
				const float mn = dot(an, vb) + pl.w;
				n = -an;
				cB = vb;
				cA = vb - an * mn;
			}
		}
		dist = -best;
		return;
	}
	// separated: closest features (B faces x A vertices, A faces x B vertices,
	// edge pairs), the first minimum in that order
	float d2 = BIG;
	uint32_t di = 0xFFFFFFFFu;
	FlatPair fb(lane, A.nv);
	for (int k = lane; k < B.np * A.nv; k += 64, fb.next())
	{
		const int f = fb.i, i = fb.j;
		const float4 pl = L.pb[f];
		const f3 bn = mk(pl.x, pl.y, pl.z), va = ldsV(L.va, i);
		const float h = dot(bn, va) + pl.w;
		if (h < 0.0f || h * h >= d2)
			continue;
		if (!inPolygon(B, f, va - bn * h))
			continue;
		d2 = h * h;
		di = uint32_t(k);
	}
	FlatPair fa(lane, B.nv);
	for (int k = lane; k < A.np * B.nv; k += 64, fa.next())
	{
		const int f = fa.i, i = fa.j;
		const float4 pl = L.pa[f];
		const f3 an = mk(pl.x, pl.y, pl.z), vb = ldsV(L.vb, i);
		const float h = dot(an, vb) + pl.w;
		if (h < 0.0f || h * h >= d2)
			continue;
		if (!inPolygon(A, f, mxtransinv(aToB, vb - an * h)))
			continue;
		d2 = h * h;
		di = (1u << 24) | uint32_t(k);
	}
	// Edge pairs. The face-vertex pairs bound the squared distance by U; an
	// edge whose bounding sphere lies farther than sqrt(U) (plus a margin well
	// above rounding) from the other polytope's bounding box cannot be part of
	// a pair that reaches U, so only the remaining edges are paired. Pair
	// indices keep their unpruned order, so the first minimum is unchanged.
	float U = d2;
	for (int off = 32; off >= 1; off >>= 1)
		U = fminf(U, __shfl_xor(U, off));
	const float lim = U < BIG ? sqrtf(U) + 1e-4f : BIG;
	int nA = 0, nB = 0;
	nearEdges(A, B, L, lane, lim, nA, nB);
	waveSync();
	FlatPair fe(lane, nB);
	for (int k = lane; k < nA * nB; k += 64, fe.next())
	{
		const int ea = L.elA[fe.i], ebi = L.elB[fe.j];
		const uint32_t e = L.ea[ea], g = L.eb[ebi];
		f3 c1, c2;
		const float s = segClosestSq(ldsV(L.va, int(e & 0xFF)), ldsV(L.va, int((e >> 8) & 0xFF)),
		                             ldsV(L.vb, int(g & 0xFF)), ldsV(L.vb, int((g >> 8) & 0xFF)), c1, c2);
		if (s < d2)
		{
			d2 = s;
			di = (2u << 24) | uint32_t(ea * B.ne + ebi);
		}
	}
	waveArgMin(d2, di);
	if (di == 0xFFFFFFFFu)
	{
		dist = FLT_BIG;   // no feature pair (an empty polytope)
		return;
	}
	const uint32_t kind = di >> 24, k = di & 0xFFFFFFu;
	if (kind == 0u)
	{
		const int f = int(k) / A.nv, i = int(k) - f * A.nv;
		const float4 pl = L.pb[f];
		const f3 bn = mk(pl.x, pl.y, pl.z), va = ldsV(L.va, i);
		cA = va;
		cB = va - bn * (dot(bn, va) + pl.w);
	}
	else if (kind == 1u)
	{
		const int f = int(k) / B.nv, i = int(k) - f * B.nv;
		const float4 pl = L.pa[f];
		const f3 an = mk(pl.x, pl.y, pl.z), vb = ldsV(L.vb, i);
		cB = vb;
		cA = vb - an * (dot(an, vb) + pl.w);
	}
	else
	{
		const int ea = int(k) / B.ne, ebi = int(k) - ea * B.ne;
		const uint32_t e = L.ea[ea], g = L.eb[ebi];
		segClosestSq(ldsV(L.va, int(e & 0xFF)), ldsV(L.va, int((e >> 8) & 0xFF)), ldsV(L.vb, int(g & 0xFF)),
		             ldsV(L.vb, int((g >> 8) & 0xFF)), cA, cB);
	}
	dist = sqrtf(d2);
	n = (cA - cB) * (1.0f / dist);
}

// polySeparation of a box- or convex-convex pair over the wave: dist and n
// (sep = dist, n xyz) and the witnesses cA, cB in B's frame, the same on every lane
PXH_D void convexSeparationWave(const Poly& A, const xf& x0, const Poly& B, const xf& x1, float cd, CvxLds& L,
                                int lane, float4& sep, f3& wA, f3& wB)
{
	const mxf aToB = mxfOf(xinvmul(x1, x0));
	for (int i = lane; i < A.nv; i += 64)
		L.va[i] = f4(mxtrans(aToB, A.vert(i)), 0.0f);
	for (int i = lane; i < B.nv; i += 64)
		L.vb[i] = f4(B.vert(i), 0.0f);
	for (int f = lane; f < A.np; f += 64)
	{
		const f3 an = mmul(aToB.r, A.pn(f));
		L.pa[f] = f4(an, A.pd(f) - dot(an, aToB.p));
	}
	for (int f = lane; f < B.np; f += 64)
		L.pb[f] = f4(B.pn(f), B.pd(f));
	waveSync();
	float best;
	uint32_t bi;
	bool sepd;
	if (faceAxesWave(A, B, cd, L, lane, best, bi, sepd))
	{
		// farther apart than the contact distance
		sep = make_float4(2.0f * cd + 1.0f, 1.0f, 0.0f, 0.0f);
		wA = wB = mk(0.0f, 0.0f, 0.0f);
		waveSync();
		return;
	}
	// The edges in LDS for both searches, and for separated polytopes (the
	// polygon tests of the closest-feature search) the polygon and vertex-ref
	// arrays too, read through the same accessors; pairs rejected by a face
	// axis above need neither, overlapping ones no polygons. A's vertices and
	// planes in its own frame stay in global memory (cached): keeping them
	// out of LDS lets four waves share a SIMD.
	for (int e = lane; e < A.ne; e += 64)
		L.ea[e] = A.edge(e);
	for (int e = lane; e < B.ne; e += 64)
		L.eb[e] = B.edge(e);
	Poly Al = A, Bl = B;
	if (!A.box)
	{
		if (sepd)
		{
			for (int f = lane; f < A.np; f += 64)
				L.polyA[f] = A.poly[f];
			for (int k = lane; k < 2 * A.ne; k += 64)
				L.vrA[k] = A.vr[k];
			Al.poly = L.polyA; Al.vr = L.vrA;
		}
		Al.ed = L.ea;
	}
	if (sepd)
	{
		for (int f = lane; f < B.np; f += 64)
			L.polyB[f] = B.poly[f];
		for (int k = lane; k < 2 * B.ne; k += 64)
			L.vrB[k] = B.vr[k];
		Bl.poly = L.polyB; Bl.vr = L.vrB;
	}
	Bl.v = L.vb; Bl.pl = L.pb; Bl.ed = L.eb;
	waveSync();
	float dist = 0.0f;
	f3 n = mk(1.0f, 0.0f, 0.0f), cA = mk(0.0f, 0.0f, 0.0f), cB = cA;
	polySeparationWave(Al, aToB, Bl, L, lane, best, bi, sepd, dist, n, cA, cB);
	waveSync();   // the staged data stay in use until every lane is done
	sep.x = dist;
	sep.y = n.x; sep.z = n.y; sep.w = n.z;
	wA = cA;
	wB = cB;
}

}  // namespace pxh
