// Narrowphase contact generation, one pair per lane. Each generator follows
// the reference's PCM contact generation on a fresh manifold, so a pair gives
// the same contacts as PxGenerateContacts with an empty PxCache
// (physx/include/PxImmediateMode.h:267); normals point from shape1 to shape0.
#pragma once
#include "pxh_math.h"

namespace pxh {

enum { GEOM_SPHERE = 0, GEOM_PLANE = 1, GEOM_CAPSULE = 2, GEOM_BOX = 3 };  // GEOM_CONVEXMESH = 5 in pxh_convex.h
constexpr float FLT_BIG = 3.402823466e+38f;

struct Contacts
{
	f3 n;
	int count;
	f3 p[4];
	float sep[4];
};

struct PC  // manifold point before reduction (reference PersistentContact)
{
	f3 la, lb, ln;
	float pen;
};

// geomutils/src/pcm/GuPCMContactSphereSphere.cpp:36-67
PXH_D int sphereSphere(float r0, f3 p0, float r1, f3 p1, float cd, Contacts& o)
{
	const f3 d = p0 - p1;
	const float dist2 = dot(d, d);
	const float rs = r0 + r1;
	const float inf = rs + cd;
	if (!(inf * inf > dist2))
		return 0;
	const float dist = sqrtf(dist2);
	const f3 n = (1e-5f >= dist) ? mk(1.0f, 0.0f, 0.0f) : d * (1.0f / dist);
	o.n = n;
	o.p[0] = n * r1 + p1;
	o.sep[0] = dist - rs;
	o.count = 1;
	return 1;
}

// geomutils/src/pcm/GuPCMContactSpherePlane.cpp:36-75 (plane n.x + d = 0)
PXH_D int spherePlane(float r, f3 p0, f3 pn, float pd, float cd, Contacts& o)
{
	const float sep = (dot(pn, p0) + pd) - r;
	if (!(cd >= sep))
		return 0;
	o.n = pn;
	o.p[0] = p0 - pn * r;
	o.sep[0] = sep;
	o.count = 1;
	return 1;
}

// geomutils/src/pcm/GuPCMContactSphereBox.cpp:37-130
PXH_D int sphereBox(float r, f3 p0, f3 h, const xf& t1, float cd, Contacts& o)
{
	const f3 c = xtransinv(t1, p0);
	const f3 p = mk(fminf(fmaxf(c.x, -h.x), h.x), fminf(fmaxf(c.y, -h.y), h.y), fminf(fmaxf(c.z, -h.z), h.z));
	const f3 v = c - p;
	const float l2 = dot(v, v);
	const float inf = r + cd;
	if (!(inf * inf > l2))
		return 0;
	if (h.x >= fabsf(c.x) && h.y >= fabsf(c.y) && h.z >= fabsf(c.z))
	{
		const f3 ds = h - vabs(p);
		const bool con0 = ds.x >= ds.z && ds.y >= ds.z;
		const bool con1 = ds.y >= ds.x && ds.z >= ds.x;
		f3 ln;
		float dist;
		if (con0) { ln = mk(0.0f, 0.0f, p.z < 0.0f ? -1.0f : 1.0f); dist = -ds.z; }
		else if (con1) { ln = mk(p.x < 0.0f ? -1.0f : 1.0f, 0.0f, 0.0f); dist = -ds.x; }
		else { ln = mk(0.0f, p.y < 0.0f ? -1.0f : 1.0f, 0.0f); dist = -ds.y; }
		const f3 n = qrot(t1.q, ln);
		o.n = n;
		o.sep[0] = dist - r;
		o.p[0] = p0 - n * dist;
	}
	else
	{
		const float rl = 1.0f / sqrtf(l2);
		const float l = 1.0f / rl;
		o.n = qrot(t1.q, v * rl);
		o.sep[0] = l - r;
		o.p[0] = xtrans(t1, p);
	}
	o.count = 1;
	return 1;
}

// Capsules: the segment along the local x axis, endpoints p +- x*halfHeight
// (geomutils/src/GuInternal.h:60-70), inflated by the radius. The vector
// library's min/max take the second operand when the comparison is unordered.
PXH_D float vmin(float a, float b) { return a < b ? a : b; }
PXH_D float vmax(float a, float b) { return a > b ? a : b; }
PXH_D float clamp01(float x) { return vmax(vmin(x, 1.0f), 0.0f); }

PXH_D void capsuleSegment(const xf& t, float hh, f3& s, f3& e)
{
	const f3 tmp = qrot(t.q, mk(1.0f, 0.0f, 0.0f)) * hh;
	s = t.p + tmp;
	e = t.p - tmp;
}

// Gu::distanceSegmentSegmentSquared, vector form
// (geomutils/src/distance/GuDistanceSegmentSegment.cpp:411-470): parallel
// segments (a e - b^2 below FLT_EPSILON) take s = 1/2 before the clamped
// refinement
PXH_D float segSegSq(f3 p1, f3 d1, f3 p2, f3 d2, float& s, float& t)
{
	const float eps = 1.1920929e-07f;
	const f3 r = p1 - p2;
	const float a = dot(d1, d1), e = dot(d2, d2), b = dot(d1, d2), c = dot(d1, r);
	const float aR = a > eps ? 1.0f / a : 0.0f;
	const float eR = e > eps ? 1.0f / e : 0.0f;
	const float f = dot(d2, r);
	const float denom = a * e - b * b;
	const float temp = b * f - c * e;
	const float sT = eps > denom ? 0.5f : clamp01(temp / denom);
	t = clamp01((b * sT + f) * eR);
	s = clamp01((b * t - c) * aR);
	const f3 vv = (d1 * s + p1) - (d2 * t + p2);
	return dot(vv, vv);
}

// geomutils/src/pcm/GuPCMContactSphereCapsule.cpp:38-96
PXH_D int sphereCapsule(float r, f3 c, float cr, float hh, const xf& t1, float cd, Contacts& o)
{
	f3 s, e;
	capsuleSegment(t1, hh, s, e);
	const float rs = r + cr, inf = rs + cd;
	const f3 ap = c - s, ab = e - s;
	const float nom = dot(ap, ab), den = dot(ab, ab);
	const float t = den == 0.0f ? 0.0f : clamp01(nom / den);
	const f3 v = ap - ab * t;
	const float d2 = dot(v, v);
	if (!(inf * inf > d2))
		return 0;
	const f3 dir = c - ((e - s) * t + s);
	const float l = sqrtf(dot(dir, dir));
	const f3 n = l > 1.1920929e-07f ? dir * (1.0f / l) : mk(1.0f, 0.0f, 0.0f);
	o.n = n;
	o.p[0] = c - n * r;
	o.sep[0] = sqrtf(d2) - rs;
	o.count = 1;
	return 1;
}

// geomutils/src/pcm/GuPCMContactCapsuleCapsule.cpp:81-283: the closest points
// of the two segments, or, for segments within 0.9998 of parallel, the
// endpoints projecting inside the other segment (up to four points), about the
// midpoint of the two centres
PXH_D int capsuleCapsule(float r0, float hh0, const xf& x0, float r1, float hh1, const xf& x1, float cd, Contacts& o)
{
	const f3 off = (x0.p + x1.p) * 0.5f;
	xf a = x0, b = x1;
	a.p = x0.p - off;
	b.p = x1.p - off;
	f3 s0, e0, s1, e1;
	capsuleSegment(a, hh0, s0, e0);
	capsuleSegment(b, hh1, s1, e1);
	const f3 d0 = e0 - s0, d1 = e1 - s1;
	const float sumR = r0 + r1, inf = sumR + cd, inf2 = inf * inf;
	const float la = dot(d0, d0), le = dot(d1, d1);
	const float eps = 1e-6f;
	float t0, t1;
	const float sq0 = segSegSq(s0, d0, s1, d1, t0, t1);
	if (!(inf2 >= sq0))
		return 0;
	o.count = 0;
	const f3 dir0 = eps > la ? mk(0.0f, 0.0f, 0.0f) : d0 * (1.0f / sqrtf(la));
	const f3 dir1 = eps > le ? mk(0.0f, 0.0f, 0.0f) : d1 * (1.0f / sqrtf(le));
	if (fabsf(dot(dir0, dir1)) > 0.9998f)
	{
		// s1, e1 projected on s0e0; s0, e0 projected on s1e1 (:40-74)
		for (int k = 0; k < 4; k++)
		{
			const f3 pt = k == 0 ? s1 : (k == 1 ? e1 : (k == 2 ? s0 : e0));
			const float tv = k < 2 ? (la == 0.0f ? 0.0f : dot(pt - s0, d0) / la)
			                       : (le == 0.0f ? 0.0f : dot(pt - s1, d1) / le);
			if (!(tv >= 0.0f && 1.0f >= tv))
				continue;
			const f3 on0 = k < 2 ? d0 * tv + s0 : pt;
			const f3 on1 = k < 2 ? pt : d1 * tv + s1;
			const f3 v = on0 - on1;
			const float sq = dot(v, v);
			if (!(sq > eps && inf2 > sq))
				continue;
			const float dist = sqrtf(sq);
			const f3 n = v * (1.0f / dist);
			o.n = n;
			o.p[o.count] = (on0 - n * r0) + off;
			o.sep[o.count] = dist - sumR;
			o.count++;
		}
		if (o.count)
			return o.count;
	}
	const f3 cA = d0 * t0 + s0, cB = d1 * t1 + s1;
	const bool con = eps > sq0;
	const f3 n = normalize(con ? (la > eps ? d0 : mk(1.0f, 0.0f, 0.0f)) : cA - cB);
	o.n = n;
	o.p[0] = (cA - n * r0) + off;
	o.sep[0] = (con ? 0.0f : sqrtf(sq0)) - sumR;
	o.count = 1;
	return 1;
}

// geomutils/src/pcm/GuPCMContactPlaneCapsule.cpp:36-114 with a fresh
// manifold: each endpoint closer to the plane than radius + contact distance,
// the second replacing the first when they coincide within radius/1000
// (addManifoldPoint2, GuPersistentContactManifold.cpp:1269); points on the
// capsule surface, separation = endpoint height - radius, normal from the
// capsule towards the plane
PXH_D int planeCapsule(f3 pn, float pd, float r, float hh, const xf& tb, float cd, Contacts& o)
{
	f3 ends[2];
	capsuleSegment(tb, hh, ends[0], ends[1]);
	const float repl = r * 0.001f;
	f3 pt[2];
	float sd[2];
	int n = 0;
	for (int k = 0; k < 2; k++)
	{
		const float h = dot(pn, ends[k]) + pd;
		if (!(r + cd > h))
			continue;
		if (n == 1)
		{
			const f3 da = pt[0] - ends[k];
			const f3 db = (pt[0] - pn * sd[0]) - (ends[k] - pn * h);
			if (repl * repl > fminf(dot(db, db), dot(da, da)))
			{
				pt[0] = ends[k];
				sd[0] = h;
				continue;
			}
		}
		pt[n] = ends[k];
		sd[n] = h;
		n++;
	}
	o.n = -pn;
	o.count = 0;
	for (int k = 0; k < n; k++)
		if (cd >= sd[k] - r)
		{
			o.p[o.count] = pt[k] - pn * r;
			o.sep[o.count] = sd[k] - r;
			o.count++;
		}
	return o.count;
}

// PersistentContactManifold::reduceBatchContactsCluster
// (geomutils/src/pcm/GuPersistentContactManifold.cpp:859-1004)
PXH_D void reduceCluster(const PC* m, int n, int* idx)
{
	bool chosen[8] = {false, false, false, false, false, false, false, false};
	float maxDist = FLT_BIG;
	int index = 0;
	for (int i = 0; i < n; i++)
		if (maxDist > m[i].pen) { maxDist = m[i].pen; index = i; }
	const int i0 = index;
	chosen[i0] = true;
	f3 v = m[0].lb - m[i0].lb;
	maxDist = dot(v, v);
	index = 0;
	for (int i = 1; i < n; i++)
	{
		v = m[i].lb - m[i0].lb;
		const float d = dot(v, v);
		if (d > maxDist) { maxDist = d; index = i; }
	}
	const int i1 = index;
	chosen[i1] = true;
	maxDist = -FLT_BIG;
	index = 0;
	v = m[i1].lb - m[i0].lb;
	f3 nrm = cross(v, m[i0].ln);
	const float sq = dot(nrm, nrm);
	if (sq > 0.0f) { const float L = sqrtf(sq); nrm = mk(nrm.x / L, nrm.y / L, nrm.z / L); } else nrm = m[i0].ln;
	float minDist = FLT_BIG;
	int index1 = 0;
	for (int i = 0; i < n; i++)
		if (!chosen[i])
		{
			v = m[i].lb - m[i0].lb;
			const float d = dot(v, nrm);
			if (d > maxDist) { maxDist = d; index = i; }
			if (minDist > d) { minDist = d; index1 = i; }
		}
	const int i2 = index;
	chosen[i2] = true;
	if (minDist * maxDist > 0.0f)
	{
		maxDist = -FLT_BIG;
		for (int i = 0; i < n; i++)
			if (!chosen[i])
			{
				v = m[i].lb - m[i0].lb;
				const float d = dot(v, nrm);
				if (d > maxDist) { maxDist = d; index1 = i; }
			}
	}
	const int i3 = index1;
	chosen[i3] = true;
	idx[0] = i0; idx[1] = i1; idx[2] = i2; idx[3] = i3;
	const int orig[4] = {i0, i1, i2, i3};
	for (int i = 0; i < n; i++)
		if (!chosen[i])
		{
			float md = FLT_BIG;
			int k = 0;
			for (int j = 0; j < 4; j++)
			{
				const f3 w = m[i].lb - m[orig[j]].lb;
				const float d = dot(w, w);
				if (md > d) { md = d; k = j; }
			}
			if (m[idx[k]].pen > m[i].pen)
				idx[k] = i;
		}
}

// geomutils/src/pcm/GuPCMContactPlaneBox.cpp:36-208 with shape0 the plane,
// shape1 the box; the normal returned points from the box to the plane.
PXH_D int planeBox(f3 pn, float pd, f3 h, const xf& tb, float cd, Contacts& o)
{
	PC mc[8];
	int nc = 0;
	const m33 R = qmat(tb.q);
	const f3 r0 = R.c0 * h.x, r1 = R.c1 * h.y, r2 = R.c2 * h.z;
	const float px = dot(pn, tb.p) + pd;
#pragma unroll
	for (int k = 0; k < 8; k++)
	{
		// corner order of the reference: (x,y,z) (x,y,-z) (x,-y,z) (x,-y,-z) (-x,y,z) ...
		const float sx = (k & 4) ? -1.0f : 1.0f, sy = (k & 2) ? -1.0f : 1.0f, sz = (k & 1) ? -1.0f : 1.0f;
		const f3 off = r0 * sx + r1 * sy + r2 * sz;
		const float s = dot(pn, off);
		if (cd - px > s)
		{
			const float pen = s + px;
			mc[nc].lb = (tb.p + off) - pn * pen;
			mc[nc].ln = pn;
			mc[nc].pen = pen;
			nc++;
		}
	}
	if (nc == 0)
		return 0;
	int idx[4] = {0, 1, 2, 3};
	if (nc > 4)
	{
		reduceCluster(mc, nc, idx);
		nc = 4;
	}
	o.n = -pn;
	o.count = 0;
	for (int i = 0; i < nc; i++)
		if (cd >= mc[idx[i]].pen)
		{
			o.p[o.count] = mc[idx[i]].lb;
			o.sep[o.count] = mc[idx[i]].pen;
			o.count++;
		}
	return o.count;
}

// ------------------------------------------------------------------------
// box-box (geomutils/src/pcm/GuPCMContactBoxBox.cpp)

// getIncidentPolygon (:42-124)
PXH_D void incidentPolygon(f3* pts, f3& faceNormal, f3 axis, const mxf& t, f3 e)
{
	const f3 u0 = t.r.c0, u1 = t.r.c1, u2 = t.r.c2;
	const float d0 = dot(u0, axis), d1 = dot(u1, axis), d2 = dot(u2, axis);
	const float a0 = fabsf(d0), a1 = fabsf(d1), a2 = fabsf(d2);
	f3 ua, ub, uc;     // incident face axis, the two in-face axes
	float ea, eb, ec, dd;
	if (a0 >= a1 && a0 >= a2) { ua = u0; ub = u1; uc = u2; ea = e.x; eb = e.y; ec = e.z; dd = d0; }
	else if (a1 >= a2) { ua = u1; ub = u0; uc = u2; ea = e.y; eb = e.x; ec = e.z; dd = d1; }
	else { ua = u2; ub = u0; uc = u1; ea = e.z; eb = e.x; ec = e.y; dd = d2; }
	const bool con = dd > 0.0f;
	faceNormal = con ? -ua : ua;
	ea = con ? -ea : ea;
	const f3 r0 = ua * ea, r1 = ub * eb, r2 = uc * ec;
	const f3 t0 = t.p + r0, t1 = r1 + r2, t2 = r1 - r2;
	pts[0] = t0 + t1; pts[1] = t0 + t2; pts[2] = t0 - t1; pts[3] = t0 - t2;
}

// intersectSegmentAABB (:127-167); exact division where the reference uses a
// fast reciprocal estimate
PXH_D bool segAabb(f3 p0, f3 d, f3 mx, f3 mn, float& tmin, float& tmax)
{
	float ft1 = -FLT_BIG, ft2 = FLT_BIG;
#pragma unroll
	for (int i = 0; i < 3; i++)
	{
		const float pi = comp(p0, i), di = comp(d, i), mxi = comp(mx, i), mni = comp(mn, i);
		const bool par = 1e-6f > fabsf(di);
		if (par && ((pi > mxi) || (mni > pi)))
			return false;
		const float od = 1.0f / di;
		const float a = par ? 0.0f : (mni - pi) * od;
		const float b = par ? FLT_BIG : (mxi - pi) * od;
		ft1 = fmaxf(ft1, fminf(a, b));
		ft2 = fminf(ft2, fmaxf(a, b));
	}
	tmin = fmaxf(ft1, 0.0f);
	tmax = fminf(1.0f, ft2);
	return !((tmin > tmax) || (tmin > 1.0f));
}

// Gu::contains (geomutils/src/pcm/GuPCMContactGenUtil.cpp:35-103)
PXH_D bool polyContains(const f3* verts, f3 p, f3 mn, f3 mx)
{
	if ((mn.x > p.x || p.x > mx.x) || (mn.y > p.y || p.y > mx.y))
		return false;
	const float tx = p.x, ty = p.y, eps = 1.1920928955078125e-7f;
	int inter = 0;
	for (int i = 0, j = 3; i < 4; j = i++)
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

// calculateContacts (:170-330): up to 16 manifold points in reference-face space
PXH_D void calcContacts(float exIn, float eyIn, const f3* pts, f3 incN, f3 ln, PC* mc, int& nc, float cd)
{
	const float ex = exIn * 1.0001f, ey = eyIn * 1.0001f;
	bool pen[4], area[4];
	f3 bmin = mk(FLT_BIG, FLT_BIG, FLT_BIG), bmax = mk(-FLT_BIG, -FLT_BIG, -FLT_BIG);
#pragma unroll
	for (int i = 0; i < 4; i++)
	{
		bmin = mk(fminf(bmin.x, pts[i].x), fminf(bmin.y, pts[i].y), fminf(bmin.z, pts[i].z));
		bmax = mk(fmaxf(bmax.x, pts[i].x), fmaxf(bmax.y, pts[i].y), fmaxf(bmax.z, pts[i].z));
		const float z = -pts[i].z;
		pen[i] = cd > z;
		area[i] = pen[i] && ex >= fabsf(pts[i].x) && ey >= fabsf(pts[i].y);
		if (area[i])
		{
			mc[nc].la = mk(pts[i].x, pts[i].y, 0.0f);
			mc[nc].lb = pts[i];
			mc[nc].ln = ln;
			mc[nc].pen = z;
			nc++;
		}
	}
	if (nc == 4)
		return;
	const float denom = incN.z;
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		const f3 q0 = mk((k & 2) ? -ex : ex, (k & 1) ? -ey : ey, 0.0f);
		if (polyContains(pts, q0, bmin, bmax))
		{
			const float t = dot(incN, pts[0] - q0) / denom;
			if (cd > -t)
			{
				mc[nc].la = q0;
				mc[nc].lb = mk(q0.x, q0.y, t);
				mc[nc].ln = ln;
				mc[nc].pen = -t;
				nc++;
			}
		}
	}
	const f3 ext = mk(ex, ey, FLT_BIG);
	const f3 negExt = mk(-ex, -ey, -(cd + 1.1920928955078125e-7f));
	for (int rs = 0, re = 3; rs < 4; re = rs++)
	{
		if (!pen[rs] && !pen[re])
			continue;
		const bool c0 = area[rs], c1 = area[re];
		if (c0 && c1)
			continue;
		const f3 p0 = pts[rs];
		const f3 d = pts[re] - p0;
		float tmin, tmax;
		if (segAabb(p0, d, ext, negExt, tmin, tmax))
		{
			if (!c0)
			{
				const f3 ip = d * tmin + p0;
				mc[nc].la = mk(ip.x, ip.y, 0.0f); mc[nc].lb = ip; mc[nc].ln = ln; mc[nc].pen = -ip.z; nc++;
			}
			if (!c1)
			{
				const f3 ip = d * tmax + p0;
				mc[nc].la = mk(ip.x, ip.y, 0.0f); mc[nc].lb = ip; mc[nc].ln = ln; mc[nc].pen = -ip.z; nc++;
			}
		}
	}
}

// PersistentContactManifold::reduceBatchContacts
// (geomutils/src/pcm/GuPersistentContactManifold.cpp:1008-1172)
template <int MAXN>
PXH_D void reduceBatchT(const PC* m, int n, float tol, int* ch)
{
	int cand[MAXN];
	float maxPen = m[0].pen, minPen = maxPen;
	int index = 0, ci = 0;
	cand[0] = 0;
	int nc = n;
	for (int i = 1; i < n; i++)
	{
		cand[i] = i;
		const float pen = m[i].pen;
		minPen = fmaxf(minPen, pen);
		if (maxPen > pen) { maxPen = pen; index = i; ci = i; }
	}
	ch[0] = index;
	nc--; cand[ci] = cand[nc];
	f3 v = m[cand[0]].lb - m[ch[0]].lb;
	float maxDist = dot(v, v);
	index = cand[0]; ci = 0;
	for (int i = 1; i < nc; i++)
	{
		v = m[cand[i]].lb - m[ch[0]].lb;
		const float d = dot(v, v);
		if (d > maxDist) { maxDist = d; index = cand[i]; ci = i; }
	}
	ch[1] = index;
	nc--; cand[ci] = cand[nc];
	v = m[ch[1]].lb - m[ch[0]].lb;
	const f3 cn0 = m[ch[0]].ln;
	f3 nrm = cross(v, cn0);
	const float sq = dot(nrm, nrm);
	if (sq > 0.0f) { const float L = sqrtf(sq); nrm = mk(nrm.x / L, nrm.y / L, nrm.z / L); } else nrm = cn0;
	maxDist = -FLT_BIG; index = cand[0]; ci = 0;
	float minDist = FLT_BIG;
	int index1 = cand[0], ci1 = 0;
	for (int i = 0; i < nc; i++)
	{
		v = m[cand[i]].lb - m[ch[0]].lb;
		const float d = dot(v, nrm);
		if (d > maxDist) { maxDist = d; index = cand[i]; ci = i; }
		if (minDist > d) { minDist = d; index1 = cand[i]; ci1 = i; }
	}
	ch[2] = index;
	nc--; cand[ci] = cand[nc];
	if (nc == ci1)
		ci1 = ci;
	if (minDist * maxDist > 0.0f)
	{
		maxDist = -FLT_BIG;
		for (int i = 0; i < nc; i++)
		{
			v = m[cand[i]].lb - m[ch[0]].lb;
			const float d = dot(v, nrm);
			if (d > maxDist) { maxDist = d; index1 = cand[i]; ci1 = i; }
		}
	}
	ch[3] = index1;
	nc--; cand[ci1] = cand[nc];
	const float eps = tol * 0.02f;
	if (eps > maxPen && minPen > eps)
	{
		for (int i = 0; i < 4; i++)
		{
			float pen = m[ch[i]].pen;
			if (pen > eps)
			{
				int cidx = -1;
				for (int j = 0; j < nc; j++)
				{
					const float pen1 = m[cand[j]].pen;
					if (pen > pen1 && eps > pen1) { pen = pen1; cidx = j; }
				}
				if (cidx >= 0)
				{
					const int orig = ch[i];
					ch[i] = cand[cidx];
					cand[cidx] = orig;
				}
			}
		}
	}
}

PXH_D void reduceBatch(const PC* m, int n, float tol, int* ch) { reduceBatchT<16>(m, n, tol, ch); }

PXH_D void segClosest(f3 p1, f3 q1, f3 p2, f3 q2, f3& c1, f3& c2)
{
	// closest points of two segments (Ericson, Real-Time Collision Detection 5.1.9)
	const f3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
	const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
	const float c = dot(d1, r), b = dot(d1, d2), den = a * e - b * b;
	float s = den > 1e-12f ? fminf(fmaxf((b * f - c * e) / den, 0.0f), 1.0f) : 0.0f;
	float t = (b * s + f) / e;
	if (t < 0.0f) { t = 0.0f; s = fminf(fmaxf(-c / a, 0.0f), 1.0f); }
	else if (t > 1.0f) { t = 1.0f; s = fminf(fmaxf((b - c) / a, 0.0f), 1.0f); }
	c1 = p1 + d1 * s;
	c2 = p2 + d2 * t;
}

// The reference's GJK/EPA fallback for box pairs that face clipping leaves
// without a point (GuPCMContactBoxBox.cpp:922-958), restated for boxes as the
// minimum-translation axis over the 15 SAT axes with the witness on box1 at the
// closest points of the supporting edges.
PXH_D int boxBoxEdge(f3 e0, const mxf& t0, f3 e1, const mxf& t1, float cd, Contacts& o)
{
	const f3 T = t0.p - t1.p;
	float best = -FLT_BIG;
	f3 bestN = mk(0.0f, 1.0f, 0.0f);
	int bi = -1, bj = -1;
	for (int k = 0; k < 15; k++)
	{
		f3 L;
		int i = -1, j = -1;
		if (k < 3) L = k == 0 ? t0.r.c0 : (k == 1 ? t0.r.c1 : t0.r.c2);
		else if (k < 6) L = k == 3 ? t1.r.c0 : (k == 4 ? t1.r.c1 : t1.r.c2);
		else
		{
			i = (k - 6) / 3; j = (k - 6) % 3;
			const f3 a = i == 0 ? t0.r.c0 : (i == 1 ? t0.r.c1 : t0.r.c2);
			const f3 b = j == 0 ? t1.r.c0 : (j == 1 ? t1.r.c1 : t1.r.c2);
			L = cross(a, b);
		}
		const float l2 = dot(L, L);
		if (l2 < 1e-10f)
			continue;
		L = L * (1.0f / sqrtf(l2));
		const float ra = e0.x * fabsf(dot(t0.r.c0, L)) + e0.y * fabsf(dot(t0.r.c1, L)) + e0.z * fabsf(dot(t0.r.c2, L));
		const float rb = e1.x * fabsf(dot(t1.r.c0, L)) + e1.y * fabsf(dot(t1.r.c1, L)) + e1.z * fabsf(dot(t1.r.c2, L));
		const float s = dot(T, L);
		const float sep = fabsf(s) - (ra + rb);
		if (sep > best)
		{
			best = sep;
			bestN = s < 0.0f ? -L : L;
			bi = i;
			bj = j;
		}
	}
	if (best > cd)
		return 0;
	if (best > 0.0f)
	{
		// separated: GJK's closest points, here by exhaustive feature pairs
		// (vertex against box, edge against edge)
		float dmin = FLT_BIG;
		f3 qa = t0.p, qb = t1.p;
		for (int side = 0; side < 2; side++)
		{
			const mxf& tv = side ? t1 : t0;
			const mxf& tb = side ? t0 : t1;
			const f3 ev = side ? e1 : e0, eb = side ? e0 : e1;
			for (int k = 0; k < 8; k++)
			{
				const f3 lv = mk((k & 4) ? -ev.x : ev.x, (k & 2) ? -ev.y : ev.y, (k & 1) ? -ev.z : ev.z);
				const f3 wv = mxtrans(tv, lv);
				const f3 lb = mtmul(tb.r, wv - tb.p);
				const f3 cl = mk(fminf(fmaxf(lb.x, -eb.x), eb.x), fminf(fmaxf(lb.y, -eb.y), eb.y),
				                 fminf(fmaxf(lb.z, -eb.z), eb.z));
				const f3 wq = mxtrans(tb, cl);
				const f3 dv = wv - wq;
				const float dd = dot(dv, dv);
				if (dd < dmin) { dmin = dd; qa = side ? wq : wv; qb = side ? wv : wq; }
			}
		}
		for (int ia = 0; ia < 12; ia++)
		{
			f3 s0a, s0b;
			{
				const int ax = ia / 4;
				const float ev[3] = {e0.x, e0.y, e0.z};
				float cc[3] = {0, 0, 0}, dd[3] = {0, 0, 0};
				cc[(ax + 1) % 3] = ((ia & 1) ? -1.0f : 1.0f) * ev[(ax + 1) % 3];
				cc[(ax + 2) % 3] = ((ia & 2) ? -1.0f : 1.0f) * ev[(ax + 2) % 3];
				dd[ax] = ev[ax];
				s0a = mxtrans(t0, mk(cc[0] - dd[0], cc[1] - dd[1], cc[2] - dd[2]));
				s0b = mxtrans(t0, mk(cc[0] + dd[0], cc[1] + dd[1], cc[2] + dd[2]));
			}
			for (int ib = 0; ib < 12; ib++)
			{
				const int ax = ib / 4;
				const float ev[3] = {e1.x, e1.y, e1.z};
				float cc[3] = {0, 0, 0}, dd[3] = {0, 0, 0};
				cc[(ax + 1) % 3] = ((ib & 1) ? -1.0f : 1.0f) * ev[(ax + 1) % 3];
				cc[(ax + 2) % 3] = ((ib & 2) ? -1.0f : 1.0f) * ev[(ax + 2) % 3];
				dd[ax] = ev[ax];
				const f3 s1a = mxtrans(t1, mk(cc[0] - dd[0], cc[1] - dd[1], cc[2] - dd[2]));
				const f3 s1b = mxtrans(t1, mk(cc[0] + dd[0], cc[1] + dd[1], cc[2] + dd[2]));
				f3 c0, c1;
				segClosest(s0a, s0b, s1a, s1b, c0, c1);
				const f3 dv = c0 - c1;
				const float dd2 = dot(dv, dv);
				if (dd2 < dmin) { dmin = dd2; qa = c0; qb = c1; }
			}
		}
		const float dist = sqrtf(dmin);
		if (dist > cd)
			return 0;
		const f3 nrm = dist > 0.0f ? (qa - qb) * (1.0f / dist) : bestN;
		const f3 lb = mtmul(t1.r, qb - t1.p);
		const f3 ln = mtmul(t1.r, nrm);
		o.n = normalize(mmul(t1.r, ln));
		o.count = 1;
		o.p[0] = mxtrans(t1, lb);
		o.sep[0] = dist;
		return 1;
	}
	const f3 a[3] = {t0.r.c0, t0.r.c1, t0.r.c2}, b[3] = {t1.r.c0, t1.r.c1, t1.r.c2};
	const float ea[3] = {e0.x, e0.y, e0.z}, eb[3] = {e1.x, e1.y, e1.z};
	f3 pB;
	if (bi >= 0)
	{
		f3 c0 = t0.p, c1 = t1.p;
		for (int k = 0; k < 3; k++)
		{
			if (k != bi) c0 = c0 - a[k] * (dot(a[k], bestN) > 0.0f ? ea[k] : -ea[k]);
			if (k != bj) c1 = c1 + b[k] * (dot(b[k], bestN) > 0.0f ? eb[k] : -eb[k]);
		}
		f3 q0, q1;
		segClosest(c0 - a[bi] * ea[bi], c0 + a[bi] * ea[bi], c1 - b[bj] * eb[bj], c1 + b[bj] * eb[bj], q0, q1);
		pB = q1;
	}
	else
	{
		pB = t1.p;
		for (int k = 0; k < 3; k++)
			pB = pB + b[k] * (dot(b[k], bestN) > 0.0f ? eb[k] : -eb[k]);
	}
	// the reference stores and re-expresses the point/normal in box1's frame
	const f3 lb = mtmul(t1.r, pB - t1.p);
	const f3 ln = mtmul(t1.r, bestN);
	o.n = normalize(mmul(t1.r, ln));
	o.count = 1;
	o.p[0] = mxtrans(t1, lb);
	o.sep[0] = best;
	return 1;
}

// doBoxBoxGenerateContacts + pcmContactBoxBox on a fresh manifold (:332-968)
PXH_D int boxBox(f3 e0, const xf& x0, f3 e1, const xf& x1, float cd, float tol, Contacts& o)
{
	mxf t0, t1;
	t0.r = qmat(x0.q); t0.p = x0.p;
	t1.r = qmat(x1.q); t1.p = x1.p;
	t0.r.c0 = normalize(t0.r.c0); t0.r.c1 = normalize(t0.r.c1); t0.r.c2 = normalize(t0.r.c2);
	t1.r.c0 = normalize(t1.r.c0); t1.r.c1 = normalize(t1.r.c1); t1.r.c2 = normalize(t1.r.c2);

	const mxf t1To0 = mxinvmul(t0, t1);
	const m33 r0To1 = mtrans(t1To0.r);
	const f3 ue = mk(1e-6f, 1e-6f, 1e-6f);
	const float ea0 = e0.x, ea1 = e0.y, ea2 = e0.z, eb0 = e1.x, eb1 = e1.y, eb2 = e1.z;
	const float tx = t1To0.p.x, ty = t1To0.p.y, tz = t1To0.p.z;
	const f3 col0 = t1To0.r.c0, col1 = t1To0.r.c1, col2 = t1To0.r.c2;
	const f3 a10c0 = vabs(col0) + ue, a10c1 = vabs(col1) + ue, a10c2 = vabs(col2) + ue;
	const f3 a01c0 = vabs(r0To1.c0) + ue, a01c1 = vabs(r0To1.c1) + ue, a01c2 = vabs(r0To1.c2) + ue;
	float sign[6], overlap[6];
	f3 vt;
	vt = mulv(a01c0, e1); sign[0] = tx; overlap[0] = ((ea0 + (vt.x + (vt.y + vt.z))) - fabsf(tx)) + cd;
	if (0.0f > overlap[0]) return 0;
	vt = mulv(a01c1, e1); sign[1] = ty; overlap[1] = ((ea1 + (vt.x + (vt.y + vt.z))) - fabsf(ty)) + cd;
	if (0.0f > overlap[1]) return 0;
	vt = mulv(a01c2, e1); sign[2] = tz; overlap[2] = ((ea2 + (vt.x + (vt.y + vt.z))) - fabsf(tz)) + cd;
	if (0.0f > overlap[2]) return 0;
	vt = mulv(a10c0, e0); sign[3] = dot(t1To0.p, col0); overlap[3] = (((vt.x + (vt.y + vt.z)) + eb0) - fabsf(sign[3])) + cd;
	if (0.0f > overlap[3]) return 0;
	vt = mulv(a10c1, e0); sign[4] = dot(t1To0.p, col1); overlap[4] = (((vt.x + (vt.y + vt.z)) + eb1) - fabsf(sign[4])) + cd;
	if (0.0f > overlap[4]) return 0;
	vt = mulv(a10c2, e0); sign[5] = dot(t1To0.p, col2); overlap[5] = (((vt.x + (vt.y + vt.z)) + eb2) - fabsf(sign[5])) + cd;
	if (0.0f > overlap[5]) return 0;
#define PXH_EDGE(absSign, raA, raB, rbA, rbB) \
	if ((absSign) > ((((raA) + (raB)) + ((rbA) + (rbB))) + cd)) return 0;
	PXH_EDGE(fabsf(col0.y * tz - col0.z * ty), a10c0.z * ea1, a10c0.y * ea2, a01c0.z * eb1, a01c0.y * eb2);
	PXH_EDGE(fabsf(col1.y * tz - col1.z * ty), a10c1.z * ea1, a10c1.y * ea2, a01c0.z * eb0, a01c0.x * eb2);
	PXH_EDGE(fabsf(col2.y * tz - col2.z * ty), a10c2.z * ea1, a10c2.y * ea2, a01c0.y * eb0, a01c0.x * eb1);
	PXH_EDGE(fabsf(col0.z * tx - col0.x * tz), a10c0.z * ea0, a10c0.x * ea2, a01c1.z * eb1, a01c1.y * eb2);
	PXH_EDGE(fabsf(col1.z * tx - col1.x * tz), a10c1.z * ea0, a10c1.x * ea2, a01c1.z * eb0, a01c1.x * eb2);
	PXH_EDGE(fabsf(col2.z * tx - col2.x * tz), a10c2.z * ea0, a10c2.x * ea2, a01c1.y * eb0, a01c1.x * eb1);
	PXH_EDGE(fabsf(col0.x * ty - col0.y * tx), a10c0.y * ea0, a10c0.x * ea1, a01c2.z * eb1, a01c2.y * eb2);
	PXH_EDGE(fabsf(col1.x * ty - col1.y * tx), a10c1.y * ea0, a10c1.x * ea1, a01c2.z * eb0, a01c2.x * eb2);
	PXH_EDGE(fabsf(col2.x * ty - col2.y * tx), a10c2.y * ea0, a10c2.x * ea1, a01c2.y * eb0, a01c2.x * eb1);
#undef PXH_EDGE

	int feature = 0;
	float minOverlap = overlap[0];
#pragma unroll
	for (int i = 1; i < 6; i++)
		if (minOverlap > overlap[i]) { minOverlap = overlap[i]; feature = i; }

	// reference face frame: z along the contact normal into the reference box
	const bool flip = feature >= 3;
	const mxf& tr = flip ? t1 : t0;          // reference box
	const int ax = flip ? feature - 3 : feature;
	const f3 u0 = tr.r.c0, u1 = tr.r.c1, u2 = tr.r.c2;
	const float er = flip ? (ax == 0 ? eb0 : (ax == 1 ? eb1 : eb2)) : (ax == 0 ? ea0 : (ax == 1 ? ea1 : ea2));
	const bool le0 = 0.0f >= sign[feature];
	mxf nt;
	f3 mtd;
	if (!flip)
	{
		if (ax == 0)
		{
			if (le0) { mtd = u0; nt.r.c0 = -u2; nt.r.c1 = u1; nt.r.c2 = u0; nt.p = tr.p - u0 * er; }
			else { mtd = -u0; nt.r.c0 = u2; nt.r.c1 = u1; nt.r.c2 = -u0; nt.p = u0 * er + tr.p; }
		}
		else if (ax == 1)
		{
			if (le0) { mtd = u1; nt.r.c0 = u0; nt.r.c1 = -u2; nt.r.c2 = u1; nt.p = tr.p - u1 * er; }
			else { mtd = -u1; nt.r.c0 = u0; nt.r.c1 = u2; nt.r.c2 = -u1; nt.p = u1 * er + tr.p; }
		}
		else
		{
			if (le0) { mtd = u2; nt.r.c0 = u0; nt.r.c1 = u1; nt.r.c2 = u2; nt.p = tr.p - u2 * er; }
			else { mtd = -u2; nt.r.c0 = u0; nt.r.c1 = -u1; nt.r.c2 = -u2; nt.p = u2 * er + tr.p; }
		}
	}
	else
	{
		if (ax == 0)
		{
			if (le0) { mtd = u0; nt.r.c0 = u2; nt.r.c1 = u1; nt.r.c2 = -u0; nt.p = u0 * er + tr.p; }
			else { mtd = -u0; nt.r.c0 = -u2; nt.r.c1 = u1; nt.r.c2 = u0; nt.p = tr.p - u0 * er; }
		}
		else if (ax == 1)
		{
			if (le0) { mtd = u1; nt.r.c0 = u0; nt.r.c1 = u2; nt.r.c2 = -u1; nt.p = u1 * er + tr.p; }
			else { mtd = -u1; nt.r.c0 = u0; nt.r.c1 = -u2; nt.r.c2 = u1; nt.p = tr.p - u1 * er; }
		}
		else
		{
			if (le0) { mtd = u2; nt.r.c0 = u0; nt.r.c1 = -u1; nt.r.c2 = -u2; nt.p = u2 * er + tr.p; }
			else { mtd = -u2; nt.r.c0 = u0; nt.r.c1 = u1; nt.r.c2 = u2; nt.p = tr.p - u2 * er; }
		}
	}
	const f3 ln = mtmul(nt.r, mtd);
	f3 pts[4], incN;
	PC mc[16];
	int nc = 0;
	// in-face extents of the reference face, in the order of the reference's cases
	float fx, fy;
	if (ax == 0) { fx = flip ? eb2 : ea2; fy = flip ? eb1 : ea1; }
	else if (ax == 1) { fx = flip ? eb0 : ea0; fy = flip ? eb2 : ea2; }
	else { fx = flip ? eb0 : ea0; fy = flip ? eb1 : ea1; }
	if (!flip)
		incidentPolygon(pts, incN, -ln, mxinvmul(nt, t1), e1);
	else
		incidentPolygon(pts, incN, ln, mxinvmul(nt, t0), e0);
	calcContacts(fx, fy, pts, incN, ln, mc, nc, cd);

	if (nc == 0)
		return boxBoxEdge(e0, t0, e1, t1, cd, o);

	const mxf newTo1 = mxinvmul(t1, nt);
	const f3 lnB = mmul(newTo1.r, ln);
	// the reference swaps the manifold's A/B points for a box1 reference face;
	// the output only needs the box1-side point (lb before the swap is on the
	// incident box, so after a swap the box1 point is la)
	int idx[4] = {0, 1, 2, 3};
	for (int i = 0; i < nc; i++)
	{
		if (flip) { mc[i].lb = mxtrans(newTo1, mc[i].la); }
		else { mc[i].lb = mxtrans(newTo1, mc[i].lb); }
		mc[i].ln = lnB;
	}
	if (nc > 4)
	{
		reduceBatch(mc, nc, tol, idx);
		nc = 4;
	}
	o.n = normalize(mmul(t1.r, lnB));
	o.count = nc;
	for (int i = 0; i < nc; i++)
	{
		o.p[i] = mxtrans(t1, mc[idx[i]].lb);
		o.sep[i] = mc[idx[i]].pen;
	}
	return nc;
}

}  // namespace pxh
