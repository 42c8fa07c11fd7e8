// Contacts of one body against a static triangle mesh, one lane per body.
//
// A restatement of the reference's PCM mesh contact generation
// (geomutils/src/pcm/GuPCMContactSphereMesh.cpp:160-501,
// GuPCMTriangleContactGen.cpp, GuPCMContactConvexCommon.h:131-230,
// GuPersistentContactManifold.cpp:1315-1420, 1678-2160) in the shape this
// build gives it:
//  * candidates: the triangles whose bounds meet the body's bounds inflated
//    by the contact distance, found through the mesh's bounding-volume tree
//    and taken in triangle order (at most MESH_MAX_TRIS);
//  * a triangle whose plane has the body's centre behind it is culled (the
//    mesh is one-sided);
//  * a sphere takes the closest point of each triangle: face contacts at
//    once, edge and vertex contacts deferred and taken by distance unless a
//    face or a closer contact already holds that edge or vertex (the
//    reference's edge and vertex caches);
//  * a capsule takes the part of its segment over the triangle (its prism),
//    else the closest boundary point, deferred like a sphere's;
//  * a box or hull takes its face most opposed to the triangle normal,
//    clipped to the triangle's prism, and its deepest vertex when over it;
//  * contacts group into patches by normal (within 0.996, the reference's
//    acceptance); a patch holds MESH_PC contacts before it is reduced to 4
//    (the deepest, the farthest from it, then the extremes across their line);
//    the deepest MESH_MAXP patches become one constraint each (a sphere keeps
//    one contact per patch).
// The C oracle (oracle/pxo.c, tm_*) is the same algorithm in the same
// arithmetic order.
#pragma once
#include "pxh_convex.h"
#include "pxh_kernels.h"

namespace pxh {

constexpr int MESH_PC = 8;
constexpr int MESH_NPATCH = 6;

struct MC { f3 p; float s; };
struct MPatch { f3 n; int count; float maxPen; MC c[MESH_PC + 1]; };
struct MPatches { MPatch p[MESH_NPATCH]; int np; };

PXH_D int meshReduce4(MC* c, int n, f3 nrm)
{
	if (n <= 4)
		return n;
	bool used[MESH_PC + 1];
	for (int i = 0; i < n; i++)
		used[i] = false;
	int i0 = 0;
	for (int i = 1; i < n; i++)
		if (c[i0].s > c[i].s)
			i0 = i;
	used[i0] = true;
	int i1 = -1;
	float best = -1.0f;
	for (int i = 0; i < n; i++)
	{
		if (used[i]) continue;
		const f3 d = c[i].p - c[i0].p;
		const float dd = dot(d, d);
		if (dd > best) { best = dd; i1 = i; }
	}
	used[i1] = true;
	const f3 ax = cross(c[i1].p - c[i0].p, nrm);
	int i2 = -1, i3 = -1;
	float mx = -3.402823466e+38f, mn = 3.402823466e+38f;
	for (int i = 0; i < n; i++)
	{
		if (used[i]) continue;
		const float d = dot(c[i].p - c[i0].p, ax);
		if (d > mx) { mx = d; i2 = i; }
		if (mn > d) { mn = d; i3 = i; }
	}
	used[i2] = true;
	if (i3 == i2 || mx * mn > 0.0f)
	{
		mx = -3.402823466e+38f;
		i3 = -1;
		for (int i = 0; i < n; i++)
		{
			if (used[i]) continue;
			const float d = dot(c[i].p - c[i0].p, ax);
			if (d > mx) { mx = d; i3 = i; }
		}
	}
	const MC a = c[i0], b = c[i1], e = c[i2], f = c[i3];
	c[0] = a; c[1] = b; c[2] = e; c[3] = f;
	return 4;
}

PXH_D void meshAdd(MPatches& P, f3 n, f3 p, float s)
{
	for (int k = 0; k < P.np; k++)
	{
		MPatch& q = P.p[k];
		if (!(dot(q.n, n) > 0.996f))
			continue;
		for (int j = 0; j < q.count; j++)
		{
			const f3 d = q.c[j].p - p;
			if (1e-10f > dot(d, d))
			{
				if (q.c[j].s > s)
				{
					q.c[j].p = p;
					q.c[j].s = s;
				}
				if (q.maxPen > s)
					q.maxPen = s;
				return;
			}
		}
		q.c[q.count].p = p;
		q.c[q.count].s = s;
		q.count++;
		if (q.maxPen > s)
			q.maxPen = s;
		if (q.count > MESH_PC)
			q.count = meshReduce4(q.c, q.count, q.n);
		return;
	}
	int k = P.np;
	if (k == MESH_NPATCH)
	{
		k = 0;
		for (int j = 1; j < MESH_NPATCH; j++)
			if (P.p[j].maxPen > P.p[k].maxPen)
				k = j;
		if (!(P.p[k].maxPen > s))
			return;
	}
	else
		P.np++;
	P.p[k].n = n;
	P.p[k].count = 1;
	P.p[k].maxPen = s;
	P.p[k].c[0].p = p;
	P.p[k].c[0].s = s;
}

// pcmDistancePointTriangleSquared (GuPCMContactSphereMesh.cpp:160-288)
enum { TF_FACE = 0, TF_V0 = 1, TF_V1 = 2, TF_V2 = 3, TF_E01 = 4, TF_E12 = 5, TF_E20 = 6 };
PXH_D float meshPointTri(f3 p, f3 a, f3 b, f3 c, f3& cp, int& fc)
{
	const f3 ab = b - a, ac = c - a, bc = c - b;
	const f3 ap = p - a, bp = p - b, cq = p - c;
	const float d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cq),
	            d6 = dot(ac, cq);
	const float unom = d4 - d3, udenom = d5 - d6;
	const f3 n = cross(ab, ac);
	if (0.0f > d1 && 0.0f > d2) { fc = TF_V0; cp = a; return dot(ap, ap); }
	if (d3 >= 0.0f && d3 >= d4) { fc = TF_V1; cp = b; return dot(bp, bp); }
	if (d6 >= 0.0f && d6 >= d5) { fc = TF_V2; cp = c; return dot(cq, cq); }
	const float vc = dot(n, cross(ap, bp));
	if (0.0f > vc && d1 >= 0.0f && 0.0f > d3)
	{
		fc = TF_E01;
		const f3 q = ab * (d1 / (d1 - d3)) + a;
		cp = q;
		const f3 vv = p - q;
		return dot(vv, vv);
	}
	const float va = dot(n, cross(bp, cq));
	if (0.0f > va && d4 >= d3 && d5 >= d6)
	{
		fc = TF_E12;
		const f3 q = bc * (unom / (unom + udenom)) + b;
		cp = q;
		const f3 vv = p - q;
		return dot(vv, vv);
	}
	const float vb = dot(n, cross(cq, ap));
	if (0.0f > vb && d2 >= 0.0f && 0.0f > d6)
	{
		fc = TF_E20;
		const f3 q = ac * (d2 / (d2 - d6)) + a;
		cp = q;
		const f3 vv = p - q;
		return dot(vv, vv);
	}
	fc = TF_FACE;
	const float t = dot(n, a - p) / dot(n, n);
	const f3 vv = n * t;
	cp = p + vv;
	return dot(vv, vv);
}

// edge and vertex caches (CachedEdge / CachedVertex,
// geomutils/src/mesh/GuTriangleCache.h:38-90): add returns true when new
struct MeshCache
{
	uint64_t e[3 * MESH_MAX_TRIS];
	int v[3 * MESH_MAX_TRIS];
	int ne, nv;
	PXH_D bool edge(int a, int b)
	{
		const uint64_t k = (uint64_t(uint32_t(a < b ? a : b)) << 32) | uint32_t(a < b ? b : a);
		for (int i = 0; i < ne; i++)
			if (e[i] == k) return false;
		e[ne++] = k;
		return true;
	}
	PXH_D bool vertex(int a)
	{
		for (int i = 0; i < nv; i++)
			if (v[i] == a) return false;
		v[nv++] = a;
		return true;
	}
};

struct MeshTri
{
	f3 v[3], n;
	int vi[3];
};

PXH_D MeshTri meshTri(const MeshSet& M, int t)
{
	MeshTri r;
	const float4 a = M.v0[t], b = M.v1[t], c = M.v2[t];
	r.v[0] = mk(a); r.v[1] = mk(b); r.v[2] = mk(c);
	r.vi[0] = __float_as_int(a.w); r.vi[1] = __float_as_int(b.w); r.vi[2] = __float_as_int(c.w);
	r.n = mk(M.n[t]);
	return r;
}

PXH_D void meshCacheTri(MeshCache& C, const MeshTri& t)
{
	C.edge(t.vi[0], t.vi[1]);
	C.edge(t.vi[1], t.vi[2]);
	C.edge(t.vi[2], t.vi[0]);
	C.vertex(t.vi[0]);
	C.vertex(t.vi[1]);
	C.vertex(t.vi[2]);
}

PXH_D bool meshCulled(const MeshTri& t, f3 c) { return 0.0f > dot(c, t.n) - dot(t.v[0], t.n); }

struct MeshDef { f3 p, n; float sq; int fc, t; };

// deferred edge and vertex contacts by distance, each unless its feature is
// held (PCMSphereVsMeshContactGeneration::generateLastContacts); reported on
// the sphere or capsule: p - n r
PXH_D void meshFlushDeferred(const MeshSet& M, MeshDef* d, int nd, MeshCache& C, float r, MPatches& P)
{
	for (int i = 1; i < nd; i++)
	{
		const MeshDef x = d[i];
		int j = i - 1;
		while (j >= 0 && d[j].sq > x.sq)
		{
			d[j + 1] = d[j];
			j--;
		}
		d[j + 1] = x;
	}
	for (int i = 0; i < nd; i++)
	{
		const MeshTri t = meshTri(M, d[i].t);
		const bool e01 = C.edge(t.vi[0], t.vi[1]);
		const bool e12 = C.edge(t.vi[1], t.vi[2]);
		const bool e20 = C.edge(t.vi[2], t.vi[0]);
		const bool v0 = C.vertex(t.vi[0]);
		const bool v1 = C.vertex(t.vi[1]);
		const bool v2 = C.vertex(t.vi[2]);
		const int fc = d[i].fc;
		const bool need = fc == TF_V0 ? v0 : fc == TF_V1 ? v1 : fc == TF_V2 ? v2 : fc == TF_E01 ? e01 : fc == TF_E12 ? e12 : e20;
		if (need)
		{
			const float dist = sqrtf(d[i].sq);
			meshAdd(P, d[i].n, d[i].p - d[i].n * r, dist - r);
		}
	}
}

PXH_D void meshSphere(const MeshSet& M, const int* idx, int nc, f3 c, float r, float cd, MeshCache& C, MeshDef* def,
                      MPatches& P)
{
	const float infl = r + cd;
	const float inflSq = infl * infl;
	C.ne = C.nv = 0;
	int nd = 0;
	for (int k = 0; k < nc; k++)
	{
		const MeshTri t = meshTri(M, idx[k]);
		if (meshCulled(t, c))
			continue;
		f3 cp;
		int fc;
		const float sq = meshPointTri(c, t.v[0], t.v[1], t.v[2], cp, fc);
		if (!(inflSq > sq))
			continue;
		if (fc == TF_FACE)
		{
			meshCacheTri(C, t);
			const float dist = sqrtf(sq);
			meshAdd(P, t.n, c - t.n * r, dist - r);
		}
		else
		{
			def[nd].p = c;
			def[nd].n = normalize(c - cp);
			def[nd].sq = sq;
			def[nd].fc = fc;
			def[nd].t = idx[k];
			nd++;
		}
	}
	meshFlushDeferred(M, def, nd, C, r, P);
}

// inward normals of the triangle's edge planes
PXH_D void meshPrism(const MeshTri& t, f3* m)
{
	for (int k = 0; k < 3; k++)
		m[k] = cross(t.n, t.v[(k + 1) % 3] - t.v[k]);
}

// closest points of segments p1q1, p2q2 (Ericson, Real-Time Collision
// Detection, 5.1.9), as the oracle's seg_closest2
PXH_D float meshSegClosest(f3 p1, f3 q1, f3 p2, f3 q2, f3& c1, f3& c2)
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

PXH_D void meshCapsule(const MeshSet& M, const int* idx, int nc, f3 p0, f3 p1, f3 c, float r, float cd, MeshCache& C,
                       MeshDef* def, MPatches& P)
{
	const float infl = r + cd;
	C.ne = C.nv = 0;
	int nd = 0;
	for (int k = 0; k < nc; k++)
	{
		const MeshTri t = meshTri(M, idx[k]);
		if (meshCulled(t, c))
			continue;
		f3 m[3];
		meshPrism(t, m);
		float t0 = 0.0f, t1 = 1.0f;
		bool empty = false;
		for (int e = 0; e < 3 && !empty; e++)
		{
			const float a = dot(m[e], p0 - t.v[e]), b = dot(m[e], p1 - t.v[e]);
			if (0.0f > a && 0.0f > b)
				empty = true;
			else if (0.0f > a)
			{
				const float u = a / (a - b);
				if (u > t0) t0 = u;
			}
			else if (0.0f > b)
			{
				const float u = a / (a - b);
				if (t1 > u) t1 = u;
			}
		}
		if (!empty && t1 >= t0)
		{
			const f3 d = p1 - p0;
			f3 q[2];
			q[0] = p0 + d * t0;
			q[1] = p0 + d * t1;
			const int nq = t1 > t0 ? 2 : 1;
			bool any = false;
			for (int j = 0; j < nq; j++)
			{
				const float s = dot(q[j] - t.v[0], t.n);
				if (infl > s)
				{
					meshAdd(P, t.n, q[j] - t.n * r, s - r);
					any = true;
				}
			}
			if (any)
				meshCacheTri(C, t);
			continue;
		}
		float best = 3.402823466e+38f;
		f3 bs = p0, bt = p0;
		int bf = TF_E01;
		for (int e = 0; e < 3; e++)
		{
			f3 cs, ct;
			const float sq = meshSegClosest(p0, p1, t.v[e], t.v[(e + 1) % 3], cs, ct);
			if (best > sq)
			{
				best = sq;
				bs = cs;
				bt = ct;
				bf = e == 0 ? TF_E01 : e == 1 ? TF_E12 : TF_E20;
			}
		}
		if (!(infl * infl > best) || !(best > 0.0f))
			continue;
		const f3 nn = normalize(bs - bt);
		if (!(dot(nn, t.n) > 0.0f))
			continue;
		def[nd].p = bs;
		def[nd].n = nn;
		def[nd].sq = best;
		def[nd].fc = bf;
		def[nd].t = idx[k];
		nd++;
	}
	meshFlushDeferred(M, def, nd, C, r, P);
}

PXH_D void meshPoly(const MeshSet& M, const int* idx, int nc, const Poly& H, const xf& pose, float cd, MPatches& P)
{
	const m33 R = qmat(pose.q);
	for (int k = 0; k < nc; k++)
	{
		const MeshTri t = meshTri(M, idx[k]);
		if (meshCulled(t, pose.p))
			continue;
		int f = 0;
		float best = 3.402823466e+38f;
		for (int j = 0; j < H.np; j++)
		{
			const float d = dot(mmul(R, H.pn(j)), t.n);
			if (best > d) { best = d; f = j; }
		}
		f3 buf0[64 + 3], buf1[64 + 3];
		int n0 = H.pcount(f) > 64 ? 64 : H.pcount(f);
		for (int j = 0; j < n0; j++)
			buf0[j] = mmul(R, H.vert(H.pvert(f, j))) + pose.p;
		f3 m[3];
		meshPrism(t, m);
		f3* in = buf0;
		f3* out = buf1;
		for (int e = 0; e < 3 && n0 > 0; e++)
		{
			int n1 = 0;
			for (int j = 0; j < n0; j++)
			{
				const f3 a = in[j], b = in[(j + 1) % n0];
				const float da = dot(m[e], a - t.v[e]), db = dot(m[e], b - t.v[e]);
				if (da >= 0.0f)
					out[n1++] = a;
				if ((da >= 0.0f) != (db >= 0.0f))
					out[n1++] = a + (b - a) * (da / (da - db));
			}
			f3* tmp = in;
			in = out;
			out = tmp;
			n0 = n1;
		}
		for (int j = 0; j < n0; j++)
		{
			const float s = dot(in[j] - t.v[0], t.n);
			if (cd > s)
				meshAdd(P, t.n, in[j] - t.n * s, s);
		}
		int vbest = 0;
		float dbest = 3.402823466e+38f;
		for (int j = 0; j < H.nv; j++)
		{
			const float d = dot(mmul(R, H.vert(j)), t.n);
			if (dbest > d) { dbest = d; vbest = j; }
		}
		const f3 sv = mmul(R, H.vert(vbest)) + pose.p;
		if (dot(m[0], sv - t.v[0]) >= 0.0f && dot(m[1], sv - t.v[1]) >= 0.0f && dot(m[2], sv - t.v[2]) >= 0.0f)
		{
			const float s = dot(sv - t.v[0], t.n);
			if (cd > s)
				meshAdd(P, t.n, sv - t.n * s, s);
		}
	}
}

}  // namespace pxh
