/* CPU oracle: plain-C restatement of the reference PhysX 5.6.1 CPU rigid-body
 * step. TEST INFRASTRUCTURE ONLY (see pxo.h): loaded by tests/, smoke() and
 * bench.py's CPU-baseline leg as the checker, never by the product.
 *
 * Every function cites the reference file:line it restates. Paths are
 * relative to physx/source/ unless they start with physx/include.
 * Compiled with -ffp-contract=off so float arithmetic follows the source order.
 */
#include "pxo.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float x, y, z; } v3;
typedef struct { float x, y, z, w; } qt;
typedef struct { v3 c0, c1, c2; } m33;      /* columns */
typedef struct { qt q; v3 p; } xf;          /* PxTransform */
typedef struct { m33 r; v3 p; } mxf;        /* PxMatTransformV */

static inline v3 V(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
static inline v3 neg(v3 a) { return V(-a.x, -a.y, -a.z); }
static inline v3 mulv(v3 a, v3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) { return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static inline float len2(v3 a) { return dot(a, a); }
static inline v3 vabs(v3 a) { return V(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
static inline v3 normalize(v3 a) { return scl(a, 1.0f / sqrtf(dot(a, a))); }
static inline float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

/* PxQuat::rotate / rotateInv / operator* (physx/include/foundation/PxQuat.h) */
static inline v3 qrot(qt q, v3 v)
{
	const float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
	const float w2 = q.w * q.w - 0.5f;
	const float dot2 = (q.x * vx + q.y * vy + q.z * vz);
	return V((vx * w2 + (q.y * vz - q.z * vy) * q.w + q.x * dot2), (vy * w2 + (q.z * vx - q.x * vz) * q.w + q.y * dot2),
	         (vz * w2 + (q.x * vy - q.y * vx) * q.w + q.z * dot2));
}
static inline v3 qrotinv(qt q, v3 v)
{
	const float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
	const float w2 = q.w * q.w - 0.5f;
	const float dot2 = (q.x * vx + q.y * vy + q.z * vz);
	return V((vx * w2 - (q.y * vz - q.z * vy) * q.w + q.x * dot2), (vy * w2 - (q.z * vx - q.x * vz) * q.w + q.y * dot2),
	         (vz * w2 - (q.x * vy - q.y * vx) * q.w + q.z * dot2));
}
static inline qt qmul(qt a, qt b)
{
	qt r = {a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z, a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x,
	        a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
	return r;
}
static inline qt qnormalize(qt q)
{
	const float m = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
	const float s = m > 0.0f ? 1.0f / m : 0.0f;
	qt r = {q.x * s, q.y * s, q.z * s, q.w * s};
	return r;
}
static inline qt qconj(qt q) { qt r = {-q.x, -q.y, -q.z, q.w}; return r; }

/* PxMat33(const PxQuat&) (physx/include/foundation/PxMat33.h) */
static m33 qmat(qt q)
{
	const float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
	const float xx = x2 * q.x, yy = y2 * q.y, zz = z2 * q.z;
	const float xy = x2 * q.y, xz = x2 * q.z, xw = x2 * q.w;
	const float yz = y2 * q.z, yw = y2 * q.w, zw = z2 * q.w;
	m33 m;
	m.c0 = V(1.0f - yy - zz, xy + zw, xz - yw);
	m.c1 = V(xy - zw, 1.0f - xx - zz, yz + xw);
	m.c2 = V(xz + yw, yz - xw, 1.0f - xx - yy);
	return m;
}
static inline v3 mmul(m33 m, v3 v) { return add(add(scl(m.c0, v.x), scl(m.c1, v.y)), scl(m.c2, v.z)); }
static inline v3 mtmul(m33 m, v3 v) { return V(dot(m.c0, v), dot(m.c1, v), dot(m.c2, v)); }
static inline m33 mtrans(m33 m)
{
	m33 t;
	t.c0 = V(m.c0.x, m.c1.x, m.c2.x);
	t.c1 = V(m.c0.y, m.c1.y, m.c2.y);
	t.c2 = V(m.c0.z, m.c1.z, m.c2.z);
	return t;
}
static inline m33 mmm(m33 a, m33 b)
{
	m33 r;
	r.c0 = mmul(a, b.c0);
	r.c1 = mmul(a, b.c1);
	r.c2 = mmul(a, b.c2);
	return r;
}

static inline v3 xtrans(xf t, v3 v) { return add(qrot(t.q, v), t.p); }
static inline v3 xtransinv(xf t, v3 v) { return qrotinv(t.q, sub(v, t.p)); }
static inline xf xinvmul(xf a, xf b) /* a^-1 * b  (PxTransform::transformInv(PxTransform)) */
{
	const qt qinv = qconj(a.q);
	xf r;
	r.q = qmul(qinv, b.q);
	r.p = qrot(qinv, sub(b.p, a.p));
	return r;
}
static inline v3 mxtrans(mxf t, v3 v) { return add(mmul(t.r, v), t.p); }
static inline mxf mxinvmul(mxf a, mxf b) /* PxMatTransformV::transformInv(PxMatTransformV) */
{
	mxf r;
	r.r = mmm(mtrans(a.r), b.r);
	r.p = mtmul(a.r, sub(b.p, a.p));
	return r;
}

/* ------------------------------------------------------------------------- */
/* Contact generation (PCM, fresh cache)                                      */
/* ------------------------------------------------------------------------- */

typedef struct { v3 la, lb; v3 ln; float pen; } pcontact; /* PersistentContact */

typedef struct {
	v3 n;           /* from shape1 towards shape0 */
	int count;
	v3 p[4];
	float sep[4];
} cset;

/* geomutils/src/pcm/GuPCMContactSphereSphere.cpp:36-67 */
static int c_sphere_sphere(float r0, v3 p0, float r1, v3 p1, float cd, cset* o)
{
	const v3 d = sub(p0, p1);
	const float dist2 = dot(d, d);
	const float rs = r0 + r1;
	const float inf = rs + cd;
	if (!(inf * inf > dist2))
		return 0;
	const float dist = sqrtf(dist2);
	const v3 n = (1e-5f >= dist) ? V(1.0f, 0.0f, 0.0f) : scl(d, 1.0f / dist);
	o->n = n;
	o->p[0] = add(scl(n, r1), p1);
	o->sep[0] = dist - rs;
	o->count = 1;
	return 1;
}

/* geomutils/src/pcm/GuPCMContactSpherePlane.cpp:36-75; plane n.x + d = 0 */
static int c_sphere_plane(float r, v3 p0, v3 pn, float pd, float cd, cset* o)
{
	const float sep = (dot(pn, p0) + pd) - r;
	if (!(cd >= sep))
		return 0;
	o->n = pn;
	o->p[0] = sub(p0, scl(pn, r));
	o->sep[0] = sep;
	o->count = 1;
	return 1;
}

/* geomutils/src/pcm/GuPCMContactSphereBox.cpp:37-130 */
static int c_sphere_box(float r, v3 p0, v3 h, xf t1, float cd, cset* o)
{
	const v3 c = xtransinv(t1, p0);
	const v3 p = V(fminf(fmaxf(c.x, -h.x), h.x), fminf(fmaxf(c.y, -h.y), h.y), fminf(fmaxf(c.z, -h.z), h.z));
	const v3 v = sub(c, p);
	const float l2 = dot(v, v);
	const float inf = r + cd;
	if (!(inf * inf > l2))
		return 0;
	if (h.x >= fabsf(c.x) && h.y >= fabsf(c.y) && h.z >= fabsf(c.z))
	{
		const v3 ap = vabs(p);
		const v3 ds = sub(h, ap);
		const int con0 = ds.x >= ds.z && ds.y >= ds.z && ds.z >= ds.z;
		const int con1 = ds.x >= ds.x && ds.y >= ds.x && ds.z >= ds.x;
		const v3 sg = V(p.x < 0.0f ? -1.0f : 1.0f, p.y < 0.0f ? -1.0f : 1.0f, p.z < 0.0f ? -1.0f : 1.0f);
		v3 ln;
		float dist;
		if (con0) { ln = V(0, 0, sg.z); dist = -ds.z; }
		else if (con1) { ln = V(sg.x, 0, 0); dist = -ds.x; }
		else { ln = V(0, sg.y, 0); dist = -ds.y; }
		const v3 n = qrot(t1.q, ln);
		o->n = n;
		o->sep[0] = dist - r;
		o->p[0] = sub(p0, scl(n, dist));
	}
	else
	{
		const float rl = 1.0f / sqrtf(l2);
		const float l = 1.0f / rl;
		const v3 n = qrot(t1.q, scl(v, rl));
		o->n = n;
		o->sep[0] = l - r;
		o->p[0] = xtrans(t1, p);
	}
	o->count = 1;
	return 1;
}

/* ------------------------------------------------------------------------- */
/* Capsules: the segment along the local x axis, endpoints p +- x*halfHeight  */
/* (geomutils/src/GuInternal.h:60-70), inflated by the radius                 */
/* ------------------------------------------------------------------------- */

/* the vector library's min/max/clamp (SSE minps/maxps: the second operand
 * when the comparison is unordered) */
static inline float vmin(float a, float b) { return a < b ? a : b; }
static inline float vmax(float a, float b) { return a > b ? a : b; }
static inline float clamp01(float x) { return vmax(vmin(x, 1.0f), 0.0f); }

static inline void capsule_segment(xf t, float hh, v3* s, v3* e)
{
	const v3 tmp = scl(qrot(t.q, V(1.0f, 0.0f, 0.0f)), hh);
	*s = add(t.p, tmp);
	*e = sub(t.p, tmp);
}

/* Gu::distanceSegmentSegmentSquared, vector form
 * (geomutils/src/distance/GuDistanceSegmentSegment.cpp:411-470): closest
 * parameters (s on p1 + s d1, t on p2 + t d2); parallel segments (a e - b^2
 * below FLT_EPSILON) take s = 1/2 before the clamped refinement */
static float seg_seg_sq(v3 p1, v3 d1, v3 p2, v3 d2, float* s, float* t)
{
	const float eps = 1.1920929e-07f;
	const v3 r = sub(p1, p2);
	const float a = dot(d1, d1), e = dot(d2, d2), b = dot(d1, d2), c = dot(d1, r);
	const float aR = a > eps ? 1.0f / a : 0.0f;
	const float eR = e > eps ? 1.0f / e : 0.0f;
	const float f = dot(d2, r);
	const float denom = a * e - b * b;
	const float temp = b * f - c * e;
	const float s0 = clamp01(temp / denom);
	const float sT = eps > denom ? 0.5f : s0;
	const float t2 = clamp01((b * sT + f) * eR);
	const float s2 = clamp01((b * t2 - c) * aR);
	*s = s2;
	*t = t2;
	const v3 vv = sub(add(scl(d1, s2), p1), add(scl(d2, t2), p2));
	return dot(vv, vv);
}

/* geomutils/src/pcm/GuPCMContactSphereCapsule.cpp:38-96 */
static int c_sphere_capsule(float r, v3 c, float cr, float hh, xf t1, float cd, cset* o)
{
	v3 s, e;
	capsule_segment(t1, hh, &s, &e);
	const float rs = r + cr, inf = rs + cd;
	/* distancePointSegmentSquared (:38-53) */
	const v3 ap = sub(c, s), ab = sub(e, s);
	const float nom = dot(ap, ab), den = dot(ab, ab);
	const float t = den == 0.0f ? 0.0f : clamp01(nom / den);
	const v3 v = sub(ap, scl(ab, t));
	const float d2 = dot(v, v);
	if (!(inf * inf > d2))
		return 0;
	const v3 p = add(scl(sub(e, s), t), s);
	const v3 dir = sub(c, p);
	const float l = sqrtf(dot(dir, dir));
	const v3 n = l > 1.1920929e-07f ? scl(dir, 1.0f / l) : V(1.0f, 0.0f, 0.0f); /* V3NormalizeSafe */
	o->n = n;
	o->p[0] = sub(c, scl(n, r));
	o->sep[0] = sqrtf(d2) - rs;
	o->count = 1;
	return 1;
}

/* geomutils/src/pcm/GuPCMContactCapsuleCapsule.cpp:81-283: the closest points
 * of the two segments, or, for segments within 0.9998 of parallel, the
 * endpoints projecting inside the other segment (up to four points), all
 * computed about the midpoint of the two centres */
static int c_capsule_capsule(float r0, float hh0, xf x0, float r1, float hh1, xf x1, float cd, cset* o)
{
	const v3 off = scl(add(x0.p, x1.p), 0.5f);
	xf a = x0, b = x1;
	a.p = sub(x0.p, off);
	b.p = sub(x1.p, off);
	v3 s0, e0, s1, e1;
	capsule_segment(a, hh0, &s0, &e0);
	capsule_segment(b, hh1, &s1, &e1);
	const v3 d0 = sub(e0, s0), d1 = sub(e1, s1);
	const float sumR = r0 + r1, inf = sumR + cd, inf2 = inf * inf;
	const float la = dot(d0, d0), le = dot(d1, d1);
	const float eps = 1e-6f;
	float t0, t1;
	const float sq0 = seg_seg_sq(s0, d0, s1, d1, &t0, &t1);
	if (!(inf2 >= sq0))
		return 0;
	o->count = 0;
	const v3 dir0 = eps > la ? V(0, 0, 0) : scl(d0, 1.0f / sqrtf(la));
	const v3 dir1 = eps > le ? V(0, 0, 0) : scl(d1, 1.0f / sqrtf(le));
	if (fabsf(dot(dir0, dir1)) > 0.9998f)
	{
		/* pcmDistancePointSegmentTValue22 (:40-74): s1, e1 onto s0e0; s0, e0 onto s1e1 */
		const float tv[4] = {
			la == 0.0f ? 0.0f : dot(sub(s1, s0), d0) / la, la == 0.0f ? 0.0f : dot(sub(e1, s0), d0) / la,
			le == 0.0f ? 0.0f : dot(sub(s0, s1), d1) / le, le == 0.0f ? 0.0f : dot(sub(e0, s1), d1) / le};
		for (int k = 0; k < 4; k++)
		{
			if (!(tv[k] >= 0.0f && 1.0f >= tv[k]))
				continue;
			/* k < 2: a point of segment 1 against its projection on segment 0;
			 * k >= 2: an endpoint of segment 0 against its projection on segment 1 */
			const v3 on0 = k < 2 ? add(scl(d0, tv[k]), s0) : (k == 2 ? s0 : e0);
			const v3 on1 = k < 2 ? (k == 0 ? s1 : e1) : add(scl(d1, tv[k]), s1);
			const v3 v = sub(on0, on1);
			const float sq = dot(v, v);
			if (!(sq > eps && inf2 > sq))
				continue;
			const float dist = sqrtf(sq);
			const v3 n = scl(v, 1.0f / dist);
			o->n = n;
			o->p[o->count] = add(sub(on0, scl(n, r0)), off);
			o->sep[o->count] = dist - sumR;
			o->count++;
		}
		if (o->count)
			return o->count;
	}
	const v3 cA = add(scl(d0, t0), s0), cB = add(scl(d1, t1), s1);
	const int con = eps > sq0;
	const v3 nn = con ? (la > eps ? d0 : V(1.0f, 0.0f, 0.0f)) : sub(cA, cB);
	const v3 n = normalize(nn);
	o->n = n;
	o->p[0] = add(sub(cA, scl(n, r0)), off);
	o->sep[0] = (con ? 0.0f : sqrtf(sq0)) - sumR;
	o->count = 1;
	return 1;
}

/* geomutils/src/pcm/GuPCMContactPlaneCapsule.cpp:36-114 with a fresh manifold:
 * each endpoint closer to the plane than radius + contact distance, the second
 * replacing the first when they coincide within radius/1000
 * (addManifoldPoint2, GuPersistentContactManifold.cpp:1269); output with the
 * radius variant of addManifoldContactsToContactBuffer (:783-806): point on
 * the capsule surface, separation = endpoint height - radius, normal from the
 * capsule towards the plane */
static int c_plane_capsule(v3 pn, float pd, float r, float hh, xf tb, float cd, cset* o)
{
	v3 ends[2];
	capsule_segment(tb, hh, &ends[0], &ends[1]);
	const float repl = r * 0.001f;
	v3 pt[2];
	float sd[2];
	int n = 0;
	for (int k = 0; k < 2; k++)
	{
		const float sdist = dot(pn, ends[k]) + pd;
		if (!(r + cd > sdist))
			continue;
		if (n == 1)
		{
			const v3 da = sub(pt[0], ends[k]);
			const v3 db = sub(sub(pt[0], scl(pn, sd[0])), sub(ends[k], scl(pn, sdist)));
			if (repl * repl > fminf(dot(db, db), dot(da, da)))
			{
				pt[0] = ends[k];
				sd[0] = sdist;
				continue;
			}
		}
		pt[n] = ends[k];
		sd[n] = sdist;
		n++;
	}
	o->n = neg(pn);
	o->count = 0;
	for (int k = 0; k < n; k++)
		if (cd >= sd[k] - r)
		{
			o->p[o->count] = sub(pt[k], scl(pn, r));
			o->sep[o->count] = sd[k] - r;
			o->count++;
		}
	return o->count;
}

/* PersistentContactManifold::reduceBatchContactsCluster
 * (geomutils/src/pcm/GuPersistentContactManifold.cpp:859-1004) */
static void reduce_cluster(const pcontact* m, int n, pcontact* out)
{
	int chosen[64] = {0};
	int idx[4];
	float maxDist = 3.402823466e+38f;
	int index = 0;
	for (int i = 0; i < n; i++)
		if (maxDist > m[i].pen) { maxDist = m[i].pen; index = i; }
	out[0] = m[index]; chosen[index] = 1; idx[0] = index;
	v3 v = sub(m[0].lb, out[0].lb);
	maxDist = dot(v, v); index = 0;
	for (int i = 1; i < n; i++)
	{
		v = sub(m[i].lb, out[0].lb);
		const float d = dot(v, v);
		if (d > maxDist) { maxDist = d; index = i; }
	}
	out[1] = m[index]; chosen[index] = 1; idx[1] = index;
	maxDist = -3.402823466e+38f; index = -1;
	v = sub(out[1].lb, out[0].lb);
	v3 nrm = cross(v, out[0].ln);
	const float sq = dot(nrm, nrm);
	if (sq > 0.0f) { const float L = sqrtf(sq); nrm = V(nrm.x / L, nrm.y / L, nrm.z / L); } else nrm = out[0].ln; /* V3ScaleInv divides */
	float minDist = 3.402823466e+38f;
	int index1 = -1;
	for (int i = 0; i < n; i++)
		if (!chosen[i])
		{
			v = sub(m[i].lb, out[0].lb);
			const float d = dot(v, nrm);
			if (d > maxDist) { maxDist = d; index = i; }
			if (minDist > d) { minDist = d; index1 = i; }
		}
	out[2] = m[index]; chosen[index] = 1; idx[2] = index;
	if (minDist * maxDist > 0.0f)
	{
		maxDist = -3.402823466e+38f;
		for (int i = 0; i < n; i++)
			if (!chosen[i])
			{
				v = sub(m[i].lb, out[0].lb);
				const float d = dot(v, nrm);
				if (d > maxDist) { maxDist = d; index1 = i; }
			}
	}
	out[3] = m[index1]; chosen[index1] = 1; idx[3] = index1;
	for (int i = 0; i < n; i++)
		if (!chosen[i])
		{
			float md = 3.402823466e+38f;
			const float pen = m[i].pen;
			int k = 0;
			for (int j = 0; j < 4; j++)
			{
				const v3 w = sub(m[i].lb, out[j].lb);
				const float d = dot(w, w);
				if (md > d) { md = d; k = j; }
			}
			if (m[idx[k]].pen > pen)
				idx[k] = i;
		}
	for (int j = 0; j < 4; j++)
		out[j] = m[idx[j]];
}

/* PersistentContactManifold::reduceBatchContacts
 * (geomutils/src/pcm/GuPersistentContactManifold.cpp:1008-1172) */
static void reduce_batch(const pcontact* m, int n, float tol, pcontact* out)
{
	int ch[4];
	int cand[64];
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
	v3 v = sub(m[cand[0]].lb, m[ch[0]].lb);
	float maxDist = dot(v, v);
	index = cand[0]; ci = 0;
	for (int i = 1; i < nc; i++)
	{
		v = sub(m[cand[i]].lb, m[ch[0]].lb);
		const float d = dot(v, v);
		if (d > maxDist) { maxDist = d; index = cand[i]; ci = i; }
	}
	ch[1] = index;
	nc--; cand[ci] = cand[nc];
	v = sub(m[ch[1]].lb, m[ch[0]].lb);
	const v3 cn0 = m[ch[0]].ln;
	v3 nrm = cross(v, cn0);
	const float sq = dot(nrm, nrm);
	if (sq > 0.0f) { const float L = sqrtf(sq); nrm = V(nrm.x / L, nrm.y / L, nrm.z / L); } else nrm = cn0; /* V3ScaleInv divides */
	maxDist = -3.402823466e+38f; index = -1; ci = -1;
	float minDist = 3.402823466e+38f;
	int index1 = -1, ci1 = -1;
	for (int i = 0; i < nc; i++)
	{
		v = sub(m[cand[i]].lb, m[ch[0]].lb);
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
		maxDist = -3.402823466e+38f;
		for (int i = 0; i < nc; i++)
		{
			v = sub(m[cand[i]].lb, m[ch[0]].lb);
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
				if (cidx >= 0 && cidx < nc)
				{
					const int orig = ch[i];
					ch[i] = cand[cidx];
					cand[cidx] = orig;
				}
			}
			out[i] = m[ch[i]];
		}
	}
	else
		for (int i = 0; i < 4; i++)
			out[i] = m[ch[i]];
}

/* geomutils/src/pcm/GuPCMContactPlaneBox.cpp:36-208 (shape0 plane, shape1 box).
 * Returns contacts with the normal from the box towards the plane. */
static int c_plane_box(v3 pn, float pd, v3 h, xf tb, float cd, cset* o)
{
	/* box corner order of the reference: (x,y,z) (x,y,-z) (x,-y,z) (x,-y,-z) (-x,y,z) ... */
	static const float sg[8][3] = {{1, 1, 1}, {1, 1, -1}, {1, -1, 1}, {1, -1, -1}, {-1, 1, 1}, {-1, 1, -1}, {-1, -1, 1}, {-1, -1, -1}};
	pcontact mc[8];
	int nc = 0;
	const m33 R = qmat(tb.q);
	const v3 r0 = scl(R.c0, h.x), r1 = scl(R.c1, h.y), r2 = scl(R.c2, h.z);
	const float px = dot(pn, tb.p) + pd;  /* box centre height above the plane */
	for (int k = 0; k < 8; k++)
	{
		const v3 off = add(add(scl(r0, sg[k][0]), scl(r1, sg[k][1])), scl(r2, sg[k][2]));
		const float s = dot(pn, off);
		if (cd - px > s)
		{
			const float pen = s + px;
			const v3 corner = add(tb.p, off);
			mc[nc].la = V(sg[k][0] * h.x, sg[k][1] * h.y, sg[k][2] * h.z);
			mc[nc].lb = sub(corner, scl(pn, pen)); /* world point on the plane */
			mc[nc].ln = pn;
			mc[nc].pen = pen;
			nc++;
		}
	}
	if (nc == 0)
		return 0;
	pcontact red[4];
	const pcontact* use = mc;
	if (nc > 4)
	{
		reduce_cluster(mc, nc, red);
		use = red;
		nc = 4;
	}
	o->n = neg(pn);
	o->count = 0;
	for (int i = 0; i < nc; i++)
		if (cd >= use[i].pen)
		{
			o->p[o->count] = use[i].lb;
			o->sep[o->count] = use[i].pen;
			o->count++;
		}
	return o->count;
}

/* getIncidentPolygon (geomutils/src/pcm/GuPCMContactBoxBox.cpp:42-124) */
static void incident_polygon(v3* pts, v3* faceNormal, v3 axis, mxf t1To0, v3 e)
{
	const v3 u0 = t1To0.r.c0, u1 = t1To0.r.c1, u2 = t1To0.r.c2;
	const float d0 = dot(u0, axis), d1 = dot(u1, axis), d2 = dot(u2, axis);
	const float a0 = fabsf(d0), a1 = fabsf(d1), a2 = fabsf(d2);
	float ex = e.x, ey = e.y, ez = e.z;
	if (a0 >= a1 && a0 >= a2)
	{
		const int con = d0 > 0.0f;
		*faceNormal = con ? neg(u0) : u0;
		ex = con ? -ex : ex;
		const v3 r0 = scl(u0, ex), r1 = scl(u1, ey), r2 = scl(u2, ez);
		const v3 t0 = add(t1To0.p, r0), t1 = add(r1, r2), t2 = sub(r1, r2);
		pts[0] = add(t0, t1); pts[1] = add(t0, t2); pts[2] = sub(t0, t1); pts[3] = sub(t0, t2);
	}
	else if (a1 >= a2)
	{
		const int con = d1 > 0.0f;
		*faceNormal = con ? neg(u1) : u1;
		ey = con ? -ey : ey;
		const v3 r0 = scl(u0, ex), r1 = scl(u1, ey), r2 = scl(u2, ez);
		const v3 t0 = add(t1To0.p, r1), t1 = add(r0, r2), t2 = sub(r0, r2);
		pts[0] = add(t0, t1); pts[1] = add(t0, t2); pts[2] = sub(t0, t1); pts[3] = sub(t0, t2);
	}
	else
	{
		const int con = d2 > 0.0f;
		*faceNormal = con ? neg(u2) : u2;
		ez = con ? -ez : ez;
		const v3 r0 = scl(u0, ex), r1 = scl(u1, ey), r2 = scl(u2, ez);
		const v3 t0 = add(t1To0.p, r2), t1 = add(r0, r1), t2 = sub(r0, r1);
		pts[0] = add(t0, t1); pts[1] = add(t0, t2); pts[2] = sub(t0, t1); pts[3] = sub(t0, t2);
	}
}

/* intersectSegmentAABB (GuPCMContactBoxBox.cpp:127-167); the reference uses
 * a fast reciprocal (V3RecipFast, ~12-bit), this restatement divides exactly. */
static int seg_aabb(v3 p0, v3 d, v3 mx, v3 mn, float* tmin, float* tmax)
{
	const float eps = 1e-6f;
	float t1[3], t2[3];
	for (int i = 0; i < 3; i++)
	{
		const float pi = comp(p0, i), di = comp(d, i), mxi = comp(mx, i), mni = comp(mn, i);
		const int par = eps > fabsf(di);
		const int out = (pi > mxi) || (mni > pi);
		if (par && out)
			return 0;
		const float od = 1.0f / di;
		t1[i] = par ? 0.0f : (mni - pi) * od;
		t2[i] = par ? 3.402823466e+38f : (mxi - pi) * od;
	}
	float ft1 = -3.402823466e+38f, ft2 = 3.402823466e+38f;
	for (int i = 0; i < 3; i++)
	{
		ft1 = fmaxf(ft1, fminf(t1[i], t2[i]));
		ft2 = fminf(ft2, fmaxf(t1[i], t2[i]));
	}
	const float a = fmaxf(ft1, 0.0f), b = fminf(1.0f, ft2);
	*tmin = a;
	*tmax = b;
	return !((a > b) || (a > 1.0f));
}

/* Gu::contains (geomutils/src/pcm/GuPCMContactGenUtil.cpp:35-103) */
static int poly_contains(const v3* verts, int nv, v3 p, v3 mn, v3 mx)
{
	if ((mn.x > p.x || p.x > mx.x) || (mn.y > p.y || p.y > mx.y))
		return 0;
	const float tx = p.x, ty = p.y, eps = 1.1920928955078125e-7f;
	int inter = 0;
	for (int i = 0, j = nv - 1; i < nv; j = i++)
	{
		const float jy = verts[j].y, iy = verts[i].y, jx = verts[j].x, ix = verts[i].x;
		if ((tx == jx && ty == jy) || (tx == ix && ty == iy))
			return 1;
		const int f0 = jy > ty, f1 = iy > ty;
		if (f0 != f1)
		{
			const float jix = ix - jx, jiy = iy - jy, jty = ty - jy;
			const float part1 = jty * jix;
			const float part2 = (jx + eps) * jiy;
			const float part3 = tx * jiy;
			const int cmp = jiy > 0.0f;
			const float tmp = part1 + part2;
			const float c1 = cmp ? tmp : part3, c2 = cmp ? part3 : tmp;
			if (c1 >= c2)
			{
				if (inter == 1)
					return 0;
				inter++;
			}
		}
	}
	return inter > 0;
}

/* calculateContacts (GuPCMContactBoxBox.cpp:170-330) */
static void calc_contacts(float exIn, float eyIn, const v3* pts, v3 incN, v3 ln, pcontact* mc, int* nc, float cd)
{
	const float big = 3.402823466e+38f;
	const float ex = exIn * 1.0001f, ey = eyIn * 1.0001f;
	int pen[4], area[4];
	v3 bmin = V(big, big, big), bmax = V(-big, -big, -big);
	for (int i = 0; i < 4; i++)
	{
		bmin = V(fminf(bmin.x, pts[i].x), fminf(bmin.y, pts[i].y), fminf(bmin.z, pts[i].z));
		bmax = V(fmaxf(bmax.x, pts[i].x), fmaxf(bmax.y, pts[i].y), fmaxf(bmax.z, pts[i].z));
		const float z = -pts[i].z;
		if (cd > z)
		{
			pen[i] = 1;
			if (ex >= fabsf(pts[i].x) && ey >= fabsf(pts[i].y))
			{
				area[i] = 1;
				mc[*nc].la = V(pts[i].x, pts[i].y, 0.0f);
				mc[*nc].lb = pts[i];
				mc[*nc].ln = ln;
				mc[*nc].pen = z;
				(*nc)++;
			}
			else
				area[i] = 0;
		}
		else
		{
			pen[i] = 0;
			area[i] = 0;
		}
	}
	if (*nc == 4)
		return;
	const float denom = incN.z;
	const float qs[4][2] = {{ex, ey}, {ex, -ey}, {-ex, ey}, {-ex, -ey}};
	for (int k = 0; k < 4; k++)
	{
		const v3 q0 = V(qs[k][0], qs[k][1], 0.0f);
		if (poly_contains(pts, 4, q0, bmin, bmax))
		{
			const float nom = dot(incN, sub(pts[0], q0));
			const float t = nom / denom;
			const float pn = -t;
			if (cd > pn)
			{
				mc[*nc].la = q0;
				mc[*nc].lb = V(q0.x, q0.y, t);
				mc[*nc].ln = ln;
				mc[*nc].pen = pn;
				(*nc)++;
			}
		}
	}
	const v3 ext = V(ex, ey, big);
	const v3 negExt = V(-ex, -ey, -(cd + 1.1920928955078125e-7f));
	for (int rs = 0, re = 3; rs < 4; re = rs++)
	{
		const v3 p0 = pts[rs], p1 = pts[re];
		if (!pen[rs] && !pen[re])
			continue;
		const int c0 = pen[rs] && area[rs];
		const int c1 = pen[re] && area[re];
		if (c0 && c1)
			continue;
		const v3 d = sub(p1, p0);
		float tmin, tmax;
		if (seg_aabb(p0, d, ext, negExt, &tmin, &tmax))
		{
			if (!c0)
			{
				const v3 ip = add(scl(d, tmin), p0);
				mc[*nc].la = V(ip.x, ip.y, 0.0f);
				mc[*nc].lb = ip;
				mc[*nc].ln = ln;
				mc[*nc].pen = -ip.z;
				(*nc)++;
			}
			if (!c1)
			{
				const v3 ip = add(scl(d, tmax), p0);
				mc[*nc].la = V(ip.x, ip.y, 0.0f);
				mc[*nc].lb = ip;
				mc[*nc].ln = ln;
				mc[*nc].pen = -ip.z;
				(*nc)++;
			}
		}
	}
}

static void seg_closest(v3 p1, v3 q1, v3 p2, v3 q2, v3* c1, v3* c2)
{
	/* closest points of two segments (Ericson, Real-Time Collision Detection 5.1.9) */
	const v3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
	const float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
	float s, t;
	const float c = dot(d1, r), b = dot(d1, d2), den = a * e - b * b;
	s = den > 1e-12f ? fminf(fmaxf((b * f - c * e) / den, 0.0f), 1.0f) : 0.0f;
	t = (b * s + f) / e;
	if (t < 0.0f) { t = 0.0f; s = fminf(fmaxf(-c / a, 0.0f), 1.0f); }
	else if (t > 1.0f) { t = 1.0f; s = fminf(fmaxf((b - c) / a, 0.0f), 1.0f); }
	*c1 = add(p1, scl(d1, s));
	*c2 = add(p2, scl(d2, t));
}

/* Restatement of the GJK/EPA fallback of pcmContactBoxBox
 * (GuPCMContactBoxBox.cpp:922-958) for the case where face clipping leaves no
 * point: the minimum-translation axis over the 15 box SAT axes, with the
 * witness on box1 taken from the closest points of the supporting edges. */
static int box_box_edge(v3 e0, mxf t0, v3 e1, mxf t1, float cd, pcontact* mc)
{
	const v3 a[3] = {t0.r.c0, t0.r.c1, t0.r.c2}, b[3] = {t1.r.c0, t1.r.c1, t1.r.c2};
	const float ea[3] = {e0.x, e0.y, e0.z}, eb[3] = {e1.x, e1.y, e1.z};
	const v3 T = sub(t0.p, t1.p); /* from box1 to box0 */
	float best = -3.402823466e+38f;
	v3 bestN = V(0, 1, 0);
	int bi = -1, bj = -1;
	for (int k = 0; k < 15; k++)
	{
		v3 L;
		int i = -1, j = -1;
		if (k < 3) L = a[k];
		else if (k < 6) L = b[k - 3];
		else { i = (k - 6) / 3; j = (k - 6) % 3; L = cross(a[i], b[j]); }
		const float l2 = dot(L, L);
		if (l2 < 1e-10f)
			continue;
		L = scl(L, 1.0f / sqrtf(l2));
		const float ra = ea[0] * fabsf(dot(a[0], L)) + ea[1] * fabsf(dot(a[1], L)) + ea[2] * fabsf(dot(a[2], L));
		const float rb = eb[0] * fabsf(dot(b[0], L)) + eb[1] * fabsf(dot(b[1], L)) + eb[2] * fabsf(dot(b[2], L));
		const float s = dot(T, L);
		const float sep = fabsf(s) - (ra + rb);
		if (sep > best)
		{
			best = sep;
			bestN = s < 0.0f ? neg(L) : L;
			bi = i;
			bj = j;
		}
	}
	if (best > cd)
		return 0;
	if (best > 0.0f)
	{
		/* separated: GJK's closest points, here by exhaustive feature pairs
		 * (vertex against box, edge against edge) */
		float dmin = 3.402823466e+38f;
		v3 qa = t0.p, qb = t1.p;
		for (int side = 0; side < 2; side++)
		{
			const mxf tv = side ? t1 : t0, tb = side ? t0 : t1;
			const v3 ev = side ? e1 : e0, eb = side ? e0 : e1;
			for (int k = 0; k < 8; k++)
			{
				const v3 lv = V((k & 4) ? -ev.x : ev.x, (k & 2) ? -ev.y : ev.y, (k & 1) ? -ev.z : ev.z);
				const v3 wv = mxtrans(tv, lv);
				const v3 lb = mtmul(tb.r, sub(wv, tb.p));
				const v3 cl = V(fminf(fmaxf(lb.x, -eb.x), eb.x), fminf(fmaxf(lb.y, -eb.y), eb.y), fminf(fmaxf(lb.z, -eb.z), eb.z));
				const v3 wq = mxtrans(tb, cl);
				const float dd = len2(sub(wv, wq));
				if (dd < dmin) { dmin = dd; qa = side ? wq : wv; qb = side ? wv : wq; }
			}
		}
		for (int ia = 0; ia < 12; ia++)
			for (int ib = 0; ib < 12; ib++)
			{
				v3 s0[2], s1[2];
				for (int side = 0; side < 2; side++)
				{
					const int e = side ? ib : ia;
					const mxf t = side ? t1 : t0;
					const v3 ex = side ? e1 : e0;
					const int ax = e / 4, u = (e & 1) ? -1 : 1, w = (e & 2) ? -1 : 1;
					v3 c = V(0, 0, 0), dlt = V(0, 0, 0);
					const float ev[3] = {ex.x, ex.y, ex.z};
					float cc[3] = {0, 0, 0}, dd[3] = {0, 0, 0};
					cc[(ax + 1) % 3] = u * ev[(ax + 1) % 3];
					cc[(ax + 2) % 3] = w * ev[(ax + 2) % 3];
					dd[ax] = ev[ax];
					c = V(cc[0], cc[1], cc[2]);
					dlt = V(dd[0], dd[1], dd[2]);
					(side ? s1 : s0)[0] = mxtrans(t, sub(c, dlt));
					(side ? s1 : s0)[1] = mxtrans(t, add(c, dlt));
				}
				v3 c0, c1;
				seg_closest(s0[0], s0[1], s1[0], s1[1], &c0, &c1);
				const float dd = len2(sub(c0, c1));
				if (dd < dmin) { dmin = dd; qa = c0; qb = c1; }
			}
		const float dist = sqrtf(dmin);
		if (dist > cd)
			return 0;
		const v3 nrm = dist > 0.0f ? scl(sub(qa, qb), 1.0f / dist) : bestN;
		mc[0].la = mtmul(t0.r, sub(qa, t0.p));
		mc[0].lb = mtmul(t1.r, sub(qb, t1.p));
		mc[0].ln = mtmul(t1.r, nrm);
		mc[0].pen = dist;
		return 1;
	}
	v3 pB;
	if (bi >= 0)
	{
		/* supporting edge of box0 along -n, of box1 along +n */
		v3 c0 = t0.p, c1 = t1.p;
		for (int k = 0; k < 3; k++)
		{
			if (k != bi) c0 = sub(c0, scl(a[k], dot(a[k], bestN) > 0.0f ? ea[k] : -ea[k]));
			if (k != bj) c1 = add(c1, scl(b[k], dot(b[k], bestN) > 0.0f ? eb[k] : -eb[k]));
		}
		v3 q0, q1;
		seg_closest(sub(c0, scl(a[bi], ea[bi])), add(c0, scl(a[bi], ea[bi])), sub(c1, scl(b[bj], eb[bj])),
		            add(c1, scl(b[bj], eb[bj])), &q0, &q1);
		pB = q1;
	}
	else
	{
		/* face axis: deepest vertex of the incident box */
		v3 c = bestN;
		(void)c;
		pB = t1.p;
		for (int k = 0; k < 3; k++)
			pB = add(pB, scl(b[k], dot(b[k], bestN) > 0.0f ? eb[k] : -eb[k]));
	}
	/* stored in box-local spaces like the reference manifold */
	mc[0].la = mtmul(t0.r, sub(add(pB, scl(bestN, -best)), t0.p));
	mc[0].lb = mtmul(t1.r, sub(pB, t1.p));
	mc[0].ln = mtmul(t1.r, bestN);
	mc[0].pen = best;
	return 1;
}

/* doBoxBoxGenerateContacts + pcmContactBoxBox on a fresh manifold
 * (geomutils/src/pcm/GuPCMContactBoxBox.cpp:332-818, 820-968) */
static int c_box_box(v3 e0, xf x0, v3 e1, xf x1, float cd, float tol, cset* o)
{
	mxf t0, t1;
	t0.r = qmat(x0.q); t0.p = x0.p;
	t1.r = qmat(x1.q); t1.p = x1.p;
	t0.r.c0 = normalize(t0.r.c0); t0.r.c1 = normalize(t0.r.c1); t0.r.c2 = normalize(t0.r.c2);
	t1.r.c0 = normalize(t1.r.c0); t1.r.c1 = normalize(t1.r.c1); t1.r.c2 = normalize(t1.r.c2);

	const mxf t1To0 = mxinvmul(t0, t1);
	const m33 r0To1 = mtrans(t1To0.r);
	const float uEps = 1e-6f;
	const float ea0 = e0.x, ea1 = e0.y, ea2 = e0.z, eb0 = e1.x, eb1 = e1.y, eb2 = e1.z;
	const float tx = t1To0.p.x, ty = t1To0.p.y, tz = t1To0.p.z;
	const v3 col0 = t1To0.r.c0, col1 = t1To0.r.c1, col2 = t1To0.r.c2;
	const v3 a10c0 = add(vabs(col0), V(uEps, uEps, uEps)), a10c1 = add(vabs(col1), V(uEps, uEps, uEps)),
	         a10c2 = add(vabs(col2), V(uEps, uEps, uEps));
	const v3 a01c0 = add(vabs(r0To1.c0), V(uEps, uEps, uEps)), a01c1 = add(vabs(r0To1.c1), V(uEps, uEps, uEps)),
	         a01c2 = add(vabs(r0To1.c2), V(uEps, uEps, uEps));
	float sign[6], overlap[6], ra, rb, rs;
	v3 vt;

	sign[0] = tx; vt = mulv(a01c0, e1); rb = vt.x + (vt.y + vt.z);
	overlap[0] = ((ea0 + rb) - fabsf(sign[0])) + cd; if (0.0f > overlap[0]) return 0;
	sign[1] = ty; vt = mulv(a01c1, e1); rb = vt.x + (vt.y + vt.z);
	overlap[1] = ((ea1 + rb) - fabsf(sign[1])) + cd; if (0.0f > overlap[1]) return 0;
	sign[2] = tz; vt = mulv(a01c2, e1); rb = vt.x + (vt.y + vt.z);
	overlap[2] = ((ea2 + rb) - fabsf(sign[2])) + cd; if (0.0f > overlap[2]) return 0;
	sign[3] = dot(t1To0.p, col0); vt = mulv(a10c0, e0); ra = vt.x + (vt.y + vt.z);
	overlap[3] = ((ra + eb0) - fabsf(sign[3])) + cd; if (0.0f > overlap[3]) return 0;
	sign[4] = dot(t1To0.p, col1); vt = mulv(a10c1, e0); ra = vt.x + (vt.y + vt.z);
	overlap[4] = ((ra + eb1) - fabsf(sign[4])) + cd; if (0.0f > overlap[4]) return 0;
	sign[5] = dot(t1To0.p, col2); vt = mulv(a10c2, e0); ra = vt.x + (vt.y + vt.z);
	overlap[5] = ((ra + eb2) - fabsf(sign[5])) + cd; if (0.0f > overlap[5]) return 0;

#define EDGE_TEST(absSign, raA, raB, rbA, rbB) \
	{ ra = (raA) + (raB); rb = (rbA) + (rbB); rs = (ra + rb) + cd; if ((absSign) > rs) return 0; }
	EDGE_TEST(fabsf(col0.y * tz - col0.z * ty), a10c0.z * ea1, a10c0.y * ea2, a01c0.z * eb1, a01c0.y * eb2);
	EDGE_TEST(fabsf(col1.y * tz - col1.z * ty), a10c1.z * ea1, a10c1.y * ea2, a01c0.z * eb0, a01c0.x * eb2);
	EDGE_TEST(fabsf(col2.y * tz - col2.z * ty), a10c2.z * ea1, a10c2.y * ea2, a01c0.y * eb0, a01c0.x * eb1);
	EDGE_TEST(fabsf(col0.z * tx - col0.x * tz), a10c0.z * ea0, a10c0.x * ea2, a01c1.z * eb1, a01c1.y * eb2);
	EDGE_TEST(fabsf(col1.z * tx - col1.x * tz), a10c1.z * ea0, a10c1.x * ea2, a01c1.z * eb0, a01c1.x * eb2);
	EDGE_TEST(fabsf(col2.z * tx - col2.x * tz), a10c2.z * ea0, a10c2.x * ea2, a01c1.y * eb0, a01c1.x * eb1);
	EDGE_TEST(fabsf(col0.x * ty - col0.y * tx), a10c0.y * ea0, a10c0.x * ea1, a01c2.z * eb1, a01c2.y * eb2);
	EDGE_TEST(fabsf(col1.x * ty - col1.y * tx), a10c1.y * ea0, a10c1.x * ea1, a01c2.z * eb0, a01c2.x * eb2);
	EDGE_TEST(fabsf(col2.x * ty - col2.y * tx), a10c2.y * ea0, a10c2.x * ea1, a01c2.y * eb0, a01c2.x * eb1);
#undef EDGE_TEST

	int feature = 0;
	float minOverlap = overlap[0];
	for (int i = 1; i < 6; i++)
		if (minOverlap > overlap[i]) { minOverlap = overlap[i]; feature = i; }

	const v3 ax00 = t0.r.c0, ax01 = t0.r.c1, ax02 = t0.r.c2, ax10 = t1.r.c0, ax11 = t1.r.c1, ax12 = t1.r.c2;
	mxf nt;
	v3 mtd, incN, pts[4];
	int flip = 0;
	pcontact mc[16];
	int nc = 0;
	const int le0 = 0.0f >= sign[feature];
	switch (feature)
	{
	case 0:
		if (le0) { mtd = ax00; nt.r.c0 = neg(ax02); nt.r.c1 = ax01; nt.r.c2 = ax00; nt.p = sub(t0.p, scl(ax00, ea0)); }
		else { mtd = neg(ax00); nt.r.c0 = ax02; nt.r.c1 = ax01; nt.r.c2 = neg(ax00); nt.p = add(scl(ax00, ea0), t0.p); }
		break;
	case 1:
		if (le0) { mtd = ax01; nt.r.c0 = ax00; nt.r.c1 = neg(ax02); nt.r.c2 = ax01; nt.p = sub(t0.p, scl(ax01, ea1)); }
		else { mtd = neg(ax01); nt.r.c0 = ax00; nt.r.c1 = ax02; nt.r.c2 = neg(ax01); nt.p = add(scl(ax01, ea1), t0.p); }
		break;
	case 2:
		if (le0) { mtd = ax02; nt.r.c0 = ax00; nt.r.c1 = ax01; nt.r.c2 = ax02; nt.p = sub(t0.p, scl(ax02, ea2)); }
		else { mtd = neg(ax02); nt.r.c0 = ax00; nt.r.c1 = neg(ax01); nt.r.c2 = neg(ax02); nt.p = add(scl(ax02, ea2), t0.p); }
		break;
	case 3:
		flip = 1;
		if (le0) { mtd = ax10; nt.r.c0 = ax12; nt.r.c1 = ax11; nt.r.c2 = neg(ax10); nt.p = add(scl(ax10, eb0), t1.p); }
		else { mtd = neg(ax10); nt.r.c0 = neg(ax12); nt.r.c1 = ax11; nt.r.c2 = ax10; nt.p = sub(t1.p, scl(ax10, eb0)); }
		break;
	case 4:
		flip = 1;
		if (le0) { mtd = ax11; nt.r.c0 = ax10; nt.r.c1 = ax12; nt.r.c2 = neg(ax11); nt.p = add(scl(ax11, eb1), t1.p); }
		else { mtd = neg(ax11); nt.r.c0 = ax10; nt.r.c1 = neg(ax12); nt.r.c2 = ax11; nt.p = sub(t1.p, scl(ax11, eb1)); }
		break;
	default:
		flip = 1;
		if (le0) { mtd = ax12; nt.r.c0 = ax10; nt.r.c1 = neg(ax11); nt.r.c2 = neg(ax12); nt.p = add(scl(ax12, eb2), t1.p); }
		else { mtd = neg(ax12); nt.r.c0 = ax10; nt.r.c1 = ax11; nt.r.c2 = ax12; nt.p = sub(t1.p, scl(ax12, eb2)); }
		break;
	}
	const v3 ln = mtmul(nt.r, mtd);
	if (!flip)
	{
		const mxf t1ToNew = mxinvmul(nt, t1);
		incident_polygon(pts, &incN, neg(ln), t1ToNew, e1);
		if (feature == 0) calc_contacts(ea2, ea1, pts, incN, ln, mc, &nc, cd);
		else if (feature == 1) calc_contacts(ea0, ea2, pts, incN, ln, mc, &nc, cd);
		else calc_contacts(ea0, ea1, pts, incN, ln, mc, &nc, cd);
	}
	else
	{
		const mxf t0ToNew = mxinvmul(nt, t0);
		incident_polygon(pts, &incN, ln, t0ToNew, e0);
		if (feature == 3) calc_contacts(eb2, eb1, pts, incN, ln, mc, &nc, cd);
		else if (feature == 4) calc_contacts(eb0, eb2, pts, incN, ln, mc, &nc, cd);
		else calc_contacts(eb0, eb1, pts, incN, ln, mc, &nc, cd);
	}

	v3 worldN;
	if (nc != 0)
	{
		if (flip)
			for (int i = 0; i < nc; i++) { const v3 t = mc[i].lb; mc[i].lb = mc[i].la; mc[i].la = t; }
		const mxf newTo1 = mxinvmul(t1, nt);
		const mxf newTo0 = mxinvmul(t0, nt);
		const v3 lnB = mmul(newTo1.r, mc[0].ln);
		for (int i = 0; i < nc; i++)
		{
			mc[i].la = mxtrans(newTo0, mc[i].la);
			mc[i].lb = mxtrans(newTo1, mc[i].lb);
			mc[i].ln = lnB;
		}
		pcontact red[4];
		const pcontact* use = mc;
		if (nc > 4)
		{
			reduce_batch(mc, nc, tol, red);
			use = red;
			nc = 4;
		}
		worldN = normalize(mmul(t1.r, use[0].ln));
		o->n = worldN;
		o->count = nc;
		for (int i = 0; i < nc; i++)
		{
			o->p[i] = mxtrans(t1, use[i].lb);
			o->sep[i] = use[i].pen;
		}
		return nc;
	}
	if (!box_box_edge(e0, t0, e1, t1, cd, mc))
		return 0;
	o->n = normalize(mmul(t1.r, mc[0].ln));
	o->count = 1;
	o->p[0] = mxtrans(t1, mc[0].lb);
	o->sep[0] = mc[0].pen;
	return 1;
}

/* ------------------------------------------------------------------------- */
/* Convex meshes (PxConvexMeshGeometry, identity scale) and the box seen as a */
/* polytope, PCM contact generation on a fresh manifold                       */
/* ------------------------------------------------------------------------- */

/* Gu::PolygonalData (geomutils/src/pcm/GuPCMShapeConvex.h) plus the margins
 * the contact generation reads */
typedef struct {
	int nv, np, ne;
	const v3* v;
	const v3* pn;
	const float* pd;
	const int* pf;      /* polygon first index into vr */
	const int* pc;      /* polygon vertex count */
	const int* vr;
	const int* ev;      /* 2 per edge: vertices */
	const int* ef;      /* 2 per edge: faces (v0 -> v1 runs counter-clockwise in the first) */
	v3 bc, be;          /* local bounds centre and extents */
	float margin;       /* ConvexHullV / BoxV getMarginF */
} polyd;

typedef struct {
	polyd d;
	v3* v; v3* pn; float* pd; int* pf; int* pc; int* vr; int* ev; int* ef;
} ohull;

static ohull* g_hull = NULL;
static int g_nhull = 0, g_hcap = 0;

/* edges of a polygon soup, in order of first appearance */
static int build_edges(int np, const int* pf, const int* pc, const int* vr, int* ev, int* ef)
{
	int ne = 0;
	for (int p = 0; p < np; p++)
		for (int k = 0; k < pc[p]; k++)
		{
			const int a = vr[pf[p] + k], b = vr[pf[p] + (k + 1) % pc[p]];
			int found = -1;
			for (int e = 0; e < ne; e++)
				if (ev[2 * e] == b && ev[2 * e + 1] == a)
				{
					found = e;
					break;
				}
			if (found >= 0)
				ef[2 * found + 1] = p;
			else
			{
				ev[2 * ne] = a;
				ev[2 * ne + 1] = b;
				ef[2 * ne] = p;
				ef[2 * ne + 1] = p;
				ne++;
			}
		}
	return ne;
}

int pxo_hull_add(int nv, const float* verts, int np, const float* planes, const int32_t* polys, const int32_t* vrefs,
                 const float* internal_extents)
{
	if (g_nhull == g_hcap)
	{
		g_hcap = g_hcap ? 2 * g_hcap : 64;
		g_hull = (ohull*)realloc(g_hull, sizeof(ohull) * (size_t)g_hcap);
	}
	ohull* h = &g_hull[g_nhull];
	int nr = 0;
	for (int p = 0; p < np; p++)
		if (polys[2 * p] + polys[2 * p + 1] > nr)
			nr = polys[2 * p] + polys[2 * p + 1];
	h->v = (v3*)malloc(sizeof(v3) * (size_t)nv);
	h->pn = (v3*)malloc(sizeof(v3) * (size_t)np);
	h->pd = (float*)malloc(sizeof(float) * (size_t)np);
	h->pf = (int*)malloc(sizeof(int) * (size_t)np);
	h->pc = (int*)malloc(sizeof(int) * (size_t)np);
	h->vr = (int*)malloc(sizeof(int) * (size_t)(nr ? nr : 1));
	h->ev = (int*)malloc(sizeof(int) * 2 * (size_t)nr);
	h->ef = (int*)malloc(sizeof(int) * 2 * (size_t)nr);
	v3 mn = V(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f), mx = neg(mn);
	for (int i = 0; i < nv; i++)
	{
		h->v[i] = V(verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
		mn = V(fminf(mn.x, h->v[i].x), fminf(mn.y, h->v[i].y), fminf(mn.z, h->v[i].z));
		mx = V(fmaxf(mx.x, h->v[i].x), fmaxf(mx.y, h->v[i].y), fmaxf(mx.z, h->v[i].z));
	}
	for (int p = 0; p < np; p++)
	{
		h->pn[p] = V(planes[4 * p], planes[4 * p + 1], planes[4 * p + 2]);
		h->pd[p] = planes[4 * p + 3];
		h->pf[p] = polys[2 * p];
		h->pc[p] = polys[2 * p + 1];
	}
	for (int k = 0; k < nr; k++)
		h->vr[k] = vrefs[k];
	polyd* d = &h->d;
	d->nv = nv; d->np = np;
	d->ne = build_edges(np, h->pf, h->pc, h->vr, h->ev, h->ef);
	d->v = h->v; d->pn = h->pn; d->pd = h->pd; d->pf = h->pf; d->pc = h->pc; d->vr = h->vr;
	d->ev = h->ev; d->ef = h->ef;
	/* PxBounds3 centre / extents (physx/include/foundation/PxBounds3.h) */
	d->bc = scl(add(mx, mn), 0.5f);
	d->be = scl(sub(mx, mn), 0.5f);
	/* ConvexHullV margin: 10% of the smallest internal extent
	 * (CalculateConvexMargin, geomutils/src/gjk/GuVecConvexHull.h:76-93) */
	const float me = fminf(internal_extents[0], fminf(internal_extents[1], internal_extents[2]));
	d->margin = me * 0.1f;
	return g_nhull++;
}

void pxo_hull_clear(void)
{
	for (int i = 0; i < g_nhull; i++)
	{
		ohull* h = &g_hull[i];
		free(h->v); free(h->pn); free(h->pd); free(h->pf); free(h->pc); free(h->vr); free(h->ev); free(h->ef);
	}
	g_nhull = 0;
}

static const polyd* hull_of(float id) { return &g_hull[(int)id].d; }

/* PCMPolygonalBox (geomutils/src/pcm/GuPCMShapeConvex.cpp:40-136): corners,
 * the six quads of gPCMBoxPolygonData and their planes; BoxV margin is 15%
 * of the smallest half extent (geomutils/src/gjk/GuVecBox.h:46, :56-75) */
typedef struct { polyd d; v3 v[8]; v3 pn[6]; float pd[6]; int pf[6], pc[6], vr[24], ev[24], ef[24]; } boxpoly;
static void box_poly(v3 h, boxpoly* b)
{
	static const int quads[24] = {0, 3, 2, 1, 1, 2, 6, 5, 5, 6, 7, 4, 4, 7, 3, 0, 3, 7, 6, 2, 4, 0, 1, 5};
	b->v[0] = V(-h.x, -h.y, -h.z); b->v[1] = V(h.x, -h.y, -h.z); b->v[2] = V(h.x, h.y, -h.z);
	b->v[3] = V(-h.x, h.y, -h.z); b->v[4] = V(-h.x, -h.y, h.z); b->v[5] = V(h.x, -h.y, h.z);
	b->v[6] = V(h.x, h.y, h.z); b->v[7] = V(-h.x, h.y, h.z);
	for (int i = 0; i < 24; i++)
		b->vr[i] = quads[i];
	for (int i = 0; i < 6; i++) { b->pf[i] = 4 * i; b->pc[i] = 4; }
	b->pn[0] = V(0, 0, -1); b->pd[0] = -h.z;
	b->pn[1] = V(1, 0, 0); b->pd[1] = -h.x;
	b->pn[2] = V(0, 0, 1); b->pd[2] = -h.z;
	b->pn[3] = V(-1, 0, 0); b->pd[3] = -h.x;
	b->pn[4] = V(0, 1, 0); b->pd[4] = -h.y;
	b->pn[5] = V(0, -1, 0); b->pd[5] = -h.y;
	polyd* d = &b->d;
	d->nv = 8; d->np = 6;
	d->ne = build_edges(6, b->pf, b->pc, b->vr, b->ev, b->ef);
	d->v = b->v; d->pn = b->pn; d->pd = b->pd; d->pf = b->pf; d->pc = b->pc; d->vr = b->vr; d->ev = b->ev;
	d->ef = b->ef;
	d->bc = V(0, 0, 0);
	d->be = h;
	d->margin = fminf(h.x, fminf(h.y, h.z)) * 0.15f;
}

static inline v3 mxtransinv(mxf t, v3 v) { return mtmul(t.r, sub(v, t.p)); }
static inline mxf mxf_of(xf t) { mxf m; m.r = qmat(t.q); m.p = t.p; return m; }

/* point of a polygon's plane inside its loop (counter-clockwise about n) */
static int in_polygon(const polyd* P, int f, v3 q)
{
	const v3 n = P->pn[f];
	for (int k = 0; k < P->pc[f]; k++)
	{
		const v3 a = P->v[P->vr[P->pf[f] + k]];
		const v3 b = P->v[P->vr[P->pf[f] + (k + 1) % P->pc[f]]];
		if (dot(sub(q, a), cross(sub(b, a), n)) > 0.0f)
			return 0;
	}
	return 1;
}

/* closest points of segments p1q1, p2q2 (Ericson, Real-Time Collision
 * Detection, 5.1.9) */
static float seg_closest2(v3 p1, v3 q1, v3 p2, v3 q2, v3* c1, v3* c2)
{
	const v3 d1 = sub(q1, p1), d2 = sub(q2, p2), r = sub(p1, p2);
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
	*c1 = add(p1, scl(d1, s));
	*c2 = add(p2, scl(d2, t));
	const v3 w = sub(*c1, *c2);
	return dot(w, w);
}

/* Separation of polytope A (vertices va, in B's frame; planes of A mapped by
 * aToB) and polytope B. A face axis separating them by more than cd ends the
 * search (no contact is possible; dist is that separation). Separated: the exact closest points (what GJK
 * converges to on polytopes), returns 1 with dist > 0 and n = (cA - cB)/dist.
 * Overlapping: the minimum translation over the face normals of both and all
 * edge-pair directions (what EPA converges to), returns 0 with dist = -depth
 * and n the direction that separates A from B. */
static int poly_separation(const polyd* A, mxf aToB, const v3* va, const polyd* B, float cd, float* dist, v3* n,
                           v3* cA, v3* cB)
{
	/* overlap along every candidate axis (normal from B towards A) */
	float best = 3.402823466e+38f;
	int sep = 0;
	for (int f = 0; f < B->np && !sep; f++)
	{
		float mn = 3.402823466e+38f;
		int vi = 0;
		for (int i = 0; i < A->nv; i++)
		{
			const float h = dot(B->pn[f], va[i]) + B->pd[f];
			if (h < mn) { mn = h; vi = i; }
		}
		if (mn > cd)
		{
			*dist = mn; /* farther apart than the contact distance along this axis */
			return 1;
		}
		if (mn > 0.0f) sep = 1;
		else if (-mn < best)
		{
			best = -mn;
			*n = B->pn[f];
			*cA = va[vi];
			*cB = sub(va[vi], scl(B->pn[f], mn));
		}
	}
	for (int f = 0; f < A->np && !sep; f++)
	{
		const v3 an = mmul(aToB.r, A->pn[f]);
		const float ad = A->pd[f] - dot(an, aToB.p);
		float mn = 3.402823466e+38f;
		int vi = 0;
		for (int i = 0; i < B->nv; i++)
		{
			const float h = dot(an, B->v[i]) + ad;
			if (h < mn) { mn = h; vi = i; }
		}
		if (mn > cd)
		{
			*dist = mn;
			return 1;
		}
		if (mn > 0.0f) sep = 1;
		else if (-mn < best)
		{
			best = -mn;
			*n = neg(an);
			*cB = B->v[vi];
			*cA = sub(B->v[vi], scl(an, mn));
		}
	}
	/* edge pairs that form a face of the Minkowski difference: their arcs on
	 * the Gauss maps of A and -B cross (Gregorius, "The Separating Axis Test
	 * between Convex Polyhedra", GDC 2013) */
	v3 bmean = V(0, 0, 0);
	for (int i = 0; i < B->nv; i++)
		bmean = add(bmean, B->v[i]);
	bmean = scl(bmean, 1.0f / (float)B->nv);
	for (int ea = 0; ea < A->ne && !sep; ea++)
	{
		const v3 a0 = va[A->ev[2 * ea]], a1 = va[A->ev[2 * ea + 1]];
		const v3 na = mmul(aToB.r, A->pn[A->ef[2 * ea]]), nb = mmul(aToB.r, A->pn[A->ef[2 * ea + 1]]);
		const v3 bxa = cross(nb, na);
		for (int eb = 0; eb < B->ne && !sep; eb++)
		{
			const v3 nc_ = neg(B->pn[B->ef[2 * eb]]), nd = neg(B->pn[B->ef[2 * eb + 1]]);
			const v3 dxc = cross(nd, nc_);
			const float cba = dot(nc_, bxa), dba = dot(nd, bxa), adc = dot(na, dxc), bdc = dot(nb, dxc);
			if (!(cba * dba < 0.0f && adc * bdc < 0.0f && cba * bdc > 0.0f))
				continue;
			const v3 b0 = B->v[B->ev[2 * eb]], b1 = B->v[B->ev[2 * eb + 1]];
			v3 ax = cross(sub(a1, a0), sub(b1, b0));
			const float l2 = dot(ax, ax);
			if (l2 < 1e-12f)
				continue;
			ax = scl(ax, 1.0f / sqrtf(l2));
			/* outward from B (its vertex mean lies inside it); on a Minkowski
			 * face the supports along the axis are the two edges */
			if (dot(ax, sub(b0, bmean)) < 0.0f)
				ax = neg(ax);
			const float bmax = dot(ax, b0), amin = dot(ax, a0);
			const float ov = bmax - amin;
			if (ov < 0.0f) sep = 1;
			else if (ov < best)
			{
				v3 c1, c2;
				seg_closest(a0, a1, b0, b1, &c1, &c2);
				best = ov;
				*n = ax;
				*cA = c1;
				*cB = c2;
			}
		}
	}
	if (!sep)
	{
		*dist = -best;
		return 0;
	}
	/* separated: the closest features (vertex-face both ways, edge-edge) */
	float d2 = 3.402823466e+38f;
	for (int f = 0; f < B->np; f++)
		for (int i = 0; i < A->nv; i++)
		{
			const float h = dot(B->pn[f], va[i]) + B->pd[f];
			if (h < 0.0f || h * h >= d2)
				continue;
			const v3 q = sub(va[i], scl(B->pn[f], h));
			if (!in_polygon(B, f, q))
				continue;
			d2 = h * h; *cA = va[i]; *cB = q;
		}
	for (int f = 0; f < A->np; f++)
	{
		const v3 an = mmul(aToB.r, A->pn[f]);
		const float ad = A->pd[f] - dot(an, aToB.p);
		for (int i = 0; i < B->nv; i++)
		{
			const float h = dot(an, B->v[i]) + ad;
			if (h < 0.0f || h * h >= d2)
				continue;
			const v3 q = sub(B->v[i], scl(an, h));
			if (!in_polygon(A, f, mxtransinv(aToB, q)))
				continue;
			d2 = h * h; *cB = B->v[i]; *cA = q;
		}
	}
	for (int ea = 0; ea < A->ne; ea++)
		for (int eb = 0; eb < B->ne; eb++)
		{
			v3 c1, c2;
			const float s = seg_closest2(va[A->ev[2 * ea]], va[A->ev[2 * ea + 1]], B->v[B->ev[2 * eb]],
			                            B->v[B->ev[2 * eb + 1]], &c1, &c2);
			if (s < d2) { d2 = s; *cA = c1; *cB = c2; }
		}
	*dist = sqrtf(d2);
	*n = scl(sub(*cA, *cB), 1.0f / *dist);
	return 1;
}

/* Gu::getWitnessPolygonIndex (geomutils/src/pcm/GuPCMContactGenUtil.cpp:201-271) */
static int witness_polygon(const polyd* P, v3 normal, v3 closest, float tolerance)
{
	float pd[256];
	const float eps = -tolerance;
	float dist = dot(P->pn[0], closest) + P->pd[0];
	float minDist = dist >= eps ? fabsf(dist) : 3.402823466e+38f;
	pd[0] = minDist;
	float maxDist = dist;
	int maxIdx = 0, idx = 0;
	for (int i = 1; i < P->np; i++)
	{
		dist = dot(P->pn[i], closest) + P->pd[i];
		pd[i] = dist >= eps ? fabsf(dist) : 3.402823466e+38f;
		if (minDist > pd[i]) { minDist = pd[i]; idx = i; }
		if (dist > maxDist) { maxDist = dist; maxIdx = i; }
	}
	if (minDist == 3.402823466e+38f)
		return maxIdx;
	float bestProj = dot(normalize(P->pn[idx]), normal);
	const int first = idx;
	for (int i = 0; i < P->np; i++)
		if ((tolerance > (pd[i] - minDist)) && (first != i))
		{
			const float proj = dot(normalize(P->pn[i]), normal);
			if (bestProj > proj) { idx = i; bestProj = proj; }
		}
	return idx;
}

/* Gu::findRotationMatrixFromZAxis (geomutils/src/pcm/GuPersistentContactManifold.cpp:289-354):
 * a rotation whose third row is `to` */
static m33 rot_from_z(v3 to)
{
	m33 m;
	const float e = to.z, f = fabsf(e);
	if (0.9999f > f)
	{
		const float vx = -to.y, vy = to.x;
		const float h = 1.0f / (1.0f + e);
		const float hvx = h * vx, hvxy = hvx * vy;
		m.c0 = V(hvx * vx + e, hvxy, vy);
		m.c1 = V(hvxy, h * (vy * vy) + e, -vx);
		m.c2 = V(-vy, vx, e);
		return m;
	}
	const v3 from = V(0, 0, 1), absFrom = V(0, 1, 0);
	const v3 u = sub(absFrom, from), v = sub(absFrom, to);
	const float du = dot(u, u), dv = dot(v, v), duv = dot(u, v);
	const float c1 = -(2.0f / du), c2 = -(2.0f / dv), c3 = c1 * (c2 * duv);
	const v3 c1u = scl(u, c1), c2v = scl(v, c2), c3v = scl(v, c3);
	m.c0 = add(scl(u, c1u.x), add(scl(v, c2v.x), scl(u, c3v.x)));
	m.c0.x += 1.0f;
	m.c1 = add(scl(u, c1u.y), add(scl(v, c2v.y), scl(u, c3v.y)));
	m.c1.y += 1.0f;
	m.c2 = add(scl(u, c1u.z), add(scl(v, c2v.z), scl(u, c3v.z)));
	m.c2.z += 1.0f;
	return m;
}

static inline float tri_area2(v3 a, v3 b, v3 c)
{
	const v3 ca = sub(a, c), cb = sub(b, c);
	return ca.x * cb.y - ca.y * cb.x;
}

/* generatedContacts (geomutils/src/pcm/GuPCMContactGenBoxConvex.cpp:331-530):
 * reference polygon `ref` of P0, incident polygon `inc` of P1, t0To1 maps
 * P0's frame to P1's. Contacts keep la on P1 (P1's frame), lb on P0's
 * reference plane (P0's frame), ln the reference normal (P0's frame). */
static void generated_contacts(const polyd* P0, const polyd* P1, int ref, int inc, mxf t0To1, pcontact* mc, int* nc,
                               float cd)
{
	const float feps = 1.1920928955078125e-7f, fmax_ = 3.402823466e+38f;
	const v3 n = normalize(P0->pn[ref]);
	const m33 rot = rot_from_z(n);
	const int nr = P0->pc[ref], ni = P1->pc[inc];
	v3 p0[64], p1[64];
	float t1[64];
	int pen1[64];
	const v3 sPoint = P1->v[P1->vr[P1->pf[inc]]];
	v3 rMin = V(fmax_, fmax_, fmax_), rMax = neg(rMin);
	for (int i = 0; i < nr; i++)
	{
		p0[i] = mmul(rot, P0->v[P0->vr[P0->pf[ref] + i]]);
		rMin = V(fminf(rMin.x, p0[i].x), fminf(rMin.y, p0[i].y), fminf(rMin.z, p0[i].z));
		rMax = V(fmaxf(rMax.x, p0[i].x), fmaxf(rMax.y, p0[i].y), fmaxf(rMax.z, p0[i].z));
	}
	rMin = sub(rMin, V(feps, feps, feps));
	rMax = add(rMax, V(feps, feps, feps));
	const float d = p0[0].z, rd = d + cd;
	v3 iMin = V(fmax_, fmax_, fmax_), iMax = neg(iMin);
	int inside = 0;
	for (int i = 0; i < ni; i++)
	{
		const v3 vert1 = P1->v[P1->vr[P1->pf[inc] + i]];
		const v3 a = mxtransinv(t0To1, vert1);
		v3 q = mmul(rot, a);
		const float z = q.z;
		t1[i] = z - d;
		q.z = d;
		p1[i] = q;
		iMin = V(fminf(iMin.x, q.x), fminf(iMin.y, q.y), fminf(iMin.z, q.z));
		iMax = V(fmaxf(iMax.x, q.x), fmaxf(iMax.y, q.y), fmaxf(iMax.z, q.z));
		if (rd > z)
		{
			pen1[i] = 1;
			if (poly_contains(p0, nr, q, rMin, rMax))
			{
				inside++;
				if (*nc == 64)
					return;
				pcontact* c = &mc[(*nc)++];
				c->la = vert1;
				c->lb = mtmul(rot, q);
				c->ln = n;
				c->pen = t1[i];
			}
		}
		else
			pen1[i] = 0;
	}
	if (inside == ni)
		return;
	inside = 0;
	iMin = sub(iMin, V(feps, feps, feps));
	iMax = add(iMax, V(feps, feps, feps));
	const v3 incN = normalize(P1->pn[inc]);
	const v3 nIn1 = mmul(t0To1.r, n);
	for (int i = 0; i < nr; i++)
	{
		if (!poly_contains(p1, ni, p0[i], iMin, iMax))
			continue;
		const v3 vert0 = mtmul(rot, p0[i]);
		const v3 a = mxtrans(t0To1, vert0);
		const float nom = dot(incN, sub(sPoint, a));
		const float den = dot(incN, nIn1);
		const float t = nom / den;
		if (t > cd)
			continue;
		inside++;
		if (*nc == 64)
			return;
		pcontact* c = &mc[(*nc)++];
		c->la = add(scl(nIn1, t), a);
		c->lb = vert0;
		c->ln = n;
		c->pen = t;
	}
	if (inside == nr)
		return;
	for (int iS = 0, iE = ni - 1; iS < ni; iE = iS++)
	{
		if (!pen1[iS] && !pen1[iE])
			continue;
		const v3 ipA = p1[iS], ipB = p1[iE];
		const v3 ipAO = V(ipA.x, ipA.y, t1[iS] + d), ipBO = V(ipB.x, ipB.y, t1[iE] + d);
		const v3 iMn = V(fminf(ipA.x, ipB.x), fminf(ipA.y, ipB.y), 0), iMx = V(fmaxf(ipA.x, ipB.x), fmaxf(ipA.y, ipB.y), 0);
		for (int rS = 0, rE = nr - 1; rS < nr; rE = rS++)
		{
			const v3 rpA = p0[rS], rpB = p0[rE];
			const v3 rMn = V(fminf(rpA.x, rpB.x), fminf(rpA.y, rpB.y), 0),
			         rMx = V(fmaxf(rpA.x, rpB.x), fmaxf(rpA.y, rpB.y), 0);
			if (iMn.x > rMx.x || iMn.y > rMx.y || rMn.x > iMx.x || rMn.y > iMx.y)
				continue;
			const float a1 = tri_area2(rpA, rpB, ipA), a2 = tri_area2(rpA, rpB, ipB);
			if (!(0.0f > a1 * a2))
				continue;
			const float a3 = tri_area2(ipA, ipB, rpA), a4 = tri_area2(ipA, ipB, rpB);
			if (!(0.0f > a3 * a4))
				continue;
			const float t = a1 / (a2 - a1);
			const v3 pBB = sub(ipAO, scl(sub(ipBO, ipAO), t));
			const v3 pAA = V(pBB.x, pBB.y, d);
			const v3 pA = mtmul(rot, pAA);
			const v3 pB = mxtrans(t0To1, mtmul(rot, pBB));
			const float pen = pBB.z - pAA.z;
			if (pen > cd)
				continue;
			if (*nc == 64)
				return;
			pcontact* c = &mc[(*nc)++];
			c->la = pB;
			c->lb = pA;
			c->ln = n;
			c->pen = pen;
		}
	}
}

/* manifold -> contact buffer: points on B (shape1) in the world, the summed
 * normal, only points within the contact distance
 * (PersistentContactManifold::addManifoldContactsToContactBuffer and
 * getWorldNormal, geomutils/src/pcm/GuPersistentContactManifold.cpp:739-762,
 * GuPersistentContactManifold.h:405-416) */
static int manifold_out(const pcontact* m, int n, xf x1, float cd, cset* o)
{
	if (n == 0)
		return 0;
	v3 s = m[0].ln;
	for (int i = 1; i < n; i++)
		s = add(s, m[i].ln);
	o->n = normalize(qrot(x1.q, s));
	int k = 0;
	for (int i = 0; i < n && k < 4; i++)
		if (cd >= m[i].pen)
		{
			o->p[k] = xtrans(x1, m[i].lb);
			o->sep[k] = m[i].pen;
			k++;
		}
	o->count = k;
	return k;
}

/* reduction of a batch of manifold points to at most four
 * (PersistentContactManifold::addBatchManifoldContacts, :805-830) */
static int batch_reduce(pcontact* mc, int nc, float tol)
{
	if (nc <= 4)
		return nc;
	pcontact out[4];
	reduce_batch(mc, nc, tol, out);
	for (int i = 0; i < 4; i++)
		mc[i] = out[i];
	return 4;
}

/* Gu::pcmContactConvexConvex / pcmContactBoxConvex on a fresh manifold
 * (geomutils/src/pcm/GuPCMContactConvexConvex.cpp:194-276, :111-157;
 * GuPCMContactBoxConvex.cpp; the non-degenerate branch of
 * Gu::generateFullContactManifold and Gu::addGJKEPAContacts,
 * GuPCMContactGenBoxConvex.cpp:532-720). A is shape0, B is shape1. */
static int c_convex_convex(const polyd* A, xf x0, const polyd* B, xf x1, float cd, float tol, cset* o)
{
	const xf rel = xinvmul(x1, x0);
	const mxf aToB = mxf_of(rel);
	v3 va[256];
	for (int i = 0; i < A->nv; i++)
		va[i] = mxtrans(aToB, A->v[i]);
	float dist = 0.0f;
	v3 n = V(1, 0, 0), cA = V(0, 0, 0), cB = V(0, 0, 0);
	poly_separation(A, aToB, va, B, cd, &dist, &n, &cA, &cB);
	if (dist > cd)
		return 0;
	const float lo = tol * 1e-2f, hi = tol * 5e-2f;
	const float tA = fminf(fmaxf(A->margin, lo), hi), tB = fminf(fmaxf(B->margin, lo), hi);
	const v3 negN = neg(n);
	const v3 nIn0 = mtmul(aToB.r, n);
	const int f1 = witness_polygon(B, negN, cB, tB);
	const int f0 = witness_polygon(A, nIn0, mxtransinv(aToB, cA), tA);
	const v3 refN = normalize(B->pn[f1]), incN = normalize(A->pn[f0]);
	pcontact mc[64];
	int nc = 0;
	if (fabsf(dot(refN, negN)) >= fabsf(dot(incN, nIn0)))
	{
		const mxf bToA = mxf_of(xinvmul(x0, x1));
		generated_contacts(B, A, f1, f0, bToA, mc, &nc, cd);
	}
	else
	{
		generated_contacts(A, B, f0, f1, aToB, mc, &nc, cd);
		const v3 nn = neg(mmul(aToB.r, incN));
		for (int i = 0; i < nc; i++)
		{
			const v3 t = mc[i].lb;
			mc[i].lb = mc[i].la;
			mc[i].la = t;
			mc[i].ln = nn;
		}
	}
	if (nc == 0)
	{
		/* the GJK/EPA point stays the manifold's only point */
		mc[0].la = mxtransinv(aToB, cA);
		mc[0].lb = cB;
		mc[0].ln = n;
		mc[0].pen = dist;
		nc = 1;
	}
	else
		nc = batch_reduce(mc, nc, tol);
	return manifold_out(mc, nc, x1, cd, o);
}

/* Gu::pcmContactSphereConvex on a fresh manifold
 * (geomutils/src/pcm/GuPCMContactSphereConvex.cpp:110-246): sphere centre
 * against the hull, the closest hull point when outside, the least deep face
 * when inside; one contact on the sphere's surface. */
static int c_sphere_convex(float r, v3 p0, const polyd* B, xf x1, float cd, cset* o)
{
	const v3 c = xtransinv(x1, p0);
	float hmax = -3.402823466e+38f;
	int fmax_ = 0;
	for (int f = 0; f < B->np; f++)
	{
		const float h = dot(B->pn[f], c) + B->pd[f];
		if (h > hmax) { hmax = h; fmax_ = f; }
	}
	v3 n;
	float pen;
	if (hmax <= 0.0f)
	{
		n = B->pn[fmax_];
		pen = hmax - r;
	}
	else
	{
		float d2 = 3.402823466e+38f;
		v3 q = c;
		for (int f = 0; f < B->np; f++)
		{
			const float h = dot(B->pn[f], c) + B->pd[f];
			if (h < 0.0f || h * h >= d2)
				continue;
			const v3 pq = sub(c, scl(B->pn[f], h));
			if (in_polygon(B, f, pq)) { d2 = h * h; q = pq; }
		}
		for (int e = 0; e < B->ne; e++)
		{
			v3 c1, c2;
			const float s = seg_closest2(c, c, B->v[B->ev[2 * e]], B->v[B->ev[2 * e + 1]], &c1, &c2);
			if (s < d2) { d2 = s; q = c2; }
		}
		const float dist = sqrtf(d2);
		if (dist > r + cd)
			return 0;
		n = scl(sub(c, q), 1.0f / dist);
		pen = dist - r;
	}
	const v3 wn = qrot(x1.q, n);
	o->n = wn;
	o->p[0] = sub(p0, scl(wn, r));
	o->sep[0] = pen;
	o->count = 1;
	return 1;
}

/* Capsule core segment pq against polytope B (B's frame), standing in for
 * GJK / EPA on the segment (geomutils/src/gjk/GuGJKPenetration.h:95-306,
 * takeCoreShape: points on the segment, distance without the radius).
 * Apart: the exact closest points, returns dist > 0 with n = (cA - cB)/dist.
 * Entering B: the minimum translation over B's face normals and the
 * directions normal to the segment and each edge of B (the face normals of
 * their Minkowski difference), returns -depth with n the direction that moves
 * the segment out, cB = cA + n depth on B's surface. */
static float seg_poly(const polyd* B, v3 p, v3 q, v3* n, v3* cA, v3* cB)
{
	const float fmax_ = 3.402823466e+38f;
	const v3 d = sub(q, p);
	float t0 = 0.0f, t1 = 1.0f;
	int enters = 1;
	*cA = p;
	*cB = p;
	for (int f = 0; f < B->np && enters; f++)
	{
		const float hp = dot(B->pn[f], p) + B->pd[f], dn = dot(B->pn[f], d);
		if (dn == 0.0f)
		{
			if (hp > 0.0f)
				enters = 0;
			continue;
		}
		const float t = -hp / dn;
		if (dn < 0.0f)
			t0 = fmaxf(t0, t);
		else
			t1 = fminf(t1, t);
		if (t0 > t1)
			enters = 0;
	}
	if (!enters)
	{
		float best = fmax_;
		for (int k = 0; k < 2; k++)
		{
			const v3 c = k ? q : p;
			for (int f = 0; f < B->np; f++)
			{
				const float h = dot(B->pn[f], c) + B->pd[f];
				if (h < 0.0f || h * h >= best)
					continue;
				const v3 pc = sub(c, scl(B->pn[f], h));
				if (in_polygon(B, f, pc)) { best = h * h; *cA = c; *cB = pc; }
			}
		}
		for (int e = 0; e < B->ne; e++)
		{
			v3 c1, c2;
			const float s = seg_closest2(p, q, B->v[B->ev[2 * e]], B->v[B->ev[2 * e + 1]], &c1, &c2);
			if (s < best) { best = s; *cA = c1; *cB = c2; }
		}
		const float dist = sqrtf(best);
		*n = scl(sub(*cA, *cB), 1.0f / dist);
		return dist;
	}
	float best = fmax_;
	int bestEdge = -1;
	v3 axis = V(1.0f, 0.0f, 0.0f);
	for (int f = 0; f < B->np; f++)
	{
		const float depth = -B->pd[f] - fminf(dot(B->pn[f], p), dot(B->pn[f], q));
		if (depth < best) { best = depth; axis = B->pn[f]; }
	}
	for (int e = 0; e < B->ne; e++)
	{
		const v3 a = B->v[B->ev[2 * e]], b = B->v[B->ev[2 * e + 1]];
		v3 c = cross(d, sub(b, a));
		const float l2 = dot(c, c);
		if (!(l2 > 1e-12f))
			continue;
		c = scl(c, 1.0f / sqrtf(l2));
		float mnB = fmax_, mxB = -fmax_;
		for (int i = 0; i < B->nv; i++)
		{
			const float s = dot(c, B->v[i]);
			mnB = fminf(mnB, s);
			mxB = fmaxf(mxB, s);
		}
		const float sp = dot(c, p), sq = dot(c, q);
		const float dp = mxB - fminf(sp, sq), dm = fmaxf(sp, sq) - mnB;
		if (dp < best) { best = dp; axis = c; bestEdge = e; }
		if (dm < best) { best = dm; axis = neg(c); bestEdge = e; }
	}
	*n = axis;
	if (bestEdge >= 0)
	{
		v3 c1, c2;
		seg_closest2(p, q, B->v[B->ev[2 * bestEdge]], B->v[B->ev[2 * bestEdge + 1]], &c1, &c2);
		*cA = c1;
	}
	else
		*cA = dot(axis, q) < dot(axis, p) ? q : p;
	*cB = add(*cA, scl(axis, best));
	return -best;
}

/* generatedCapsuleBoxFaceContacts (geomutils/src/pcm/GuPCMContactGenSphereCapsule.cpp:221-284):
 * each endpoint moved along -n onto the reference polygon's plane, kept when
 * that distance is within radius + contact distance and the foot lies in the
 * polygon */
static void capsule_face_contacts(const polyd* B, int f, v3 p0, v3 p1, v3 n, float infl, mxf aToB, pcontact* mc,
                                  int* nc)
{
	const float fmax_ = 3.402823466e+38f;
	const v3 pn = normalize(B->pn[f]);
	const int nv = B->pc[f];
	const v3 a = B->v[B->vr[B->pf[f]]];
	const float den0 = dot(pn, sub(p0, a)), den1 = dot(pn, sub(p1, a));
	const float proj = dot(pn, n);
	const float numer = proj > 0.0f ? 1.0f / proj : 0.0f;
	const float t[2] = {den0 * numer, den1 * numer};
	if (!(infl >= t[0] || infl >= t[1]))
		return;
	const m33 rot = rot_from_z(pn);
	v3 pts[256];
	v3 mn = V(fmax_, fmax_, fmax_), mx = neg(mn);
	for (int i = 0; i < nv; i++)
	{
		pts[i] = mmul(rot, B->v[B->vr[B->pf[f] + i]]);
		mn = V(fminf(mn.x, pts[i].x), fminf(mn.y, pts[i].y), fminf(mn.z, pts[i].z));
		mx = V(fmaxf(mx.x, pts[i].x), fmaxf(mx.y, pts[i].y), fmaxf(mx.z, pts[i].z));
	}
	for (int k = 0; k < 2; k++)
	{
		if (!(infl >= t[k]))
			continue;
		const v3 e = k ? p1 : p0;
		const v3 foot = sub(e, scl(n, t[k]));
		if (poly_contains(pts, nv, mmul(rot, foot), mn, mx))
		{
			mc[*nc].la = mxtransinv(aToB, e);
			mc[*nc].lb = foot;
			mc[*nc].ln = n;
			mc[*nc].pen = t[k];
			(*nc)++;
		}
	}
}

/* intersectRayPolyhedron (GuPCMContactGenSphereCapsule.cpp:98-141): the entry
 * parameter of the ray a + t dir, t >= 0, into B */
static int ray_polyhedron(const polyd* B, v3 a, v3 dir, float* tEnter)
{
	float tFirst = 0.0f, tLast = 3.402823466e+38f;
	for (int k = 0; k < B->np; k++)
	{
		const float den = dot(B->pn[k], dir);
		const float dist = dot(B->pn[k], a) + B->pd[k];
		if (1e-7f > fabsf(den))
		{
			if (dist > 0.0f)
				return 0;
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
			return 0;
	}
	*tEnter = tFirst;
	return 1;
}

/* generatedFaceContacts (GuPCMContactGenSphereCapsule.cpp:286-310): each
 * endpoint's ray along -n into the hull, entered within radius + contact
 * distance */
static void capsule_ray_contacts(const polyd* B, v3 p0, v3 p1, v3 n, float infl, mxf aToB, pcontact* mc, int* nc)
{
	const v3 dir = neg(n);
	for (int k = 0; k < 2; k++)
	{
		const v3 e = k ? p1 : p0;
		float tE;
		if (ray_polyhedron(B, e, dir, &tE) && infl >= tE)
		{
			mc[*nc].la = mxtransinv(aToB, e);
			mc[*nc].lb = add(scl(dir, tE), e);
			mc[*nc].ln = n;
			mc[*nc].pen = tE;
			(*nc)++;
		}
	}
}

/* generateEE / generatedContactsEEContacts (GuPCMContactGenSphereCapsule.cpp:312-377):
 * where the segment crosses the plane through each reference polygon edge
 * spanned with n, a contact against that edge (its parameter within
 * [-0.005, 1.005]) when the signed distance along n is within the radius */
static void capsule_ee_contacts(const polyd* B, int f, v3 p, v3 q, v3 normal, float infl, mxf aToB, pcontact* mc,
                                int* nc)
{
	const int nv = B->pc[f];
	for (int rs = 0, re = nv - 1; rs < nv; re = rs++)
	{
		const v3 a = B->v[B->vr[B->pf[f] + rs]], b = B->v[B->vr[B->pf[f] + re]];
		const v3 ab = sub(b, a);
		const v3 n = cross(ab, normal);
		const float d = dot(n, a);
		const float sp = dot(n, p) - d, sq = dot(n, q) - d;
		if (sp * sq > 0.0f)
			continue;
		const v3 pq = sub(q, p);
		const float npq = dot(n, pq);
		if (npq == 0.0f)
			continue;
		const float segT = (d - dot(n, p)) / npq;
		const v3 lpa = add(scl(pq, segT), p);
		const v3 perN = cross(normal, pq);
		const v3 ap = sub(lpa, a);
		const float tv = dot(perN, ap) / dot(perN, ab);
		if (tv > 1.005f || -0.005f > tv)
			continue;
		const v3 v = sub(ap, scl(ab, tv));
		const float sd = dot(v, normal);
		if (infl >= sd)
		{
			mc[*nc].la = mxtransinv(aToB, lpa);
			mc[*nc].lb = sub(lpa, v);
			mc[*nc].ln = normal;
			mc[*nc].pen = sd;
			(*nc)++;
		}
	}
}

/* PersistentContactManifold::reduceBatchContacts2
 * (geomutils/src/pcm/GuPersistentContactManifold.cpp:1177-1243): the deepest
 * point, the point farthest from it, then the last point (in order) of those
 * nearer the second than the first that is shallower than the second */
static void reduce_batch2(const pcontact* m, int n, pcontact* out)
{
	int idx = 0;
	float maxD = m[0].pen;
	for (int i = 1; i < n; i++)
		if (maxD > m[i].pen) { maxD = m[i].pen; idx = i; }
	out[0] = m[idx];
	int chosen[300] = {0};
	chosen[idx] = 1;
	v3 v = sub(m[0].lb, out[0].lb);
	float far = dot(v, v);
	idx = 0;
	for (int i = 1; i < n; i++)
	{
		v = sub(m[i].lb, out[0].lb);
		const float d = dot(v, v);
		if (d > far) { far = d; idx = i; }
	}
	out[1] = m[idx];
	chosen[idx] = 1;
	int second = idx;
	const float maxDepth = m[idx].pen;
	for (int i = 0; i < n; i++)
	{
		if (chosen[i])
			continue;
		const v3 d0 = sub(out[0].lb, m[i].lb), d1 = sub(out[1].lb, m[i].lb);
		if (dot(d0, d0) > dot(d1, d1) && maxDepth > m[i].pen)
			second = i;
	}
	if (second != idx)
		out[1] = m[second];
}

/* Gu::pcmContactCapsuleBox / pcmContactCapsuleConvex on a fresh manifold
 * (geomutils/src/pcm/GuPCMContactCapsuleBox.cpp:42-200,
 * GuPCMContactCapsuleConvex.cpp:43-236, full contact generation
 * GuPCMContactGenSphereCapsule.cpp:379-468): the core segment against the
 * polytope; box: endpoint contacts on the witness polygon, convex: endpoint
 * rays into the hull; fewer than two: edge contacts on the witness polygon;
 * none: the closest-point contact; reduced to two. Points on the capsule's
 * surface, separation less the radius, normal from the polytope. */
static int c_capsule_poly(float r, float hh, xf x0, const polyd* B, int isBox, xf x1, float cd, float tol, cset* o)
{
	const mxf aToB = mxf_of(xinvmul(x1, x0));
	const v3 ext = mmul(aToB.r, V(hh, 0.0f, 0.0f));
	const v3 p0 = add(aToB.p, ext), p1 = sub(aToB.p, ext);
	v3 n, cA, cB;
	const float dist = seg_poly(B, p0, p1, &n, &cA, &cB);
	if (dist > r + cd)
		return 0;
	const float infl = r + cd;
	const float wtol = fminf(fmaxf(B->margin, tol * 1e-2f), tol * 5e-2f);
	pcontact mc[300];
	int nc = 0;
	if (isBox)
	{
		const int f = witness_polygon(B, neg(n), cB, wtol);
		capsule_face_contacts(B, f, p0, p1, n, infl, aToB, mc, &nc);
		if (nc < 2)
			capsule_ee_contacts(B, f, p0, p1, n, infl, aToB, mc, &nc);
	}
	else
	{
		capsule_ray_contacts(B, p0, p1, n, infl, aToB, mc, &nc);
		if (nc < 2)
			capsule_ee_contacts(B, witness_polygon(B, neg(n), cB, wtol), p0, p1, n, infl, aToB, mc, &nc);
	}
	pcontact red[2];
	const pcontact* use = mc;
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
		reduce_batch2(mc, nc, red);
		use = red;
		nc = 2;
	}
	const v3 wn = qrot(x1.q, n);
	o->n = wn;
	o->count = 0;
	for (int i = 0; i < nc; i++)
	{
		const float sep = use[i].pen - r;
		if (cd >= sep)
		{
			o->p[o->count] = sub(xtrans(x0, use[i].la), scl(wn, r));
			o->sep[o->count] = sep;
			o->count++;
		}
	}
	return o->count;
}

/* Gu::pcmContactPlaneConvex on a fresh manifold
 * (geomutils/src/pcm/GuPCMContactPlaneConvex.cpp:36-227): the hull face most
 * opposed to the plane normal, or the better face of the edge whose summed
 * face normals are most opposed to it (with a 5e-4 bias towards edges); the
 * vertices of the chosen face below the contact distance, taken once for the
 * face and once more when an edge was chosen (the reference loops twice over
 * the same face); reduced to four. Normal from the hull towards the plane. */
static int c_plane_convex(v3 pn, float pd, const polyd* B, xf tb, float cd, float tol, cset* o)
{
	const v3 n = normalize(qrotinv(tb.q, pn));
	const v3 nn = neg(n);
	float minProj = 3.402823466e+38f;
	int face = 0;
	for (int i = 0; i < B->np; i++)
	{
		const float proj = dot(n, B->pn[i]);
		if (minProj > proj) { minProj = proj; face = i; }
	}
	int edge = -1;
	minProj = minProj - 5e-4f;
	float maxDpSq = minProj * minProj;
	for (int e = 0; e < B->ne; e++)
	{
		const v3 en = add(B->pn[B->ef[2 * e]], B->pn[B->ef[2 * e + 1]]);
		const float enMagSq = dot(en, en);
		const float dp = dot(en, nn);
		const float sqDp = dp * dp;
		if (dp >= 0.0f && sqDp > maxDpSq * enMagSq)
		{
			maxDpSq = sqDp / enMagSq;
			edge = e;
		}
	}
	int passes = 1;
	if (edge >= 0)
	{
		const int f0 = B->ef[2 * edge], f1 = B->ef[2 * edge + 1];
		face = dot(B->pn[f0], nn) > dot(B->pn[f1], nn) ? f0 : f1;
		passes = 2;
	}
	pcontact mc[64];
	int nc = 0;
	int deepestTaken = 0;
	float deepest = 3.402823466e+38f;
	int dv = 0;
	for (int i = 0; i < B->nv; i++)
	{
		const float s = dot(pn, xtrans(tb, B->v[i])) + pd;
		if (s < deepest) { deepest = s; dv = i; }
	}
	for (int pass = 0; pass < passes; pass++)
		for (int k = 0; k < B->pc[face]; k++)
		{
			const int vi = B->vr[B->pf[face] + k];
			const v3 pw = xtrans(tb, B->v[vi]);
			const float s = dot(pn, pw) + pd;
			if (cd > s && nc < 64)
			{
				mc[nc].la = B->v[vi];
				mc[nc].lb = sub(pw, scl(pn, s));
				mc[nc].ln = pn;
				mc[nc].pen = s;
				nc++;
				deepestTaken |= vi == dv;
			}
		}
	/* Deliberate deviation: when the chosen face does not hold the hull's
	 * deepest vertex (an irregular hull can rest on a vertex that the face
	 * most opposed to the plane does not contain), that vertex joins the
	 * batch. The reference scene's persistent manifold keeps such a vertex
	 * from earlier steps; a step that generates contacts afresh has to add it. */
	if (!deepestTaken && cd > deepest && nc < 64)
	{
		const v3 pw = xtrans(tb, B->v[dv]);
		mc[nc].la = B->v[dv];
		mc[nc].lb = sub(pw, scl(pn, deepest));
		mc[nc].ln = pn;
		mc[nc].pen = deepest;
		nc++;
	}
	nc = batch_reduce(mc, nc, tol);
	o->n = neg(pn);
	int k = 0;
	for (int i = 0; i < nc; i++)
		if (cd >= mc[i].pen)
		{
			o->p[k] = mc[i].lb;
			o->sep[k] = mc[i].pen;
			k++;
		}
	o->count = k;
	return k;
}

static xf load_pose(const float* p)
{
	xf t;
	t.p = V(p[0], p[1], p[2]);
	t.q.x = p[3]; t.q.y = p[4]; t.q.z = p[5]; t.q.w = p[6];
	return t;
}

/* plane shape pose -> (n, d): the plane is the local x=0 plane of its pose
 * (PxPlaneGeometry, physx/include/geometry/PxPlaneGeometry.h) */
static void plane_from_pose(xf t, v3* n, float* d)
{
	*n = qrot(t.q, V(1.0f, 0.0f, 0.0f));
	*d = -dot(*n, t.p);
}

int pxo_generate_contacts(int type0, const float* g0, const float* pose0, int type1, const float* g1,
                          const float* pose1, float cd, float tol, float* out)
{
	cset c;
	c.count = 0;
	const xf x0 = load_pose(pose0), x1 = load_pose(pose1);
	int n = 0;
	if (type0 == PXO_SPHERE && type1 == PXO_SPHERE)
		n = c_sphere_sphere(g0[0], x0.p, g1[0], x1.p, cd, &c);
	else if (type0 == PXO_SPHERE && type1 == PXO_PLANE)
	{
		v3 pn; float pd;
		plane_from_pose(x1, &pn, &pd);
		n = c_sphere_plane(g0[0], x0.p, pn, pd, cd, &c);
	}
	else if (type0 == PXO_SPHERE && type1 == PXO_BOX)
		n = c_sphere_box(g0[0], x0.p, V(g1[0], g1[1], g1[2]), x1, cd, &c);
	else if (type0 == PXO_PLANE && type1 == PXO_BOX)
	{
		v3 pn; float pd;
		plane_from_pose(x0, &pn, &pd);
		n = c_plane_box(pn, pd, V(g1[0], g1[1], g1[2]), x1, cd, &c);
	}
	else if (type0 == PXO_BOX && type1 == PXO_BOX)
		n = c_box_box(V(g0[0], g0[1], g0[2]), x0, V(g1[0], g1[1], g1[2]), x1, cd, tol, &c);
	else if (type0 == PXO_SPHERE && type1 == PXO_CAPSULE)
		n = c_sphere_capsule(g0[0], x0.p, g1[0], g1[1], x1, cd, &c);
	else if (type0 == PXO_CAPSULE && type1 == PXO_CAPSULE)
		n = c_capsule_capsule(g0[0], g0[1], x0, g1[0], g1[1], x1, cd, &c);
	else if (type0 == PXO_CAPSULE && type1 == PXO_BOX)
	{
		boxpoly bp;
		box_poly(V(g1[0], g1[1], g1[2]), &bp);
		n = c_capsule_poly(g0[0], g0[1], x0, &bp.d, 1, x1, cd, tol, &c);
	}
	else if (type0 == PXO_CAPSULE && type1 == PXO_CONVEX)
		n = c_capsule_poly(g0[0], g0[1], x0, hull_of(g1[0]), 0, x1, cd, tol, &c);
	else if (type0 == PXO_PLANE && type1 == PXO_CAPSULE)
	{
		v3 pn; float pd;
		plane_from_pose(x0, &pn, &pd);
		n = c_plane_capsule(pn, pd, g1[0], g1[1], x1, cd, &c);
	}
	else if (type0 == PXO_PLANE && type1 == PXO_CONVEX)
	{
		v3 pn; float pd;
		plane_from_pose(x0, &pn, &pd);
		n = c_plane_convex(pn, pd, hull_of(g1[0]), x1, cd, tol, &c);
	}
	else if (type0 == PXO_SPHERE && type1 == PXO_CONVEX)
		n = c_sphere_convex(g0[0], x0.p, hull_of(g1[0]), x1, cd, &c);
	else if (type0 == PXO_BOX && type1 == PXO_CONVEX)
	{
		boxpoly bp;
		box_poly(V(g0[0], g0[1], g0[2]), &bp);
		n = c_convex_convex(&bp.d, x0, hull_of(g1[0]), x1, cd, tol, &c);
	}
	else if (type0 == PXO_CONVEX && type1 == PXO_CONVEX)
		n = c_convex_convex(hull_of(g0[0]), x0, hull_of(g1[0]), x1, cd, tol, &c);
	else
		return -1;
	for (int i = 0; i < n; i++)
	{
		float* o = out + 8 * i;
		o[0] = c.p[i].x; o[1] = c.p[i].y; o[2] = c.p[i].z; o[3] = c.sep[i];
		o[4] = c.n.x; o[5] = c.n.y; o[6] = c.n.z; o[7] = 0.0f;
	}
	return n;
}

/* ------------------------------------------------------------------------- */
/* Broadphase: ABP's inclusive overlap of bounds inflated by per-box distances */
/* (lowlevelaabb/src/BpBroadPhaseABP.cpp:1187-1192 inflation, :2246-2262 and  */
/*  :2474 inclusive Y/Z and X tests)                                          */
/* ------------------------------------------------------------------------- */

typedef struct { float minx; int idx; } sortx;
static int cmp_sortx(const void* a, const void* b)
{
	const sortx* x = (const sortx*)a;
	const sortx* y = (const sortx*)b;
	if (x->minx < y->minx) return -1;
	if (x->minx > y->minx) return 1;
	return x->idx - y->idx;
}
static int cmp_pair(const void* a, const void* b)
{
	const uint32_t* x = (const uint32_t*)a;
	const uint32_t* y = (const uint32_t*)b;
	if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
	if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
	return 0;
}

typedef struct { uint32_t* p; long long n, cap; } pairvec;
static void pv_push(pairvec* v, uint32_t a, uint32_t b)
{
	if (v->n == v->cap)
	{
		v->cap = v->cap ? v->cap * 2 : 1024;
		v->p = (uint32_t*)realloc(v->p, sizeof(uint32_t) * 2 * (size_t)v->cap);
	}
	v->p[2 * v->n] = a < b ? a : b;
	v->p[2 * v->n + 1] = a < b ? b : a;
	v->n++;
}

static void bp_core(const float* lo, const float* hi, const int32_t* world, int n, pairvec* out)
{
	sortx* s = (sortx*)malloc(sizeof(sortx) * (size_t)(n ? n : 1));
	for (int i = 0; i < n; i++) { s[i].minx = lo[3 * i]; s[i].idx = i; }
	qsort(s, (size_t)n, sizeof(sortx), cmp_sortx);
	for (int a = 0; a < n; a++)
	{
		const int i = s[a].idx;
		const float maxx = hi[3 * i];
		for (int b = a + 1; b < n && s[b].minx <= maxx; b++)
		{
			const int j = s[b].idx;
			if (world && world[i] != world[j])
				continue;
			if (hi[3 * j + 1] < lo[3 * i + 1] || hi[3 * i + 1] < lo[3 * j + 1] || hi[3 * j + 2] < lo[3 * i + 2] ||
			    hi[3 * i + 2] < lo[3 * j + 2])
				continue;
			pv_push(out, (uint32_t)i, (uint32_t)j);
		}
	}
	free(s);
	if (out->n)
		qsort(out->p, (size_t)out->n, 2 * sizeof(uint32_t), cmp_pair);
}

long long pxo_bp_pairs(const float* bounds, const float* dist, const int32_t* world, int n, uint32_t* out,
                       long long max_pairs)
{
	float* lo = (float*)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1));
	float* hi = (float*)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1));
	for (int i = 0; i < n; i++)
		for (int k = 0; k < 3; k++)
		{
			lo[3 * i + k] = bounds[6 * i + k] - dist[i];
			hi[3 * i + k] = bounds[6 * i + 3 + k] + dist[i];
		}
	pairvec pv = {0, 0, 0};
	bp_core(lo, hi, world, n, &pv);
	for (long long i = 0; i < pv.n && i < max_pairs; i++)
	{
		out[2 * i] = pv.p[2 * i];
		out[2 * i + 1] = pv.p[2 * i + 1];
	}
	free(lo);
	free(hi);
	free(pv.p);
	return pv.n;
}

/* ------------------------------------------------------------------------- */
/* World and TGS step                                                         */
/* ------------------------------------------------------------------------- */

/* Friction patch kept between steps (lowleveldynamics/src/DyFrictionPatch.h):
 * body-space normals and anchors of both bodies, the anchor count and the
 * broken flag the last solve pass wrote. */
typedef struct {
	uint64_t key;        /* pair key: a << 32 | b, planes 0x80000000 | plane */
	v3 n0, n1, a0[2], a1[2];
	int count, broken;
	int color;           /* the constraint's colour in that step */
} fpatch;

typedef struct {
	int a, b;            /* body indices; b < 0: static plane -1-b */
	int color;
	uint64_t key;
	uint64_t pkey;       /* pair key of the friction patch */
	fpatch fp;
	cset c;              /* contacts, normal from b towards a */
	/* prepared rows */
	float invMassA, invMassB;
	v3 raXnI[4], rbXnI[4];
	float sepErr[4], biasCoef[4], velMul[4], recipResp[4], force[4];
	int nFric;
	v3 fn[4], fraXnI[4], frbXnI[4];
	float ferr[4], fvelMul[4], fbiasScale, fscale, fforce[4];
	float maxPenBias;
	/* PGS rows: normal biased/unbiased error, friction bias and target
	 * velocity, friction coefficients scaled for the anchor count */
	float biasedErr[4], unbiasedErr[4], fbias[4], ftarget[4], muS, muD;
} pxo_con;

/* A joint: kind, bodies (-1: the world frame), local frames, colour, and
 * the rows its prep produced this step (SolverConstraint1DHeaderStep and
 * SolverConstraint1DStep, lowleveldynamics/src/DySolverConstraint1DStep.h:
 * 139-255). */
typedef struct {
	v3 lin0, lin1, ang0, ang1;
	float error, biasScale, velMul, velTarget, minImp, maxImp, applied, maxBias;
	int flags;           /* JF_* */
} jrow;

enum { JF_ROT_EQ = 1, JF_ORTHO_TARGET = 2, JF_ANGULAR_ERROR = 4, JF_INEQUALITY = 8, JF_KEEP_BIAS = 16, JF_SPRING = 32 };

/* limit and drive parameters of revolute and prismatic joints: flags JP_*,
 * limit lower / upper (angles about or distances along the joint x axis),
 * the limit's restitution and bounce threshold, drive velocity, force limit
 * and gear ratio, the limit's spring stiffness and damping (a soft limit
 * when either is positive) (PxRevoluteJoint::setLimit / setRevoluteJointFlag /
 * setDriveVelocity / setDriveForceLimit / setDriveGearRatio,
 * PxPrismaticJoint::setLimit / setPrismaticJointFlag) */
enum { JP_LIMIT = 1, JP_DRIVE = 2, JP_FREESPIN = 4, JP_COLLIDE = 8, JP_FORCES = 16 };
#define JOINT_MAX_ROWS 8

typedef struct {
	int type, a, b, color;
	xf f0, f1;
	int pflags;
	float lo, hi, rest, bthr, dvel, dfl, gear, stiff, damp;
	/* prepared each step */
	int count;
	float invMass0, invMass1;
	v3 rA, rB;
	v3 o0[3], o1[3];     /* angOrthoAxis0 / angOrthoAxis1 */
	float orecip[3], oerr[3];
	jrow r[JOINT_MAX_ROWS];
} pxo_joint;

/* a static triangle mesh instance: its world and triangle range; a triangle
 * in world space with its unit normal (counter-clockwise winding), bounds and
 * vertex ids */
typedef struct tmesh_s { int scene, first, count; v3 blo, bhi; } tmesh_t;
typedef struct tri_s { v3 v[3]; v3 n; v3 lo, hi; int vi[3]; } tri_t;

typedef struct {
	pxo_desc d;
	int n, cap;
	int32_t *type, *scene, *local;
	float *dims;
	xf* pose;
	v3 *lv, *av;
	float* invMass;
	v3* invI;
	/* accelerations of forces and torques set for the next step */
	v3 *extLin, *extAng;
	int m, mcap;
	v3* pn;
	float* pd;
	int32_t *pscene, *plocal;
	/* scratch of the last step */
	pxo_con* con;
	long long ncon, concap;
	long long stats[6];
	/* friction patches of the previous step, sorted by pair key */
	fpatch* prev;
	long long nprev;
	/* joints, their colours, and the sorted body-pair keys they keep from
	 * colliding */
	pxo_joint* joint;
	int nj, jcap, njcolors;
	uint64_t* jkeys;
	/* aggregates: each body's aggregate (-1: none, array of n or NULL) and
	 * whether each aggregate's members collide with each other */
	int* agg;
	char* aggSelf;
	int nagg;
	/* sleeping: wake counter, accumulated motion velocities (linear; angular
	 * in the body frame), asleep and ready-for-sleep flags, island label (the
	 * smallest body index of the island the body last belonged to), islands
	 * to wake at the next step (by label), joints per body */
	float* wc;
	v3 *accL, *accA;
	unsigned char *asleep, *ready, *wakeIsl;
	int *isl, *jcount;
	/* static triangle meshes, at most one per world: world-space triangles */
	struct tmesh_s* tm;
	int ntm;
	struct tri_s* tri;
	int ntri, ntmverts;
} pxo_world;

void pxo_default_desc(pxo_desc* d)
{
	memset(d, 0, sizeof(*d));
	d->gravity[1] = -9.81f;
	d->solver = 1;
	d->pos_iters = 4;
	d->vel_iters = 1;
	d->contact_offset = 0.02f;
	d->rest_offset = 0.0f;
	d->static_friction = 0.5f;
	d->dynamic_friction = 0.5f;
	d->restitution = 0.0f;
	d->linear_damping = 0.0f;
	d->angular_damping = 0.05f;
	d->max_lin_vel_sq = 1e32f;
	d->max_ang_vel_sq = 100.0f * 100.0f;
	d->bounce_threshold = 2.0f;
	d->friction_offset_threshold = 0.04f;
	d->correlation_distance = 0.025f;
	d->tolerance_length = 1.0f;
}

void* pxo_world_create(const pxo_desc* d)
{
	pxo_world* w = (pxo_world*)calloc(1, sizeof(pxo_world));
	w->d = *d;
	return w;
}

void pxo_world_release(void* wp)
{
	pxo_world* w = (pxo_world*)wp;
	if (!w) return;
	free(w->type); free(w->scene); free(w->local); free(w->dims); free(w->pose); free(w->lv); free(w->av);
	free(w->invMass); free(w->invI); free(w->extLin); free(w->extAng); free(w->pn); free(w->pd); free(w->pscene);
	free(w->plocal); free(w->con);
	free(w->prev);
	free(w->joint); free(w->jkeys);
	free(w->agg); free(w->aggSelf);
	free(w->wc); free(w->accL); free(w->accA); free(w->asleep); free(w->ready); free(w->wakeIsl); free(w->isl);
	free(w->jcount);
	free(w->tm); free(w->tri);
	free(w);
}

int pxo_world_add_bodies(void* wp, int n, const int32_t* type, const float* dims, const float* pose, const float* lv,
                         const float* av, const float* mass, const float* inertia, const int32_t* scene)
{
	pxo_world* w = (pxo_world*)wp;
	const int nn = w->n + n;
	if (w->agg)
	{
		w->agg = (int*)realloc(w->agg, sizeof(int) * nn);
		for (size_t i = (size_t)w->n; i < nn; i++)
			w->agg[i] = -1;
	}
	w->type = (int32_t*)realloc(w->type, sizeof(int32_t) * nn);
	w->scene = (int32_t*)realloc(w->scene, sizeof(int32_t) * nn);
	w->local = (int32_t*)realloc(w->local, sizeof(int32_t) * nn);
	w->dims = (float*)realloc(w->dims, sizeof(float) * 4 * nn);
	w->pose = (xf*)realloc(w->pose, sizeof(xf) * nn);
	w->lv = (v3*)realloc(w->lv, sizeof(v3) * nn);
	w->av = (v3*)realloc(w->av, sizeof(v3) * nn);
	w->invMass = (float*)realloc(w->invMass, sizeof(float) * nn);
	w->invI = (v3*)realloc(w->invI, sizeof(v3) * nn);
	w->extLin = (v3*)realloc(w->extLin, sizeof(v3) * nn);
	w->extAng = (v3*)realloc(w->extAng, sizeof(v3) * nn);
	w->wc = (float*)realloc(w->wc, sizeof(float) * nn);
	w->accL = (v3*)realloc(w->accL, sizeof(v3) * nn);
	w->accA = (v3*)realloc(w->accA, sizeof(v3) * nn);
	w->asleep = (unsigned char*)realloc(w->asleep, (size_t)nn);
	w->ready = (unsigned char*)realloc(w->ready, (size_t)nn);
	w->wakeIsl = (unsigned char*)realloc(w->wakeIsl, (size_t)nn);
	w->isl = (int*)realloc(w->isl, sizeof(int) * nn);
	w->jcount = (int*)realloc(w->jcount, sizeof(int) * nn);
	for (int k = w->n; k < nn; k++)
	{
		/* a new body is awake with the scene's wake counter reset value
		 * (PxSceneDesc::wakeCounterResetValue, 20 * 0.02 s) */
		w->wc[k] = 20.0f * 0.02f;
		w->accL[k] = V(0, 0, 0);
		w->accA[k] = V(0, 0, 0);
		w->asleep[k] = 0;
		w->ready[k] = 0;
		w->wakeIsl[k] = 0;
		w->isl[k] = k;
		w->jcount[k] = 0;
	}
	/* index of each body within its world: bodies of that world added before it */
	int maxScene = 0;
	for (int j = 0; j < w->n; j++) if (w->scene[j] > maxScene) maxScene = w->scene[j];
	for (int i = 0; i < n; i++) if (scene && scene[i] > maxScene) maxScene = scene[i];
	int* count = (int*)calloc((size_t)maxScene + 1, sizeof(int));
	for (int j = 0; j < w->n; j++) count[w->scene[j]]++;
	for (int i = 0; i < n; i++)
	{
		const int k = w->n + i;
		w->type[k] = type[i];
		w->scene[k] = scene ? scene[i] : 0;
		w->local[k] = count[w->scene[k]]++;
		w->extLin[k] = V(0, 0, 0);
		w->extAng[k] = V(0, 0, 0);
		memcpy(w->dims + 4 * k, dims + 4 * i, 4 * sizeof(float));
		w->pose[k] = load_pose(pose + 7 * i);
		w->lv[k] = V(lv[3 * i], lv[3 * i + 1], lv[3 * i + 2]);
		w->av[k] = V(av[3 * i], av[3 * i + 1], av[3 * i + 2]);
		w->invMass[k] = mass[i] > 0.0f ? 1.0f / mass[i] : 0.0f;
		w->invI[k] = V(inertia[3 * i] > 0.0f ? 1.0f / inertia[3 * i] : 0.0f,
		               inertia[3 * i + 1] > 0.0f ? 1.0f / inertia[3 * i + 1] : 0.0f,
		               inertia[3 * i + 2] > 0.0f ? 1.0f / inertia[3 * i + 2] : 0.0f);
	}
	free(count);
	w->n = nn;
	return nn;
}

int pxo_world_add_planes(void* wp, int m, const float* planes, const int32_t* scene)
{
	pxo_world* w = (pxo_world*)wp;
	const int nm = w->m + m;
	w->pn = (v3*)realloc(w->pn, sizeof(v3) * nm);
	w->pd = (float*)realloc(w->pd, sizeof(float) * nm);
	w->pscene = (int32_t*)realloc(w->pscene, sizeof(int32_t) * nm);
	w->plocal = (int32_t*)realloc(w->plocal, sizeof(int32_t) * nm);
	for (int i = 0; i < m; i++)
	{
		const int k = w->m + i;
		w->pn[k] = V(planes[4 * i], planes[4 * i + 1], planes[4 * i + 2]);
		w->pd[k] = planes[4 * i + 3];
		w->pscene[k] = scene ? scene[i] : 0;
		int loc = 0;
		for (int j = 0; j < k; j++)
			loc += w->pscene[j] == w->pscene[k];
		w->plocal[k] = loc;
	}
	w->m = nm;
	return nm;
}

void pxo_world_get_bodies(void* wp, float* pose, float* lv, float* av)
{
	pxo_world* w = (pxo_world*)wp;
	for (int i = 0; i < w->n; i++)
	{
		if (pose)
		{
			float* o = pose + 7 * i;
			o[0] = w->pose[i].p.x; o[1] = w->pose[i].p.y; o[2] = w->pose[i].p.z;
			o[3] = w->pose[i].q.x; o[4] = w->pose[i].q.y; o[5] = w->pose[i].q.z; o[6] = w->pose[i].q.w;
		}
		if (lv) { lv[3 * i] = w->lv[i].x; lv[3 * i + 1] = w->lv[i].y; lv[3 * i + 2] = w->lv[i].z; }
		if (av) { av[3 * i] = w->av[i].x; av[3 * i + 1] = w->av[i].y; av[3 * i + 2] = w->av[i].z; }
	}
}

/* PxRigidDynamic::wakeUp (NpRigidDynamic.cpp wakeUp -> BodyCore::wakeUp with
 * the scene's wake counter reset value): the body is awake now with a wake
 * counter of 0.4 s and not ready for sleep; the rest of its island wakes at
 * the next step (IslandSim::activateNode, then wakeIslands) */
static void wake_body(pxo_world* w, int i)
{
	w->wc[i] = 20.0f * 0.02f;
	w->ready[i] = 0;
	if (w->asleep[i])
	{
		w->asleep[i] = 0;
		w->wakeIsl[w->isl[i]] = 1;
	}
}

void pxo_world_wake(void* wp, int k, const int32_t* bodies)
{
	pxo_world* w = (pxo_world*)wp;
	for (int j = 0; j < k; j++)
		if (bodies[j] >= 0 && bodies[j] < w->n)
			wake_body(w, bodies[j]);
}

void pxo_world_get_sleep(void* wp, float* wc, int32_t* asleep)
{
	pxo_world* w = (pxo_world*)wp;
	for (int i = 0; i < w->n; i++)
	{
		if (wc) wc[i] = w->wc[i];
		if (asleep) asleep[i] = w->asleep[i];
	}
}

/* Writing poses or velocities wakes the bodies when sleeping is on (the
 * autowake of PxRigidActor::setGlobalPose, PxRigidDynamic::setLinearVelocity
 * and setAngularVelocity). */
void pxo_world_set_bodies(void* wp, const float* pose, const float* lv, const float* av)
{
	pxo_world* w = (pxo_world*)wp;
	for (int i = 0; i < w->n; i++)
	{
		if (pose) w->pose[i] = load_pose(pose + 7 * i);
		if (lv) w->lv[i] = V(lv[3 * i], lv[3 * i + 1], lv[3 * i + 2]);
		if (av) w->av[i] = V(av[3 * i], av[3 * i + 1], av[3 * i + 2]);
		if (w->d.enable_sleeping && (pose || lv || av))
			wake_body(w, i);
	}
}

/* PxRigidDynamic::addForce / addTorque with PxForceMode::eFORCE at the centre
 * of mass for the next step (NpRigidBodyTemplate.h:507-528): force[n*3] and
 * torque[n*3] (either may be NULL) become the accelerations F / m and
 * I_world^-1 tau (R diag(invI) R^T, Cm::transformInertiaTensor, with the
 * current orientation), replacing those set before. */
void pxo_world_set_forces(void* wp, const float* force, const float* torque)
{
	pxo_world* w = (pxo_world*)wp;
	for (int i = 0; i < w->n; i++)
	{
		if (w->d.enable_sleeping &&
		    ((force && (force[3 * i] != 0.0f || force[3 * i + 1] != 0.0f || force[3 * i + 2] != 0.0f)) ||
		     (torque && (torque[3 * i] != 0.0f || torque[3 * i + 1] != 0.0f || torque[3 * i + 2] != 0.0f))))
			wake_body(w, i);   /* addForce / addTorque wake the body (autowake) */
		if (force)
			w->extLin[i] = scl(V(force[3 * i], force[3 * i + 1], force[3 * i + 2]), w->invMass[i]);
		if (torque)
		{
			const m33 R = qmat(w->pose[i].q);
			m33 A;
			A.c0 = scl(R.c0, w->invI[i].x); A.c1 = scl(R.c1, w->invI[i].y); A.c2 = scl(R.c2, w->invI[i].z);
			w->extAng[i] = mmul(mmm(A, mtrans(R)), V(torque[3 * i], torque[3 * i + 1], torque[3 * i + 2]));
		}
	}
}

void pxo_world_stats(void* wp, long long* out)
{
	pxo_world* w = (pxo_world*)wp;
	for (int i = 0; i < 6; i++) out[i] = w->stats[i];
}

long long pxo_world_contacts(void* wp, float* out, long long max)
{
	pxo_world* w = (pxo_world*)wp;
	for (long long i = 0; i < w->ncon && i < max; i++)
	{
		const pxo_con* k = &w->con[i];
		float* r = out + 32 * i;
		memset(r, 0, 32 * sizeof(float));
		const int32_t hdr[3] = {k->a, k->b, k->c.count};
		memcpy(r, hdr, sizeof(hdr));
		r[4] = k->c.n.x; r[5] = k->c.n.y; r[6] = k->c.n.z;
		for (int j = 0; j < k->c.count; j++)
		{
			r[8 + 4 * j] = k->c.p[j].x; r[9 + 4 * j] = k->c.p[j].y; r[10 + 4 * j] = k->c.p[j].z;
			r[11 + 4 * j] = k->c.sep[j];
			r[24 + j] = k->force[j];
		}
	}
	return w->ncon;
}

long long pxo_world_constraints(void* wp, int32_t* out, long long max)
{
	pxo_world* w = (pxo_world*)wp;
	for (long long i = 0; i < w->ncon && i < max; i++)
	{
		out[4 * i] = w->con[i].a;
		out[4 * i + 1] = w->con[i].b;
		out[4 * i + 2] = w->con[i].color;
		out[4 * i + 3] = w->con[i].c.count;
	}
	return w->ncon;
}

/* shape AABB (geomutils/src/GuBounds.cpp:354-395: sphere centre +- r, box
 * basisExtent, physx/include/foundation/PxBounds3.h:301-314) */
static void shape_bounds(int type, const float* dims, xf t, v3* mn, v3* mx)
{
	if (type == PXO_SPHERE)
	{
		const v3 e = V(dims[0], dims[0], dims[0]);
		*mn = sub(t.p, e);
		*mx = add(t.p, e);
		return;
	}
	const m33 R = qmat(t.q);
	if (type == PXO_CAPSULE)
	{
		/* computeCapsuleBounds (geomutils/include/GuBounds.h:54-62) */
		const v3 w = V(fabsf(R.c0.x) * dims[1] + dims[0], fabsf(R.c0.y) * dims[1] + dims[0],
		               fabsf(R.c0.z) * dims[1] + dims[0]);
		*mn = sub(t.p, w);
		*mx = add(t.p, w);
		return;
	}
	v3 e = V(dims[0], dims[1], dims[2]), c = t.p;
	if (type == PXO_CONVEX)
	{
		/* hull bounds: the local box rotated (GuBounds.cpp:136-140 transformNoEmptyTest) */
		const polyd* h = hull_of(dims[0]);
		e = h->be;
		c = add(mmul(R, h->bc), t.p);
	}
	const v3 c0 = scl(R.c0, e.x), c1 = scl(R.c1, e.y), c2 = scl(R.c2, e.z);
	const v3 w = V(fabsf(c0.x) + fabsf(c1.x) + fabsf(c2.x), fabsf(c0.y) + fabsf(c1.y) + fabsf(c2.y),
	               fabsf(c0.z) + fabsf(c1.z) + fabsf(c2.z));
	*mn = sub(c, w);
	*mx = add(c, w);
}

/* Constraint ordering key: a 64-bit mix of the scene-local identities of the
 * two bodies (or body and plane), so the colouring of a scene does not depend
 * on other scenes batched with it. */
static uint64_t mix64(uint64_t z)
{
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}
static uint64_t con_key(uint32_t la, uint32_t lb) { return mix64(((uint64_t)la << 32) | lb); }

static void push_con(pxo_world* w, int a, int b, const cset* c)
{
	if (w->ncon == w->concap)
	{
		w->concap = w->concap ? 2 * w->concap : 1024;
		w->con = (pxo_con*)realloc(w->con, sizeof(pxo_con) * (size_t)w->concap);
	}
	pxo_con* k = &w->con[w->ncon++];
	memset(k, 0, sizeof(*k));
	k->a = a;
	k->b = b;
	k->c = *c;
	const uint32_t la = (uint32_t)w->local[a];
	/* statics: a plane's world-local index, or 0x400000 + patch for a mesh patch */
	const uint32_t lb = b >= 0 ? (uint32_t)w->local[b]
	                           : (0x80000000u | (-1 - b < w->m ? (uint32_t)w->plocal[-1 - b] : 0x400000u + (uint32_t)(-1 - b - w->m)));
	k->key = con_key(la, lb) ^ ((uint64_t)w->d.order_salt * 0x9E3779B97F4A7C15ull);
	/* world-relative pair identity: world 16 | local a 24 | local b or 0x800000 | plane 24 */
	k->pkey = ((uint64_t)((uint32_t)w->scene[a] & 0xFFFFu) << 48) | ((uint64_t)(la & 0xFFFFFFu) << 24) |
	          ((lb & 0x80000000u) ? (0x800000u | (lb & 0x7FFFFFu)) : (lb & 0x7FFFFFu));
}

static int cmp_patch(const void* x, const void* y)
{
	const uint64_t a = ((const fpatch*)x)->key, b = ((const fpatch*)y)->key;
	return a < b ? -1 : (a > b ? 1 : 0);
}

static const fpatch* find_patch(const pxo_world* w, uint64_t key)
{
	long long lo = 0, hi = w->nprev;
	while (lo < hi)
	{
		const long long mid = (lo + hi) / 2;
		if (w->prev[mid].key < key) lo = mid + 1;
		else hi = mid;
	}
	return lo < w->nprev && w->prev[lo].key == key ? &w->prev[lo] : NULL;
}

/* Solver body (PxTGSSolverBodyVel + PxTGSSolverBodyTxInertia + data),
 * lowleveldynamics/src/DyTGSDynamics.cpp:154-242 copyToSolverBodyDataStep */
typedef struct {
	v3 lin, ang, dlin, dang;   /* linearVelocity, angularVelocity (sqrt-inertia space), deltaLinDt, deltaAngDt */
	v3 p;                      /* body2WorldP */
	qt dq;                     /* deltaBody2WorldQ */
	m33 sqrtInvI;
	v3 origLin, origAng;
	float invMass;
} sbody;

/* integrateCoreStep (lowleveldynamics/src/DyTGSDynamics.cpp:1403-1476) */
static void integrate_step(sbody* b, float dt)
{
	const v3 delta = scl(b->lin, dt);
	const v3 unm = b->ang;
	const v3 amv = mmul(b->sqrtInvI, b->ang);
	const float w2 = dot(amv, amv);
	b->p = add(b->p, delta);
	if (w2 != 0.0f)
	{
		const float wl = sqrtf(w2);
		const float v = dt * wl * 0.5f;
		float s = sinf(v);
		const float q = cosf(v);
		s /= wl;
		const v3 pqr = scl(amv, s);
		const qt qv = {pqr.x, pqr.y, pqr.z, 0.0f};
		qt res = qmul(qv, b->dq);
		res.x += b->dq.x * q; res.y += b->dq.y * q; res.z += b->dq.z * q; res.w += b->dq.w * q;
		b->dq = qnormalize(res);
	}
	b->dang = add(b->dang, scl(unm, dt));
	b->dlin = add(b->dlin, delta);
}

/* Friction patch of a constraint; returns its anchor count. The previous
 * step's patch of this pair is kept when it is unbroken, has anchors, its body
 * normals still agree, no anchor has drifted apart along the normal
 * (getFrictionPatches, isSeparated: DyContactPrepShared.h:53-127) and its world
 * normal matches this step's contact normal (correlatePatches,
 * DyFrictionCorrelation.cpp:136-184); otherwise a new patch starts from the
 * contacts. Shared by the TGS and PGS preparation. */
static int friction_patch(const pxo_world* w, pxo_con* k, xf f0, xf f1, v3 n, float restDist)
{
	fpatch* fp = &k->fp;
	const fpatch* pp = find_patch(w, k->pkey);
	int kept = 0;
	if (pp && !pp->broken && pp->count != 0)
	{
		const xf b10 = xinvmul(f0, f1);
		if (dot(pp->n0, qrot(b10.q, pp->n1)) > 0.999f)
		{
			int separated = 0;
			for (int a = 0; a < pp->count; a++)
				if (!(fabsf(dot(sub(pp->a0[a], xtrans(b10, pp->a1[a])), pp->n0)) < w->d.correlation_distance))
					separated = 1;
			if (!separated && !(dot(n, qrot(f0.q, pp->n0)) < 0.999f))
				kept = 1;
		}
	}
	if (kept)
		*fp = *pp;
	else
	{
		fp->n0 = qrotinv(f0.q, n);
		fp->n1 = qrotinv(f1.q, n);
		fp->count = 0;
	}
	fp->key = k->pkey;
	fp->broken = 0;
	/* anchors (growPatches, DyFrictionCorrelation.cpp:189-316): two kept
	 * anchors stay unless they are closer than half the diagonal of this
	 * step's contact bounds; one kept anchor is completed from the contacts;
	 * otherwise up to two contacts closer than the offset threshold */
	if (fp->count == 2)
	{
		v3 mn = k->c.p[0], mx = k->c.p[0];
		for (int j = 1; j < k->c.count; j++)
		{
			const v3 q = k->c.p[j];
			mn = V(fminf(mn.x, q.x), fminf(mn.y, q.y), fminf(mn.z, q.z));
			mx = V(fmaxf(mx.x, q.x), fmaxf(mx.y, q.y), fmaxf(mx.z, q.z));
		}
		const float diag2 = len2(sub(mx, mn));
		const float anch2 = len2(sub(fp->a0[0], fp->a0[1]));
		if (!(anch2 * 4.0f >= diag2))
			fp->count = 0;
	}
	if (fp->count != 2)
	{
		int nAnch = 0;
		v3 wa[2];
		float pd2 = 0.0f;
		if (fp->count == 1)
			wa[nAnch++] = xtrans(f0, fp->a0[0]);
		const float thr = w->d.friction_offset_threshold + restDist;
		for (int j = 0; j < k->c.count; j++)
		{
			const v3 wp = k->c.p[j];
			if (k->c.sep[j] < thr)
			{
				if (nAnch == 0) { wa[0] = wp; nAnch = 1; }
				else if (nAnch == 1)
				{
					pd2 = len2(sub(wp, wa[0]));
					if (pd2 > 1e-8f) { wa[1] = wp; nAnch = 2; }
				}
				else
				{
					const float d0 = len2(sub(wp, wa[0])), d1 = len2(sub(wp, wa[1]));
					if (d0 > d1) { if (d0 > pd2) { wa[1] = wp; pd2 = d0; } }
					else if (d1 > pd2) { wa[0] = wp; pd2 = d1; }
				}
			}
		}
		for (int j = fp->count; j < nAnch; j++)
		{
			fp->a0[j] = xtransinv(f0, wa[j]);
			fp->a1[j] = xtransinv(f1, wa[j]);
		}
		if (nAnch == 0)
			fp->a0[0] = fp->a1[0] = V(0.0f, 0.0f, 0.0f);
		fp->count = nAnch;
	}
	return fp->count;
}

/* setupFinalizeSolverConstraints / constructContactConstraintStep for one
 * rigid-rigid (or rigid-static) contact patch:
 * lowleveldynamics/src/DyTGSContactPrep.cpp:426-812 (rows), :310-404 (normal row),
 * friction patch creation and anchors DyFrictionCorrelation.cpp:55-292 */
static void prep_con(const pxo_world* w, pxo_con* k, const sbody* sb, const sbody* world, const xf* frame,
                     float invStepDt, float stepDt, float totalDt, float scale)
{
	const sbody* b0 = &sb[k->a];
	const sbody* b1 = k->b >= 0 ? &sb[k->b] : world;
	const xf f0 = frame[k->a];
	xf f1;
	if (k->b >= 0) f1 = frame[k->b];
	else { f1.p = V(0, 0, 0); f1.q.x = 0; f1.q.y = 0; f1.q.z = 0; f1.q.w = 1; }
	const float invDtp8 = invStepDt * scale;
	const v3 n = k->c.n;
	const float nlen2 = dot(n, n);
	k->invMassA = b0->invMass;
	k->invMassB = b1->invMass;
	k->maxPenBias = -1e32f; /* max(penBiasClamp0, penBiasClamp1), defaults -1e32 (PxvDynamics.h:140) */
	const float imn0 = b0->invMass * nlen2, imn1 = -b1->invMass * nlen2;
	const float restDist = 2.0f * w->d.rest_offset;
	(void)totalDt;
	for (int i = 0; i < k->c.count; i++)
	{
		const v3 ra = sub(k->c.p[i], f0.p), rb = sub(k->c.p[i], f1.p);
		const v3 raXn = cross(ra, n), rbXn = cross(rb, n);
		const v3 raXnI = mmul(b0->sqrtInvI, raXn), rbXnI = mmul(b1->sqrtInvI, rbXn);
		const float i0 = dot(raXnI, raXnI), i1 = dot(rbXnI, rbXnI);
		const float resp0 = imn0 + i0, resp1 = i1 - imn1;
		const float unit = resp0 + resp1;
		const float pen = k->c.sep[i] - restDist;
		const float recip = unit > 0.0f ? 1.0f / unit : 0.0f;
		k->raXnI[i] = raXnI;
		k->rbXnI[i] = rbXnI;
		k->sepErr[i] = pen;
		k->biasCoef[i] = -(pen > 0.0f ? invStepDt : invDtp8);
		k->velMul[i] = recip;
		k->recipResp[i] = recip;
		k->force[i] = 0.0f;
	}
	const int nAnch = friction_patch(w, k, f0, f1, n, restDist);
	const fpatch* fp = &k->fp;
	k->nFric = 2 * nAnch;
	if (nAnch == 0)
		return;
	const v3 lvr = sub(b0->origLin, b1->origLin);
	const v3 fb1 = V(0.0f, -n.z, n.y), fb2 = V(-n.y, n.x, 0.0f);
	const v3 fb = (0.70710678f > fabsf(n.x)) ? fb1 : fb2;
	v3 t0 = sub(lvr, scl(n, dot(n, lvr)));
	t0 = len2(t0) > 0.0001f ? t0 : fb;
	t0 = normalize(t0);
	const v3 t1 = normalize(cross(n, t0));
	const v3 relTr = sub(f0.p, f1.p);
	k->fscale = nAnch == 2 ? 0.5f : 1.0f;
	k->fbiasScale = invStepDt * scale;
	for (int j = 0; j < nAnch; j++)
	{
		const v3 ra = qrot(f0.q, fp->a0[j]), rb = qrot(f1.q, fp->a1[j]);
		const v3 err = add(sub(ra, rb), relTr);
		for (int d = 0; d < 2; d++)
		{
			const v3 t = d == 0 ? t0 : t1;
			const v3 raXn = cross(ra, t), rbXn = cross(rb, t);
			const v3 raXnI = mmul(b0->sqrtInvI, raXn), rbXnI = mmul(b1->sqrtInvI, rbXn);
			const float resp0 = imn0 + dot(raXnI, raXnI), resp1 = dot(rbXnI, rbXnI) - imn1;
			const float unit = resp0 + resp1;
			const int r = 2 * j + d;
			k->fn[r] = t;
			k->ferr[r] = dot(err, t);
			k->fraXnI[r] = raXnI;
			k->frbXnI[r] = rbXnI;
			k->fvelMul[r] = unit > 0.0f ? scale / unit : 0.0f;
			k->fforce[r] = 0.0f;
		}
	}
}

/* solveContact / solveDynamicContactsStep
 * (lowleveldynamics/src/DyTGSContactPrep.cpp:1490-1843) */
static void solve_con(pxo_con* k, sbody* b0, sbody* b1, float minPen, float elapsed, float mu_s, float mu_d)
{
	v3 l0 = b0->lin, l1 = b1->lin, a0 = b0->ang, a1 = b1->ang;
	const v3 am0 = b0->dang, am1 = b1->dang;
	const v3 rel = sub(b0->dlin, b1->dlin);
	const v3 n = k->c.n;
	const v3 nim0 = scl(n, k->invMassA), nim1 = scl(n, k->invMassB);
	const float dv = dot(rel, n);
	float acc = 0.0f;
	for (int i = 0; i < k->c.count; i++)
	{
		const float ad0 = dot(am0, k->raXnI[i]), ad1 = dot(am1, k->rbXnI[i]);
		const float dang = ad0 - ad1;
		const float dbias = (dv + dang) - 0.0f * elapsed;
		const float sep = fmaxf(minPen, k->sepErr[i] + dbias);
		const float bias = fminf(-k->maxPenBias, k->biasCoef[i] * sep);
		const v3 v0 = add(mulv(l0, n), mulv(a0, k->raXnI[i]));
		const v3 v1 = add(mulv(l1, n), mulv(a1, k->rbXnI[i]));
		const v3 dvv = sub(v0, v1);
		const float nv = dvv.x + dvv.y + dvv.z;
		const float biasNV = bias * k->recipResp[i];
		const float lambda = biasNV - (nv - 0.0f) * k->velMul[i];
		const float af = k->force[i];
		const float df0 = fmaxf(lambda, -af);
		const float nf = fminf(af + df0, 1e32f);
		const float df = nf - af;
		l0 = add(scl(nim0, df), l0);
		l1 = sub(l1, scl(nim1, df));
		a0 = add(scl(k->raXnI[i], df), a0);
		a1 = sub(a1, scl(k->rbXnI[i], df));
		k->force[i] = nf;
		acc = acc + nf;
	}
	if (k->nFric)
	{
		const float maxF = mu_s * acc, maxD = mu_d * acc;
		int broken = 0;
		for (int i = 0; i < k->nFric; i += 2)
		{
			float ti[2], af[2];
			for (int d = 0; d < 2; d++)
			{
				const int r = i + d;
				float dvf = (dot(k->fraXnI[r], am0) - dot(k->frbXnI[r], am1)) + dot(k->fn[r], rel);
				dvf = dvf - 0.0f * elapsed;
				const float err = k->ferr[r] + dvf;
				const float bias = err * k->fbiasScale;
				const v3 v0 = add(mulv(l0, k->fn[r]), mulv(a0, k->fraXnI[r]));
				const v3 v1 = add(mulv(l1, k->fn[r]), mulv(a1, k->frbXnI[r]));
				const v3 dd = sub(v0, v1);
				const float nv = dd.x + dd.y + dd.z;
				af[d] = k->fforce[r];
				const float tmp = af[d] - (bias - 0.0f) * k->fvelMul[r];
				ti[d] = tmp - nv * k->fvelMul[r];
			}
			const float tot = sqrtf(ti[0] * ti[0] + ti[1] * ti[1]);
			const int clamp = tot > k->fscale * maxF;
			const float tc = clamp ? fminf(k->fscale * maxD, tot) : tot;
			const float ratio = tot > 0.0f ? tc / tot : 0.0f;
			const float n0 = ti[0] * ratio, n1 = ti[1] * ratio;
			const float d0 = n0 - af[0], d1 = n1 - af[1];
			const v3 t0 = k->fn[i], t1 = k->fn[i + 1];
			l0 = add(scl(scl(t0, k->invMassA), d0), add(scl(scl(t1, k->invMassA), d1), l0));
			l1 = sub(sub(l1, scl(scl(t1, k->invMassB), d1)), scl(scl(t0, k->invMassB), d0));
			a0 = add(scl(k->fraXnI[i], d0), add(scl(k->fraXnI[i + 1], d1), a0));
			a1 = sub(sub(a1, scl(k->frbXnI[i + 1], d1)), scl(k->frbXnI[i], d0));
			k->fforce[i] = n0;
			k->fforce[i + 1] = n1;
			broken |= clamp;
		}
		k->fp.broken = broken; /* the last pass's value is kept (DyTGSContactPrep.cpp:1837, 1907) */
	}
	b0->lin = l0; b0->ang = a0;
	if (b1->invMass > 0.0f) { b1->lin = l1; b1->ang = a1; }
}

/* PGS contact rows (setupFinalizeSolverConstraints, lowleveldynamics/src/
 * DyContactPrep.cpp:49-357; normal rows constructContactConstraint,
 * DyContactPrepShared.h:298-393). The solver works on velocity changes: the
 * bodies' pre-integrated velocities (origLin, origAng world) enter as the
 * target velocity -vrel of every row. */
static void prep_con_pgs(const pxo_world* w, pxo_con* k, const sbody* sb, const sbody* world, const xf* frame,
                         float invDt, float mu_s, float mu_d)
{
	const sbody* b0 = &sb[k->a];
	const sbody* b1 = k->b >= 0 ? &sb[k->b] : world;
	const xf f0 = frame[k->a];
	xf f1;
	if (k->b >= 0) f1 = frame[k->b];
	else { f1.p = V(0, 0, 0); f1.q.x = 0; f1.q.y = 0; f1.q.z = 0; f1.q.w = 1; }
	const v3 n = k->c.n;
	const float nl2 = dot(n, n);
	const float invMass0 = b0->invMass, invMass1 = b1->invMass;
	const float imn0 = invMass0 * nl2, imn1 = -invMass1 * nl2;
	const v3 v0 = b0->origLin, v1 = b1->origLin, w0 = b0->origAng, w1 = b1->origAng;
	const float restDist = 2.0f * w->d.rest_offset;
	const float invDtp8 = invDt * 0.8f;
	const float maxPenBias = -1e32f; /* max(penBiasClamp0, penBiasClamp1), DyDynamics.cpp:144 */
	k->invMassA = invMass0;
	k->invMassB = invMass1;
	const float norVel = (n.x * v0.x - n.x * v1.x) + (n.y * v0.y - n.y * v1.y) + (n.z * v0.z - n.z * v1.z);
	for (int i = 0; i < k->c.count; i++)
	{
		const v3 ra = sub(k->c.p[i], f0.p), rb = sub(k->c.p[i], f1.p);
		const v3 raXn = cross(ra, n), rbXn = cross(rb, n);
		const float vrel = norVel + (dot(raXn, w0) - dot(rbXn, w1));
		const v3 raXnI = mmul(b0->sqrtInvI, raXn), rbXnI = mmul(b1->sqrtInvI, rbXn);
		const float unit = (imn0 + dot(raXnI, raXnI)) + (dot(rbXnI, rbXnI) - imn1);
		const float pen = k->c.sep[i] - restDist;
		const float target = 0.0f - vrel;
		const float recip = unit > 0.0f ? 1.0f / unit : 0.0f;
		const float penScaled = pen >= 0.0f ? pen * invDt : pen * invDtp8;
		const float scaledBias = recip * fmaxf(maxPenBias, penScaled);
		k->raXnI[i] = raXnI;
		k->rbXnI[i] = rbXnI;
		k->velMul[i] = recip;
		k->biasedErr[i] = target * recip + -scaledBias;
		k->unbiasedErr[i] = target * recip + -fmaxf(scaledBias, 0.0f);
		k->force[i] = 0.0f;
	}
	const int nAnch = friction_patch(w, k, f0, f1, n, restDist);
	const fpatch* fp = &k->fp;
	k->nFric = 2 * nAnch;
	const float fc = nAnch == 2 ? 0.5f : 1.0f;
	k->muS = mu_s * fc;
	k->muD = mu_d * fc;
	if (nAnch == 0)
		return;
	const v3 lvr = sub(v0, v1);
	const v3 fb1 = V(0.0f, -n.z, n.y), fb2 = V(-n.y, n.x, 0.0f);
	const v3 fb = (0.70710678f > fabsf(n.x)) ? fb1 : fb2;
	v3 t0 = sub(lvr, scl(n, dot(n, lvr)));
	t0 = len2(t0) > 0.0001f ? t0 : fb;
	t0 = normalize(t0);
	const v3 t1 = cross(n, t0);
	for (int j = 0; j < nAnch; j++)
	{
		const v3 ra = qrot(f0.q, fp->a0[j]), rb = qrot(f1.q, fp->a1[j]);
		const v3 err = sub(add(ra, f0.p), add(rb, f1.p));
		for (int d = 0; d < 2; d++)
		{
			const v3 t = d == 0 ? t0 : t1;
			const v3 raXn = cross(ra, t), rbXn = cross(rb, t);
			const v3 raXnI = mmul(b0->sqrtInvI, raXn), rbXnI = mmul(b1->sqrtInvI, rbXn);
			const float resp = (invMass0 + dot(raXnI, raXnI)) + (dot(rbXnI, rbXnI) - -invMass1);
			const float vrel = (dot(t, v0) + dot(raXn, w0)) - (dot(t, v1) + dot(rbXn, w1));
			const int r = 2 * j + d;
			k->fn[r] = t;
			k->fraXnI[r] = raXnI;
			k->frbXnI[r] = rbXnI;
			k->fvelMul[r] = resp > 0.0f ? 0.8f / resp : 0.0f;
			k->ftarget[r] = 0.0f - vrel;
			k->fbias[r] = dot(t, err) * invDt;
			k->fforce[r] = 0.0f;
		}
	}
}

/* PGS solveContact (lowleveldynamics/src/DySolverConstraints.cpp:221-371,
 * normal rows solveDynamicContacts, DySolverConstraintsShared.h:49-105) on the
 * velocity changes l (linear) and a (sqrt-inertia angular) of both bodies;
 * friction rows are clamped one at a time. conclude: after the last position
 * pass the rows switch to their unbiased error and friction loses its bias
 * (concludeContact, DySolverConstraints.cpp:508-549). */
static void solve_con_pgs(pxo_con* k, sbody* b0, sbody* b1, int doFriction, int conclude)
{
	v3 l0 = b0->lin, l1 = b1->lin, a0 = b0->ang, a1 = b1->ang;
	const v3 n = k->c.n;
	const v3 dl0 = scl(n, k->invMassA), dl1 = scl(n, k->invMassB);
	float acc = 0.0f;
	for (int i = 0; i < k->c.count; i++)
	{
		const v3 v0 = add(mulv(l0, n), mulv(a0, k->raXnI[i]));
		const v3 v1 = add(mulv(l1, n), mulv(a1, k->rbXnI[i]));
		const v3 dv = sub(v0, v1);
		const float nv = dv.x + dv.y + dv.z;
		const float af = k->force[i];
		const float d0 = fmaxf(k->biasedErr[i] - nv * k->velMul[i], -af);
		const float nf = fminf(af + d0, 1e32f);
		const float df = nf - af;
		l0 = add(scl(dl0, df), l0);
		l1 = sub(l1, scl(dl1, df));
		a0 = add(scl(k->raXnI[i], df), a0);
		a1 = sub(a1, scl(k->rbXnI[i], df));
		k->force[i] = nf;
		acc = acc + nf;
	}
	if (doFriction && k->nFric)
	{
		const float maxF = k->muS * acc, maxD = k->muD * acc;
		int broken = 0;
		for (int r = 0; r < k->nFric; r++)
		{
			const v3 t = k->fn[r];
			const v3 v0 = add(mulv(l0, t), mulv(a0, k->fraXnI[r]));
			const v3 v1 = add(mulv(l1, t), mulv(a1, k->frbXnI[r]));
			const v3 dv = sub(v0, v1);
			const float nv = dv.x + dv.y + dv.z;
			const float af = k->fforce[r];
			const float vm = k->fvelMul[r];
			const float tmp = af - (k->fbias[r] - k->ftarget[r]) * vm;
			const float total = tmp - nv * vm;
			const int clamp = fabsf(total) > maxF;
			const float tc = fminf(maxD, fmaxf(-maxD, total));
			const float nf = clamp ? tc : total;
			broken |= clamp;
			const float df = nf - af;
			l0 = add(scl(scl(t, k->invMassA), df), l0);
			l1 = sub(l1, scl(scl(t, k->invMassB), df));
			a0 = add(scl(k->fraXnI[r], df), a0);
			a1 = sub(a1, scl(k->frbXnI[r], df));
			k->fforce[r] = nf;
		}
		k->fp.broken = broken;
	}
	if (conclude)
	{
		for (int i = 0; i < k->c.count; i++)
			k->biasedErr[i] = k->unbiasedErr[i];
		for (int r = 0; r < k->nFric; r++)
			k->fbias[r] = 0.0f;
	}
	b0->lin = l0; b0->ang = a0;
	if (b1->invMass > 0.0f) { b1->lin = l1; b1->ang = a1; }
}

static int cmp_u64(const void* x, const void* y)
{
	const uint64_t a = *(const uint64_t*)x, b = *(const uint64_t*)y;
	return a < b ? -1 : (a > b ? 1 : 0);
}

/* whether bodies i < j are joined by a joint: jointed actors do not collide
 * (PxConstraintFlag::eCOLLISION_ENABLED is off by default, physx/include/
 * PxConstraint.h; the pair is filtered out before the narrowphase) */
static int jointed(const pxo_world* w, int i, int j)
{
	if (!w->nj)
		return 0;
	const uint64_t k = ((uint64_t)(uint32_t)i << 32) | (uint32_t)j;
	return bsearch(&k, w->jkeys, (size_t)w->nj, sizeof(uint64_t), cmp_u64) != NULL;
}

static void add_pair(pxo_world* w, int i, int j, float cd, float tol)
{
	if (jointed(w, i, j))
		return;
	/* shape order: lower geometry type first (the contact-method table order);
	 * for two shapes of one type the later-added body first, as the
	 * reference's contact managers order them (shouldSwapBodies,
	 * simulationcontroller/src/ScNPhaseCore.cpp:182-249; i < j here) */
	int s0 = j, s1 = i;
	if (w->type[i] < w->type[j]) { s0 = i; s1 = j; }
	cset c;
	c.count = 0;
	int n = 0;
	const float* g0 = w->dims + 4 * s0;
	const float* g1 = w->dims + 4 * s1;
	if (w->type[s0] == PXO_SPHERE && w->type[s1] == PXO_SPHERE)
		n = c_sphere_sphere(g0[0], w->pose[s0].p, g1[0], w->pose[s1].p, cd, &c);
	else if (w->type[s0] == PXO_SPHERE && w->type[s1] == PXO_CAPSULE)
		n = c_sphere_capsule(g0[0], w->pose[s0].p, g1[0], g1[1], w->pose[s1], cd, &c);
	else if (w->type[s0] == PXO_CAPSULE && w->type[s1] == PXO_CAPSULE)
		n = c_capsule_capsule(g0[0], g0[1], w->pose[s0], g1[0], g1[1], w->pose[s1], cd, &c);
	else if (w->type[s0] == PXO_CAPSULE && w->type[s1] == PXO_BOX)
	{
		boxpoly bp;
		box_poly(V(g1[0], g1[1], g1[2]), &bp);
		n = c_capsule_poly(g0[0], g0[1], w->pose[s0], &bp.d, 1, w->pose[s1], cd, tol, &c);
	}
	else if (w->type[s0] == PXO_CAPSULE)
		n = c_capsule_poly(g0[0], g0[1], w->pose[s0], hull_of(g1[0]), 0, w->pose[s1], cd, tol, &c);
	else if (w->type[s0] == PXO_SPHERE && w->type[s1] == PXO_BOX)
		n = c_sphere_box(g0[0], w->pose[s0].p, V(g1[0], g1[1], g1[2]), w->pose[s1], cd, &c);
	else if (w->type[s0] == PXO_BOX && w->type[s1] == PXO_BOX)
		n = c_box_box(V(g0[0], g0[1], g0[2]), w->pose[s0], V(g1[0], g1[1], g1[2]), w->pose[s1], cd, tol, &c);
	else if (w->type[s0] == PXO_SPHERE)
		n = c_sphere_convex(g0[0], w->pose[s0].p, hull_of(g1[0]), w->pose[s1], cd, &c);
	else if (w->type[s0] == PXO_BOX)
	{
		boxpoly bp;
		box_poly(V(g0[0], g0[1], g0[2]), &bp);
		n = c_convex_convex(&bp.d, w->pose[s0], hull_of(g1[0]), w->pose[s1], cd, tol, &c);
	}
	else
		n = c_convex_convex(hull_of(g0[0]), w->pose[s0], hull_of(g1[0]), w->pose[s1], cd, tol, &c);
	if (n > 0)
		push_con(w, s0, s1, &c);
}

/* ---------------------------------------------------------------------------
 * Joints: 1D constraint rows of fixed, spherical, revolute and prismatic
 * joints, with the hard limits and the velocity drive of revolute and
 * prismatic joints, TGS.
 * ------------------------------------------------------------------------- */

/* Px1DConstraint (physx/include/PxConstraintDesc.h:143-230), the fields the
 * locked-axis rows set; rows start from setupConstraintRows
 * (lowleveldynamics/src/DyConstraintPrep.h:115-133): zeros, impulses
 * unbounded. */
typedef struct {
	v3 lin0, ang0, lin1, ang1;
	float geomErr, velTarget, minImp, maxImp, rest, bthr, stiff, damp;
	int hint, angular, flags;   /* flags RF_* (Px1DConstraintFlag) */
} row1d;

/* PxConstraintSolveHint (physx/include/PxConstraintDesc.h:79-110) */
enum { HINT_NONE = 0, HINT_ROT_EQ = 1024, HINT_ROT_INEQ = 1025, HINT_EQ = 2048, HINT_INEQ = 2049 };
/* Px1DConstraintFlag: eKEEPBIAS, eRESTITUTION, eSPRING, eHAS_DRIVE_LIMIT */
enum { RF_KEEP_BIAS = 1, RF_RESTITUTION = 2, RF_SPRING = 4, RF_DRIVE_LIMIT = 8 };

/* Jacobian of the relative rotation qa* qb against (omega_b - omega_a)
 * (computeJacobianAxes, physxextensions/src/ExtConstraintHelper.h:77-102) */
static void jacobian_axes(v3 row[3], qt qa, qt qb)
{
	const float wa = qa.w, wb = qb.w;
	const v3 va = V(qa.x, qa.y, qa.z), vb = V(qb.x, qb.y, qb.z);
	const v3 c = add(scl(vb, wa), scl(va, wb));
	const float d0 = wa * wb;
	const float d1 = dot(va, vb);
	const float d = d0 - d1;
	row[0] = scl(add(add(scl(va, vb.x), scl(vb, va.x)), V(d, c.z, -c.y)), 0.5f);
	row[1] = scl(add(add(scl(va, vb.y), scl(vb, va.y)), V(-c.z, d, c.x)), 0.5f);
	row[2] = scl(add(add(scl(va, vb.z), scl(vb, va.z)), V(c.y, -c.x, d)), 0.5f);
	if ((d0 + d1) == 0.0f)
	{
		row[0].x += 1.192092896e-07f;
		row[1].y += 1.192092896e-07f;
		row[2].z += 1.192092896e-07f;
	}
}

/* locked axes of each joint kind (linear, angular bit masks x y z):
 * FixedJointSolverPrep (ExtFixedJoint.cpp:70), SphericalJointSolverPrep
 * (ExtSphericalJoint.cpp:157), RevoluteJointSolverPrep
 * (ExtRevoluteJoint.cpp:221, twist about x free), PrismaticJointSolverPrep
 * (ExtPrismaticJoint.cpp:135, translation along x free) */
static void joint_axes(int type, int* lin, int* ang)
{
	switch (type)
	{
	case PXO_JOINT_FIXED: *lin = 7; *ang = 7; break;
	case PXO_JOINT_SPHERICAL: *lin = 7; *ang = 0; break;
	case PXO_JOINT_REVOLUTE: *lin = 7; *ang = 6; break;
	case PXO_JOINT_DISTANCE: *lin = 0; *ang = 0; break;
	default: *lin = 6; *ang = 7; break;
	}
}

/* soft limits: a spring with the joint's stiffness or damping */
static int soft_limit(const pxo_joint* j) { return j->damp > 0.0f || j->stiff > 0.0f; }

/* a limit row (ConstraintHelper::addLimit, ExtConstraintHelper.h:316-393):
 * soft, a spring row pushing back (no solve hint); hard, an inequality
 * keeping its bias while the limit is not reached, bouncing when
 * restitution is set; impulses at least zero */
static void limit_row(row1d* c, const pxo_joint* j)
{
	if (soft_limit(j))
	{
		c->flags |= RF_SPRING;
		c->stiff = j->stiff;
		c->damp = j->damp;
		c->minImp = 0.0f;
		return;
	}
	c->hint = c->angular ? HINT_ROT_INEQ : HINT_INEQ;
	c->rest = j->rest;
	c->bthr = j->bthr;
	if (c->geomErr > 0.0f)
		c->flags |= RF_KEEP_BIAS;
	if (j->rest > 0.0f)
		c->flags |= RF_RESTITUTION;
	c->minImp = 0.0f;
}

/* the twist angle of frame B about frame A's x axis (computePhi,
 * ExtRevoluteJoint.cpp:154-179) */
static float twist_angle(qt qa, qt qb)
{
	const qt r = qmul(qconj(qa), qb);
	float x = r.x, y = 0.0f, z = 0.0f, ww = r.w;
	const float mag = sqrtf(x * x + y * y + z * z + ww * ww);
	if (mag != 0.0f)
	{
		const float im = 1.0f / mag;
		x *= im; ww *= im;
	}
	float angle = acosf(fminf(fmaxf(ww, -1.0f), 1.0f)) * 2.0f;
	if (x < 0.0f)
		angle = -angle;
	return angle;
}

/* the closest point of the ellipse with radii (ry, rz) in the y-z plane to
 * point (PxEllipseClamp, physx/include/foundation/PxMathUtils.h:204-262:
 * Lagrange multiplier, Newton steps from Eberly's initial estimate) */
static v3 ellipse_clamp(v3 point, float ry, float rz)
{
	const float qy = fabsf(point.y), qz = fabsf(point.z);
	const float tinyEps = 1e-6f;
	if (ry >= rz)
	{
		if (qz < tinyEps)
			return V(0.0f, point.y > 0.0f ? ry : -ry, 0.0f);
	}
	else if (qy < tinyEps)
		return V(0.0f, 0.0f, point.z > 0.0f ? rz : -rz);
	const float e2y = ry * ry, e2z = rz * rz, eqy = ry * qy, eqz = rz * qz;
	float t = fmaxf(eqy - e2y, eqz - e2z);
	float dy = 0.0f, dz = 0.0f;
	for (int i = 0; i < 20; i++)
	{
		dy = 1.0f / (t + e2y);
		dz = 1.0f / (t + e2z);
		const float d2y = eqy * dy, d2z = eqz * dz;
		const float fy = d2y * d2y, fz = d2z * d2z;
		const float f = fy + fz - 1.0f;
		if (f < 1e-4f)
			return V(0.0f, e2y * point.y * dy, e2z * point.z * dz);
		const float df = (0.0f * 0.0f + fy * dy + fz * dz) * -2.0f;
		t = t - f / df;
	}
	const float ry_ = e2y * point.y * dy, rz_ = e2z * point.z * dz;
	const float a = ry_ / ry, b = rz_ / rz;
	const float s = 1.0f / sqrtf(a * a + b * b);
	return V(0.0f * s, ry_ * s, rz_ * s);
}

/* the cone limit's axis and error (ConeLimitHelperTanLess::getLimit with
 * computeAxisAndError, common/src/CmConeLimitHelper.h:50-69, 163-178):
 * swing.x = 0; the axis turns the twist axis outward from the limit, the
 * error is positive inside the cone */
static float cone_limit(qt swing, float ymax, float zmax, v3* axis)
{
	const float x2 = swing.x * 2.0f, w2 = swing.w * 2.0f;
	const v3 twistAxis = V((swing.w * w2) - 1.0f + swing.x * x2, (swing.z * w2) + swing.y * x2,
	                       (-swing.y * w2) + swing.z * x2);
	const v3 swingAngle = V(0.0f, 4.0f * atan2f(swing.y, 1.0f + swing.w), 4.0f * atan2f(swing.z, 1.0f + swing.w));
	const v3 p = ellipse_clamp(swingAngle, ymax, zmax);
	const v3 normal = V(0.0f, p.y / (ymax * ymax), p.z / (zmax * zmax));
	const v3 r = V(0.0f, tanf(p.y / 4.0f), tanf(p.z / 4.0f)), d = V(0.0f, normal.y, normal.z);
	const v3 px = V(1.0f, 0.0f, 0.0f);
	const float r2 = dot(r, r), a = 1.0f - r2, b = 1.0f / (1.0f + r2), b2 = b * b;
	const float v1 = 2.0f * a * b2;
	const v3 v2 = V(a, 2.0f * r.z, -2.0f * r.y);
	const v3 coneLine = sub(scl(v2, v1), px);
	const float rd = dot(r, d);
	const float dv1 = -4.0f * rd * (3.0f - r2) * b2 * b;
	const v3 dv2 = V(-2.0f * rd, 2.0f * d.z, -2.0f * d.y);
	const v3 coneNormal = add(scl(dv2, v1), scl(v2, dv1));
	const float inv = 1.0f / sqrtf(dot(coneNormal, coneNormal));
	*axis = scl(cross(coneLine, coneNormal), inv);
	return dot(cross(coneLine, *axis), twistAxis);
}

/* The joint's solverPrep with ConstraintHelper::prepareLockedAxes
 * (ExtConstraintHelper.h:150-208, 243-297): joint frames in the world
 * (computeJointFrames, :61-73), the neighbourhood operator (:44-48), rows
 * for the locked linear axes of frame A through the anchor, rows for the
 * locked rotation axes; then for a revolute joint its drive row and limit
 * pair (RevoluteJointSolverPrep, ExtRevoluteJoint.cpp:195-258: a locked
 * limit, lower >= upper, locks the twist instead), for a prismatic joint its
 * limit pair along x (PrismaticJointSolverPrep, ExtPrismaticJoint.cpp:
 * 119-153), limit rows through the unshifted anchor. Returns the row count;
 * *ra / *rb the anchor offsets (cA2wOut - bA2w.p, cB2wOut - bB2w.p). */
static int joint_rows(const pxo_joint* j, xf bA, xf bB, row1d* rows, v3* raOut, v3* rbOut)
{
	xf cA, cB;
	cA.q = qmul(bA.q, j->f0.q);
	cA.p = add(qrot(bA.q, j->f0.p), bA.p);
	cB.q = qmul(bB.q, j->f1.q);
	cB.p = add(qrot(bB.q, j->f1.p), bB.p);
	if (cA.q.x * cB.q.x + cA.q.y * cB.q.y + cA.q.z * cB.q.z + cA.q.w * cB.q.w < 0.0f)
	{
		cB.q.x = -cB.q.x; cB.q.y = -cB.q.y; cB.q.z = -cB.q.z; cB.q.w = -cB.q.w;
	}
	const v3 ra0 = sub(cB.p, bA.p);
	v3 ra = ra0;
	const v3 rb = sub(cB.p, bB.p);
	int lin, ang;
	joint_axes(j->type, &lin, &ang);
	const int limitOn = (j->type == PXO_JOINT_REVOLUTE || j->type == PXO_JOINT_PRISMATIC) && (j->pflags & JP_LIMIT);
	const int limitLocked = limitOn && j->lo >= j->hi;
	if (limitLocked)
	{
		if (j->type == PXO_JOINT_REVOLUTE) ang = 7;
		else lin = 7;
	}
	const v3 cB2cAp = qrotinv(cA.q, sub(cB.p, cA.p));
	int k = 0;
	for (int i = 0; i < JOINT_MAX_ROWS; i++)
	{
		memset(&rows[i], 0, sizeof(row1d));
		rows[i].minImp = -3.402823466e+38f;
		rows[i].maxImp = 3.402823466e+38f;
	}
	if (j->type == PXO_JOINT_DISTANCE)
	{
		/* DistanceJointSolverPrep (ExtDistanceJoint.cpp:179-271): rows along
		 * the anchors' direction through frame B's anchor; flags 1 max
		 * distance, 2 min distance, 4 spring; lower / upper the min / max
		 * distance, the restitution slot the tolerance */
		const int enforceMax = (j->pflags & 1) && j->hi >= 0.0f;
		const int enforceMin = (j->pflags & 2) && j->lo >= 0.0f;
		if (enforceMax || enforceMin)
		{
			v3 dir = sub(cA.p, cB.p);
			const float dist = sqrtf(dot(dir, dir));
			if (dist > 0.0f)
				dir = scl(dir, 1.0f / dist);
			if (dist < 1.192092896e-07f)
				dir = V(1.0f, 0.0f, 0.0f);
			const float tol = j->rest;
			row1d* c = &rows[k];
			const int both = enforceMin && enforceMax;
			const int nrows = both && j->lo != j->hi ? 2 : 1;
			for (int r = 0; r < nrows; r++)
			{
				row1d* e = &c[r];
				e->lin0 = dir; e->ang0 = cross(ra0, dir);
				e->lin1 = dir; e->ang1 = cross(rb, dir);
				if (j->pflags & 4)
				{
					e->flags |= RF_SPRING;
					e->stiff = j->stiff;
					e->damp = j->damp;
				}
			}
			if (both && j->lo == j->hi)
			{
				const float error = dist - j->hi;
				c->geomErr = error > tol ? error - tol : (error < -tol ? error + tol : 0.0f);
			}
			else
			{
				if (enforceMin)
				{
					c->geomErr = dist - j->lo + tol;
					c->minImp = 0.0f;
					if (dist >= j->lo)
						c->flags |= RF_KEEP_BIAS;
					c++;
				}
				if (enforceMax)
				{
					c->geomErr = dist - j->hi - tol;
					c->maxImp = 0.0f;
					if (dist <= j->hi)
						c->flags |= RF_KEEP_BIAS;
				}
			}
			k += nrows;
		}
	}
	if (j->type == PXO_JOINT_SPHERICAL && (j->pflags & JP_LIMIT))
	{
		/* the cone limit row comes first (SphericalJointSolverPrep,
		 * ExtSphericalJoint.cpp:157-190): swing of B's frame in A's
		 * (PxSeparateSwingTwist, PxMathUtils.h:272-276), its limit axis in
		 * the world */
		const qt q = qmul(qconj(cA.q), cB.q);
		qt tw = {0.0f, 0.0f, 0.0f, 1.0f};
		if (q.x != 0.0f)
		{
			const float sc = 1.0f / sqrtf(q.x * q.x + 0.0f * 0.0f + 0.0f * 0.0f + q.w * q.w);
			tw.x = q.x * sc; tw.y = 0.0f * sc; tw.z = 0.0f * sc; tw.w = q.w * sc;
		}
		const qt swing = qmul(q, qconj(tw));
		v3 ax;
		const float err = cone_limit(swing, j->lo, j->hi, &ax);
		row1d* c = &rows[k++];
		c->ang0 = qrot(cA.q, ax); c->ang1 = c->ang0; c->angular = 1;
		c->geomErr = err;
		limit_row(c, j);
	}
	const m33 axes = qmat(cA.q);
	if (lin)
	{
		v3 ev = V(0, 0, 0);
		if (lin & 1) ev = sub(ev, scl(axes.c0, cB2cAp.x));
		if (lin & 2) ev = sub(ev, scl(axes.c1, cB2cAp.y));
		if (lin & 4) ev = sub(ev, scl(axes.c2, cB2cAp.z));
		ra = add(ra, ev);
		const v3 ax[3] = {axes.c0, axes.c1, axes.c2};
		for (int i = 0; i < 3; i++)
			if (lin & (1 << i))
			{
				row1d* c = &rows[k++];
				c->hint = HINT_EQ;
				c->lin0 = ax[i];
				c->ang0 = cross(ra, ax[i]);
				c->lin1 = ax[i];
				c->ang1 = cross(rb, ax[i]);
				c->geomErr = -comp(cB2cAp, i);
			}
	}
	if (ang)
	{
		const qt qB2qA = qmul(qconj(cA.q), cB.q);
		v3 row[3];
		jacobian_axes(row, cA.q, cB.q);
		const float qe[3] = {qB2qA.x, qB2qA.y, qB2qA.z};
		for (int i = 0; i < 3; i++)
			if (ang & (1 << i))
			{
				row1d* c = &rows[k++];
				c->hint = HINT_EQ;
				c->ang0 = row[i];
				c->ang1 = row[i];
				c->geomErr = -qe[i];
				c->angular = 1;
			}
	}
	const v3 axis = axes.c0;
	if (j->type == PXO_JOINT_REVOLUTE && !limitLocked)
	{
		if (j->pflags & JP_DRIVE)
		{
			row1d* c = &rows[k++];
			c->hint = HINT_NONE;
			c->ang0 = neg(axis);
			c->ang1 = scl(neg(axis), j->gear);
			c->velTarget = j->dvel;
			c->minImp = -j->dfl;
			c->maxImp = j->dfl;
			c->angular = 1;
			c->flags |= RF_DRIVE_LIMIT;
			if (j->pflags & JP_FREESPIN)
			{
				if (j->dvel > 0.0f) c->minImp = 0.0f;
				if (j->dvel < 0.0f) c->maxImp = 0.0f;
			}
		}
		if (limitOn)
		{
			/* anglePair (ExtConstraintHelper.h:220-231): a soft limit only
			 * beyond its stop */
			const float phi = twist_angle(cA.q, cB.q);
			const int soft = soft_limit(j);
			if (!soft || phi < j->lo)
			{
				row1d* c = &rows[k++];
				c->ang0 = neg(axis); c->ang1 = neg(axis); c->angular = 1;
				c->geomErr = -(j->lo - phi);
				limit_row(c, j);
			}
			if (!soft || phi > j->hi)
			{
				row1d* c = &rows[k++];
				c->ang0 = axis; c->ang1 = axis; c->angular = 1;
				c->geomErr = j->hi - phi;
				limit_row(c, j);
			}
		}
	}
	if (j->type == PXO_JOINT_PRISMATIC && limitOn && !limitLocked)
	{
		/* linearLimit (ExtConstraintHelper.h:203-207): a soft limit only
		 * beyond its stop */
		const float ordinate = cB2cAp.x;
		const int soft = soft_limit(j);
		if (!soft || ordinate > j->hi)
		{
			row1d* c = &rows[k++];
			c->lin0 = axis; c->ang0 = cross(ra0, axis); c->lin1 = axis; c->ang1 = cross(rb, axis);
			c->geomErr = j->hi - ordinate;
			limit_row(c, j);
		}
		const v3 na = neg(axis);
		if (!soft || -ordinate > -j->lo)
		{
			row1d* c = &rows[k++];
			c->lin0 = na; c->ang0 = cross(ra0, na); c->lin1 = na; c->ang1 = cross(rb, na);
			c->geomErr = -j->lo - -ordinate;
			limit_row(c, j);
		}
	}
	*raOut = ra;
	*rbOut = rb;
	return k;
}

/* orthogonalize (lowleveldynamics/src/DyConstraintSetup.cpp:313-470):
 * Gram-Schmidt in the inverse-mass metric over one group of rows, each row
 * made orthogonal to the group's first eq rows (its equality rows; the
 * inequality rows after them are not made orthogonal to each other),
 * geometric error and target velocity carried along */
static void orthogonalize_rows(row1d** row, v3* as0, v3* as1, int count, int eq, float im0, float im1)
{
	v3 lin0m[JOINT_MAX_ROWS], ang0m[JOINT_MAX_ROWS], lin1m[JOINT_MAX_ROWS], ang1m[JOINT_MAX_ROWS];
	v3 lin0[JOINT_MAX_ROWS], ang0[JOINT_MAX_ROWS], lin1[JOINT_MAX_ROWS], ang1[JOINT_MAX_ROWS];
	float g[JOINT_MAX_ROWS], t[JOINT_MAX_ROWS];
	for (int i = 0; i < count; i++)
	{
		v3 l0 = row[i]->lin0, a0 = row[i]->ang0, l1 = row[i]->lin1, a1 = row[i]->ang1;
		float gi = row[i]->geomErr, ti = row[i]->velTarget;
		v3 s0v = as0[i], s1v = as1[i];
		const int elim = i < eq ? i : eq;
		for (int k = 0; k < elim; k++)
		{
			const v3 sa = add(mulv(l1, lin1m[k]), mulv(l0, lin0m[k]));
			const v3 sb = add(mulv(s1v, ang1m[k]), mulv(s0v, ang0m[k]));
			const v3 ss = add(sa, sb);
			const float tt = ss.x + ss.y + ss.z;
			l0 = sub(l0, scl(lin0[k], tt));
			gi = gi - g[k] * tt;
			a0 = sub(a0, scl(ang0[k], tt));
			ti = ti - t[k] * tt;
			l1 = sub(l1, scl(lin1[k], tt));
			a1 = sub(a1, scl(ang1[k], tt));
			s0v = sub(s0v, scl(as0[k], tt));
			s1v = sub(s1v, scl(as1[k], tt));
		}
		row[i]->lin0 = l0; row[i]->geomErr = gi;
		row[i]->ang0 = a0; row[i]->velTarget = ti;
		row[i]->lin1 = l1;
		row[i]->ang1 = a1;
		as0[i] = s0v;
		as1[i] = s1v;
		lin0[i] = l0; g[i] = gi; ang0[i] = a0; t[i] = ti; lin1[i] = l1; ang1[i] = a1;
		const v3 l0m = scl(l0, im0), l1m = scl(l1, im1);
		const v3 a0m = s0v, a1m = s1v;   /* angular inverse-mass scales are 1 */
		const v3 q0 = add(mulv(l0, l0m), mulv(l1, l1m));
		const v3 q1 = add(mulv(a0m, s0v), mulv(a1m, s1v));
		const v3 qq = add(q0, q1);
		const float sum = qq.x + qq.y + qq.z;
		const float a = sum > 0.0f ? 1.0f / sum : 0.0f;
		lin0m[i] = scl(l0m, a);
		ang0m[i] = scl(a0m, a);
		lin1m[i] = scl(l1m, a);
		ang1m[i] = scl(a1m, a);
	}
}

/* One joint's rows for this step: the solverPrep rows, then
 * setupSolverConstraintStep (lowleveldynamics/src/DyTGSContactPrep.cpp:
 * 1939-2164) with preprocessRows (DyConstraintSetup.cpp:492-597): angular
 * rows take the rotational hints, rows sorted by hint, the equality and
 * inequality groups (linear, rotational) orthogonalized; per row the unit
 * response, the solver constants (compute1dConstraintSolverConstantsTGS,
 * lowleveldynamics/shared/DyCpuGpu1dConstraint.h:418-486: hard rows with
 * erp = half the bias coefficient, bouncing limit rows with their bounce
 * velocity; computeBounceVelocity :141-159 from the pair's velocity along
 * the row at the start of the step), the maximum bias
 * (computeMaxBiasVelocityTGS, :332-363), the flags (raiseInternalFlagsTGS,
 * :607-623) and the orthogonal-basis terms of the rotational equality rows. */
static void prep_joint(pxo_joint* j, xf bA, xf bB, const sbody* sa, const sbody* sb_, float stepDt, float invStepDt,
                       float invDt, float erp, float lengthScale)
{
	row1d rows[JOINT_MAX_ROWS];
	v3 ra, rb;
	const int n = joint_rows(j, bA, bB, rows, &ra, &rb);
	j->count = n;
	j->invMass0 = sa->invMass;
	j->invMass1 = sb_->invMass;
	j->rA = ra;
	j->rB = rb;
	for (int k = 0; k < 3; k++)
	{
		j->o0[k] = V(0, 0, 0); j->o1[k] = V(0, 0, 0);
		j->orecip[k] = 0.0f; j->oerr[k] = 0.0f;
	}
	for (int i = 0; i < n; i++)
		if (rows[i].angular)
		{
			if (rows[i].hint == HINT_EQ) rows[i].hint = HINT_ROT_EQ;
			else if (rows[i].hint == HINT_INEQ) rows[i].hint = HINT_ROT_INEQ;
		}
	/* insertion sort by hint (stable) */
	row1d* sorted[JOINT_MAX_ROWS];
	for (int i = 0; i < n; i++)
	{
		int k = i;
		for (; k > 0 && rows[i].hint < sorted[k - 1]->hint; k--)
			sorted[k] = sorted[k - 1];
		sorted[k] = &rows[i];
	}
	v3 as0[JOINT_MAX_ROWS], as1[JOINT_MAX_ROWS];
	for (int i = 0; i < n; i++)
	{
		as0[i] = mmul(sa->sqrtInvI, sorted[i]->ang0);
		as1[i] = mmul(sb_->sqrtInvI, sorted[i]->ang1);
	}
	for (int i = 0; i < n;)
	{
		const int major = sorted[i]->hint >> 8, start = i++;
		while (i < n && (sorted[i]->hint >> 8) == major)
			i++;
		if (major == 4 || major == 8)
		{
			int eq = start;
			while (eq < i && (sorted[eq]->hint & 255) == 0)
				eq++;
			orthogonalize_rows(sorted + start, as0 + start, as1 + start, i - start, eq - start, sa->invMass,
			                   sb_->invMass);
		}
	}
	int ortho = 0;
	for (int i = 0; i < n; i++)
	{
		const row1d* c = sorted[i];
		jrow* s = &j->r[i];
		s->lin0 = c->lin0; s->lin1 = c->lin1; s->ang0 = c->ang0; s->ang1 = c->ang1;
		/* computeMinMaxImpulseOrForceAsImpulse (DyCpuGpu1dConstraint.h:71-79):
		 * a drive's limits are forces with PxConstraintFlag::
		 * eDRIVE_LIMITS_ARE_FORCES (parameter flag 16) */
		const float driveScale = ((c->flags & RF_DRIVE_LIMIT) && (j->pflags & JP_FORCES)) ? 1.0f / invDt : 1.0f;
		s->minImp = c->minImp * driveScale; s->maxImp = c->maxImp * driveScale;
		s->applied = 0.0f;
		s->flags = JF_ANGULAR_ERROR;
		const float linSum = len2(s->lin0) * sa->invMass + len2(s->lin1) * sb_->invMass;
		const float resp0 = len2(as0[i]), resp1 = len2(as1[i]);
		const float unit = resp0 + resp1 + linSum;
		/* projectVelocity of both bodies on the row as prepared (before the
		 * linear rows' angular terms are cleared) */
		const float vel0 = dot(sa->origLin, s->lin0) + dot(sa->origAng, s->ang0);
		const float vel1 = dot(sb_->origLin, s->lin1) + dot(sb_->origAng, s->ang1);
		const float jointSpeed = vel0 - vel1;
		if (!c->angular)
		{
			s->ang0 = V(0, 0, 0);
			s->ang1 = V(0, 0, 0);
			s->flags &= ~JF_ANGULAR_ERROR;
		}
		const float recip = unit <= 0.0f ? 0.0f : 1.0f / unit;
		float bounce = jointSpeed * -c->rest;
		if (!((c->flags & RF_RESTITUTION) && -jointSpeed > c->bthr && bounce * c->geomErr <= 0.0f))
			bounce = 0.0f;
		s->velMul = -1.0f;
		if (c->flags & RF_SPRING)
		{
			/* a force spring: implicit in the step */
			const float a = stepDt * (stepDt * c->stiff + c->damp);
			const float b = stepDt * (c->damp * c->velTarget);
			const float x = 1.0f / (1.0f + a * unit);
			s->maxBias = 3.402823466e+38f;
			s->biasScale = -x * c->stiff * stepDt;
			s->error = s->biasScale * c->geomErr;
			s->velMul = -x * a;
			s->velTarget = x * b - s->velMul * 0.0f;
			s->flags |= JF_SPRING;
		}
		else if (bounce != 0.0f)
		{
			s->maxBias = 0.0f;
			s->biasScale = 0.0f;
			s->error = 0.0f;
			s->velTarget = bounce - s->velMul * 0.0f;
		}
		else
		{
			s->maxBias = c->angular ? invDt * 0.75f : invDt * 15.0f * lengthScale;
			s->biasScale = -invStepDt * erp;
			s->error = c->geomErr * s->biasScale;
			s->velTarget = c->velTarget - s->velMul * 0.0f;
		}
		if (c->flags & RF_KEEP_BIAS)
			s->flags |= JF_KEEP_BIAS;
		if (c->hint & 1)
			s->flags |= JF_INEQUALITY;
		if (c->hint == HINT_ROT_EQ)
		{
			s->flags |= JF_ROT_EQ;
			j->o0[ortho] = as0[i];
			j->orecip[ortho] = recip;
			j->o1[ortho] = as1[i];
			j->oerr[ortho] = c->geomErr;
			ortho++;
		}
		else if (c->hint & HINT_EQ)
			s->flags |= JF_ORTHO_TARGET;
	}
}

/* solve1DStep (lowleveldynamics/src/DyTGSContactPrep.cpp:2418-2747): the
 * anchor offsets rotated by the bodies' rotation since the step began,
 * linear rows' angular terms recomputed from them, the rotational basis
 * projected out of the linear rows, the error resolved so far from the
 * bodies' accumulated motion, the response recomputed, and the clamped
 * impulse applied; spring rows (soft limits) take their velocity multiplier
 * and constant without the response (:2640-2660) and resolve no target over
 * the elapsed time. conclude: conclude1DStep (:2750-2780; a spring then gives
 * no impulse in the velocity iterations) after the row is
 * solved in the last position iteration. */
static void solve_joint(pxo_joint* j, sbody* b0, sbody* b1, float elapsed, int conclude)
{
	(void)elapsed;   /* target velocities are zero */
	const v3 ra = qrot(b0->dq, j->rA);
	const v3 rb = qrot(b1->dq, j->rB);
	const v3 raMotion = sub(add(ra, b0->dlin), j->rA);
	const v3 rbMotion = sub(add(rb, b1->dlin), j->rB);
	float oerr[3];
	for (int k = 0; k < 3; k++)
		oerr[k] = j->oerr[k] + (dot(j->o0[k], b0->dang) - dot(j->o1[k], b1->dang));
	v3 l0 = b0->lin, l1 = b1->lin, a0 = b0->ang, a1 = b1->ang;
	const float im0 = j->invMass0, im1 = j->invMass1;
	for (int i = 0; i < j->count; i++)
	{
		jrow* c = &j->r[i];
		const v3 cang0 = add(c->ang0, cross(ra, c->lin0));
		const v3 cang1 = add(c->ang1, cross(rb, c->lin1));
		float error = c->error;
		const float minBias = (c->flags & JF_INEQUALITY) ? -3.402823466e+38f : -c->maxBias;
		v3 raXnI = mmul(b0->sqrtInvI, cang0);
		v3 rbXnI = mmul(b1->sqrtInvI, cang1);
		if (c->flags & JF_ORTHO_TARGET)
		{
			float proj[3];
			for (int k = 0; k < 3; k++)
			{
				const v3 pp = add(mulv(raXnI, j->o0[k]), mulv(rbXnI, j->o1[k]));
				proj[k] = (pp.x + pp.y + pp.z) * j->orecip[k];
			}
			const v3 d0 = add(scl(j->o0[0], proj[0]), add(scl(j->o0[1], proj[1]), scl(j->o0[2], proj[2])));
			const v3 d1 = add(scl(j->o1[0], proj[0]), add(scl(j->o1[1], proj[1]), scl(j->o1[2], proj[2])));
			raXnI = sub(raXnI, d0);
			rbXnI = sub(rbXnI, d1);
			const float ob = c->biasScale * (oerr[0] * proj[0] + (oerr[1] * proj[1] + oerr[2] * proj[2]));
			error = error - ob;
		}
		const float angScale = (c->flags & JF_ANGULAR_ERROR) ? 1.0f : 0.0f;
		const float dAng = angScale * (dot(raXnI, b0->dang) - dot(rbXnI, b1->dang));
		const float dLin = dot(c->lin0, raMotion) - dot(c->lin1, rbMotion);
		const int spring = (c->flags & JF_SPRING) != 0;
		const float motion = dLin + dAng;
		const float errorChange = spring ? motion : motion - c->velTarget * elapsed;
		const v3 r0 = mulv(raXnI, raXnI), r1 = mulv(rbXnI, rbXnI);
		const float resp0 = im0 * dot(c->lin0, c->lin0) + (r0.x + r0.y + r0.z);
		const float resp1 = im1 * dot(c->lin1, c->lin1) - -(r1.x + r1.y + r1.z);
		const float response = resp0 + resp1;
		const float recip = response > 0.0f ? 1.0f / response : 0.0f;
		const float vMul = spring ? c->velMul : recip * c->velMul;
		const float ub = errorChange * c->biasScale + error;
		const float bias = fminf(fmaxf(ub, minBias), c->maxBias);
		const float constant = spring ? bias + c->velTarget : recip * (bias + c->velTarget);
		const v3 v0 = add(mulv(l0, c->lin0), mulv(a0, raXnI));
		const v3 v1 = add(mulv(l1, c->lin1), mulv(a1, rbXnI));
		const v3 dv = sub(v0, v1);
		const float normalVel = dv.x + dv.y + dv.z;
		const float unclamped = c->applied + (vMul * normalVel + constant);
		const float clamped = fminf(fmaxf(unclamped, c->minImp), c->maxImp);
		const float df = clamped - c->applied;
		c->applied = clamped;
		l0 = add(scl(c->lin0, df * im0), l0);
		l1 = sub(l1, scl(c->lin1, df * im1));
		a0 = add(scl(raXnI, df), a0);
		a1 = add(scl(rbXnI, df * -1.0f), a1);
		if (conclude && !(c->flags & JF_KEEP_BIAS))
		{
			c->biasScale = 0.0f;
			c->error = 0.0f;
		}
		if (conclude && spring)
		{
			/* no spring impulse in the velocity iterations */
			c->velMul = 0.0f;
			c->biasScale = 0.0f;
			c->error = 0.0f;
			c->velTarget = 0.0f;
		}
	}
	b0->lin = l0; b0->ang = a0;
	b1->lin = l1; b1->ang = a1;
}

static void joint_keys(pxo_world* w);

/* limit and drive parameters of k joints (params[10 * i]: flags JP_* as a
 * float, lower, upper, restitution, bounce threshold, drive velocity, drive
 * force limit, drive gear ratio, limit stiffness, limit damping). Only
 * revolute and prismatic joints take them, prismatic joints no drive.
 * Returns 0, or -1 on a bad argument. */
int pxo_world_set_joint_params(void* wp, int k, const int32_t* joints, const float* params)
{
	pxo_world* w = (pxo_world*)wp;
	for (int i = 0; i < k; i++)
	{
		if (joints[i] < 0 || joints[i] >= w->nj)
			return -1;
		const int t = w->joint[joints[i]].type;
		const int f = (int)params[10 * i];
		if ((f & 7) && t != PXO_JOINT_REVOLUTE && t != PXO_JOINT_PRISMATIC && t != PXO_JOINT_SPHERICAL &&
		    t != PXO_JOINT_DISTANCE)
			return -1;
		if ((f & (JP_DRIVE | JP_FREESPIN)) && t != PXO_JOINT_REVOLUTE && t != PXO_JOINT_DISTANCE)
			return -1;
	}
	for (int i = 0; i < k; i++)
	{
		pxo_joint* j = &w->joint[joints[i]];
		const float* q = params + 10 * i;
		j->pflags = (int)q[0];
		j->lo = q[1]; j->hi = q[2]; j->rest = q[3]; j->bthr = q[4];
		j->dvel = q[5]; j->dfl = q[6]; j->gear = q[7]; j->stiff = q[8]; j->damp = q[9];
	}
	joint_keys(w);
	return 0;
}

/* PxAggregate (physx/include/PxAggregate.h): the members of an aggregate
 * without self-collisions never pair with each other (the reference's
 * broadphase sees the aggregate's bounds and its own self-collision pass is
 * off); every other pair is found as before */
int pxo_world_add_aggregate(void* wp, int k, const int32_t* bodies, int self_collisions)
{
	pxo_world* w = (pxo_world*)wp;
	if (!w->agg)
	{
		w->agg = (int*)malloc(sizeof(int) * (size_t)(w->n ? w->n : 1));
		for (int i = 0; i < w->n; i++)
			w->agg[i] = -1;
	}
	for (int i = 0; i < k; i++)
		if (bodies[i] < 0 || bodies[i] >= w->n || w->agg[bodies[i]] >= 0)
			return -1;
	w->aggSelf = (char*)realloc(w->aggSelf, (size_t)(w->nagg + 1));
	w->aggSelf[w->nagg] = (char)(self_collisions != 0);
	for (int i = 0; i < k; i++)
		w->agg[bodies[i]] = w->nagg;
	return w->nagg++;
}

static int aggregate_filtered(const pxo_world* w, int i, int j)
{
	if (!w->agg)
		return 0;
	const int a = w->agg[i];
	return a >= 0 && a == w->agg[j] && !w->aggSelf[a];
}

int pxo_world_add_joints(void* wp, int k, const int32_t* type, const int32_t* body0, const int32_t* body1,
                         const float* frame0, const float* frame1)
{
	pxo_world* w = (pxo_world*)wp;
	if (k <= 0)
		return w->nj;
	for (int i = 0; i < k; i++)
	{
		if (type[i] < 0 || type[i] > PXO_JOINT_DISTANCE || body0[i] >= w->n || body1[i] >= w->n ||
		    (body0[i] < 0 && body1[i] < 0) || body0[i] == body1[i] || body0[i] < -1 || body1[i] < -1)
			return -1;
	}
	if (w->nj + k > w->jcap)
	{
		w->jcap = (w->nj + k) * 2;
		w->joint = (pxo_joint*)realloc(w->joint, sizeof(pxo_joint) * (size_t)w->jcap);
		w->jkeys = (uint64_t*)realloc(w->jkeys, sizeof(uint64_t) * (size_t)w->jcap);
	}
	const int first = w->nj;
	for (int i = 0; i < k; i++)
	{
		pxo_joint* j = &w->joint[w->nj++];
		memset(j, 0, sizeof(*j));
		j->gear = 1.0f;
		j->type = type[i];
		if (j->type == PXO_JOINT_DISTANCE)
		{
			/* PxDistanceJointCreate's defaults (ExtDistanceJoint.cpp:37-47):
			 * max distance 0 enforced, tolerance 0.025 */
			j->pflags = 1;
			j->rest = 0.025f;
		}
		j->a = body0[i];
		j->b = body1[i];
		j->f0.p = V(frame0[7 * i], frame0[7 * i + 1], frame0[7 * i + 2]);
		j->f0.q.x = frame0[7 * i + 3]; j->f0.q.y = frame0[7 * i + 4]; j->f0.q.z = frame0[7 * i + 5]; j->f0.q.w = frame0[7 * i + 6];
		j->f1.p = V(frame1[7 * i], frame1[7 * i + 1], frame1[7 * i + 2]);
		j->f1.q.x = frame1[7 * i + 3]; j->f1.q.y = frame1[7 * i + 4]; j->f1.q.z = frame1[7 * i + 5]; j->f1.q.w = frame1[7 * i + 6];
		/* each jointed body counts the joint among its interactions
		 * (BodySim::onConstraintAttach, ScBodySim.h:159) */
		if (j->a >= 0) w->jcount[j->a]++;
		if (j->b >= 0) w->jcount[j->b]++;
	}
	/* colours (the GPU host computes the same): joints between two bodies
	 * first fit, newest joint first (the order in which the reference's
	 * island lists them: its partitions of a jointed chain come out so),
	 * over the colours each body's joints hold; then each joint to the world
	 * takes the colour after every colour of its body's other joints, one
	 * colour per such joint in joint order, as the reference partitions
	 * constraints against static bodies after their body's dynamic ones
	 * (lowleveldynamics/src/DyConstraintPartition.cpp:168-179, 604-616) */
	uint64_t* used = (uint64_t*)calloc((size_t)(w->n ? w->n : 1), sizeof(uint64_t));
	int* next = (int*)calloc((size_t)(w->n ? w->n : 1), sizeof(int));
	w->njcolors = 0;
	for (int i = w->nj - 1; i >= 0; i--)
	{
		pxo_joint* j = &w->joint[i];
		if (j->a < 0 || j->b < 0)
			continue;
		const uint64_t u = used[j->a] | used[j->b];
		if (u == ~0ull)
		{
			free(used);
			free(next);
			w->nj = first;
			return -1;
		}
		int c = 0;
		while ((u >> c) & 1ull)
			c++;
		j->color = c;
		used[j->a] |= 1ull << c;
		used[j->b] |= 1ull << c;
		if (c + 1 > next[j->a]) next[j->a] = c + 1;
		if (c + 1 > next[j->b]) next[j->b] = c + 1;
		if (c + 1 > w->njcolors) w->njcolors = c + 1;
	}
	for (int i = 0; i < w->nj; i++)
	{
		pxo_joint* j = &w->joint[i];
		if (j->a >= 0 && j->b >= 0)
			continue;
		const int body = j->a >= 0 ? j->a : j->b;
		j->color = next[body]++;
		if (j->color + 1 > w->njcolors) w->njcolors = j->color + 1;
	}
	free(used);
	free(next);
	joint_keys(w);
	return first;
}

/* the sorted body-pair keys of the joints whose bodies do not collide
 * (PxConstraintFlag::eCOLLISION_ENABLED off, the default; flag 8 of the
 * joint parameters turns it on) */
static void joint_keys(pxo_world* w)
{
	int nk = 0;
	for (int i = 0; i < w->nj; i++)
	{
		const pxo_joint* j = &w->joint[i];
		if (j->a < 0 || j->b < 0 || (j->pflags & JP_COLLIDE))
			continue;
		const uint32_t lo = (uint32_t)(j->a < j->b ? j->a : j->b), hi = (uint32_t)(j->a < j->b ? j->b : j->a);
		w->jkeys[nk++] = ((uint64_t)lo << 32) | hi;
	}
	qsort(w->jkeys, (size_t)nk, sizeof(uint64_t), cmp_u64);
	/* jointed() searches nj keys: pad with keys no pair has */
	for (int i = nk; i < w->nj; i++)
		w->jkeys[i] = ~0ull;
}

/* ---- static triangle meshes ----
 * Contacts of a body against a static triangle mesh, restated after the
 * reference's PCM mesh contact generation (geomutils/src/pcm/
 * GuPCMContactSphereMesh.cpp:160-501, GuPCMTriangleContactGen.cpp,
 * GuPCMContactConvexCommon.h:131-230, GuPersistentContactManifold.cpp:
 * 1315-1420, 1678-2160). Triangles whose bounds meet the body's bounds
 * inflated by the contact distance, in triangle order; a triangle whose plane
 * has the body's centre behind it is culled (one-sided meshes). Per triangle:
 * a sphere takes the closest point (face contacts at once; edge and vertex
 * contacts deferred, by distance, and dropped when a face or a closer contact
 * already holds that edge or vertex); a capsule the part of its segment over
 * the triangle (else the closest edge point, deferred like a sphere's); a box
 * or hull its face most opposed to the triangle normal clipped to the
 * triangle's prism, and its deepest vertex when over the triangle. Contacts
 * group into patches by normal (within 0.996, about 5 degrees, the reference's
 * acceptance); a patch holds TM_PC contacts before it is reduced to 4 (the
 * deepest, the farthest from it, then the extremes across their line: the
 * first four picks of SinglePersistentContactManifold::reduceContacts); the
 * deepest TM_MAXP patches become one constraint each (a sphere keeps one
 * contact per patch, GU_SPHERE_MANIFOLD_CACHE_SIZE). */
#define TM_MAX_TRIS 128
#define TM_MAXP 4
#define TM_PC 8
#define TM_NPATCH 6

typedef struct { v3 p; float s; } tmc_t;
typedef struct { v3 n; int count; float maxPen; tmc_t c[TM_PC + 1]; } tpatch_t;
typedef struct { tpatch_t p[TM_NPATCH]; int np; } tpatches_t;

/* reduction of a patch's contacts to 4 */
static int tm_reduce4(tmc_t* c, int n, v3 nrm)
{
	if (n <= 4)
		return n;
	int used[TM_PC + 1];
	for (int i = 0; i < n; i++)
		used[i] = 0;
	int i0 = 0;
	for (int i = 1; i < n; i++)
		if (c[i0].s > c[i].s)
			i0 = i;
	used[i0] = 1;
	int i1 = -1;
	float best = -1.0f;
	for (int i = 0; i < n; i++)
	{
		if (used[i]) continue;
		const v3 d = sub(c[i].p, c[i0].p);
		const float dd = dot(d, d);
		if (dd > best) { best = dd; i1 = i; }
	}
	used[i1] = 1;
	const v3 ax = cross(sub(c[i1].p, c[i0].p), nrm);
	int i2 = -1, i3 = -1;
	float mx = -FLT_MAX, mn = FLT_MAX;
	for (int i = 0; i < n; i++)
	{
		if (used[i]) continue;
		const float d = dot(sub(c[i].p, c[i0].p), ax);
		if (d > mx) { mx = d; i2 = i; }
		if (mn > d) { mn = d; i3 = i; }
	}
	used[i2] = 1;
	if (i3 == i2 || mx * mn > 0.0f)
	{
		/* both on one side of the line (or one candidate): the next largest */
		mx = -FLT_MAX;
		i3 = -1;
		for (int i = 0; i < n; i++)
		{
			if (used[i]) continue;
			const float d = dot(sub(c[i].p, c[i0].p), ax);
			if (d > mx) { mx = d; i3 = i; }
		}
	}
	const tmc_t a = c[i0], b = c[i1], e = c[i2], f = c[i3];
	c[0] = a; c[1] = b; c[2] = e; c[3] = f;
	return 4;
}

/* adds a contact (point, separation) with normal n to its patch */
static void tm_add(tpatches_t* P, v3 n, v3 p, float s)
{
	for (int k = 0; k < P->np; k++)
	{
		tpatch_t* q = &P->p[k];
		if (!(dot(q->n, n) > 0.996f))
			continue;
		for (int j = 0; j < q->count; j++)
		{
			const v3 d = sub(q->c[j].p, p);
			if (1e-10f > dot(d, d))
			{
				if (q->c[j].s > s)
				{
					q->c[j].p = p;
					q->c[j].s = s;
				}
				if (q->maxPen > s)
					q->maxPen = s;
				return;
			}
		}
		q->c[q->count].p = p;
		q->c[q->count].s = s;
		q->count++;
		if (q->maxPen > s)
			q->maxPen = s;
		if (q->count > TM_PC)
			q->count = tm_reduce4(q->c, q->count, q->n);
		return;
	}
	int k = P->np;
	if (k == TM_NPATCH)
	{
		/* no room: a deeper contact replaces the shallowest patch */
		k = 0;
		for (int j = 1; j < TM_NPATCH; j++)
			if (P->p[j].maxPen > P->p[k].maxPen)
				k = j;
		if (!(P->p[k].maxPen > s))
			return;
	}
	else
		P->np++;
	P->p[k].n = n;
	P->p[k].count = 1;
	P->p[k].maxPen = s;
	P->p[k].c[0].p = p;
	P->p[k].c[0].s = s;
}

/* pcmDistancePointTriangleSquared (GuPCMContactSphereMesh.cpp:160-288):
 * closest point of triangle abc to p and its feature (0 face, 1-3 vertex a,
 * b, c, 4 edge ab, 5 edge bc, 6 edge ca) */
enum { TF_FACE = 0, TF_V0 = 1, TF_V1 = 2, TF_V2 = 3, TF_E01 = 4, TF_E12 = 5, TF_E20 = 6 };
static float tm_point_tri(v3 p, v3 a, v3 b, v3 c, v3* cp, int* fc)
{
	const v3 ab = sub(b, a), ac = sub(c, a), bc = sub(c, b);
	const v3 ap = sub(p, a), bp = sub(p, b), cq = sub(p, c);
	const float d1 = dot(ab, ap), d2 = dot(ac, ap), d3 = dot(ab, bp), d4 = dot(ac, bp), d5 = dot(ab, cq),
	            d6 = dot(ac, cq);
	const float unom = d4 - d3, udenom = d5 - d6;
	const v3 n = cross(ab, ac);
	if (0.0f > d1 && 0.0f > d2) { *fc = TF_V0; *cp = a; return dot(ap, ap); }
	if (d3 >= 0.0f && d3 >= d4) { *fc = TF_V1; *cp = b; return dot(bp, bp); }
	if (d6 >= 0.0f && d6 >= d5) { *fc = TF_V2; *cp = c; return dot(cq, cq); }
	const float vc = dot(n, cross(ap, bp));
	if (0.0f > vc && d1 >= 0.0f && 0.0f > d3)
	{
		*fc = TF_E01;
		const v3 q = add(scl(ab, d1 / (d1 - d3)), a);
		*cp = q;
		const v3 vv = sub(p, q);
		return dot(vv, vv);
	}
	const float va = dot(n, cross(bp, cq));
	if (0.0f > va && d4 >= d3 && d5 >= d6)
	{
		*fc = TF_E12;
		const v3 q = add(scl(bc, unom / (unom + udenom)), b);
		*cp = q;
		const v3 vv = sub(p, q);
		return dot(vv, vv);
	}
	const float vb = dot(n, cross(cq, ap));
	if (0.0f > vb && d2 >= 0.0f && 0.0f > d6)
	{
		*fc = TF_E20;
		const v3 q = add(scl(ac, d2 / (d2 - d6)), a);
		*cp = q;
		const v3 vv = sub(p, q);
		return dot(vv, vv);
	}
	*fc = TF_FACE;
	const float t = dot(n, sub(a, p)) / dot(n, n);
	const v3 vv = scl(n, t);
	*cp = add(p, vv);
	return dot(vv, vv);
}

/* edge and vertex caches of the deferred contacts (CachedEdge / CachedVertex,
 * geomutils/src/mesh/GuTriangleCache.h:38-90): add returns 1 when new */
typedef struct { uint64_t e[3 * TM_MAX_TRIS]; int ne; int v[3 * TM_MAX_TRIS]; int nv; } tm_cache_t;
static int tm_cache_edge(tm_cache_t* c, int a, int b)
{
	const uint64_t k = ((uint64_t)(uint32_t)(a < b ? a : b) << 32) | (uint32_t)(a < b ? b : a);
	for (int i = 0; i < c->ne; i++)
		if (c->e[i] == k) return 0;
	c->e[c->ne++] = k;
	return 1;
}
static int tm_cache_vertex(tm_cache_t* c, int a)
{
	for (int i = 0; i < c->nv; i++)
		if (c->v[i] == a) return 0;
	c->v[c->nv++] = a;
	return 1;
}

typedef struct { v3 p, n; float sq; int fc, t; } tm_def_t;

/* deferred edge and vertex contacts by distance, each unless its feature is
 * already held (PCMSphereVsMeshContactGeneration::generateLastContacts,
 * GuPCMContactSphereMesh.cpp:400-501); r: the sphere or capsule radius */
static void tm_flush_deferred(const pxo_world* w, tm_def_t* d, int nd, tm_cache_t* cache, float r, tpatches_t* P)
{
	for (int i = 1; i < nd; i++)   /* stable insertion sort by squared distance */
	{
		const tm_def_t x = d[i];
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
		const tri_t* t = &w->tri[d[i].t];
		const int e01 = tm_cache_edge(cache, t->vi[0], t->vi[1]);
		const int e12 = tm_cache_edge(cache, t->vi[1], t->vi[2]);
		const int e20 = tm_cache_edge(cache, t->vi[2], t->vi[0]);
		const int v0 = tm_cache_vertex(cache, t->vi[0]);
		const int v1 = tm_cache_vertex(cache, t->vi[1]);
		const int v2 = tm_cache_vertex(cache, t->vi[2]);
		const int fc = d[i].fc;
		const int need = fc == TF_V0 ? v0 : fc == TF_V1 ? v1 : fc == TF_V2 ? v2 : fc == TF_E01 ? e01 : fc == TF_E12 ? e12 : e20;
		if (need)
		{
			const float dist = sqrtf(d[i].sq);
			tm_add(P, d[i].n, sub(d[i].p, scl(d[i].n, r)), dist - r);
		}
	}
}

static void tm_cache_tri(tm_cache_t* cache, const tri_t* t)
{
	tm_cache_edge(cache, t->vi[0], t->vi[1]);
	tm_cache_edge(cache, t->vi[1], t->vi[2]);
	tm_cache_edge(cache, t->vi[2], t->vi[0]);
	tm_cache_vertex(cache, t->vi[0]);
	tm_cache_vertex(cache, t->vi[1]);
	tm_cache_vertex(cache, t->vi[2]);
}

/* one-sided: the body's centre must not be behind the triangle's plane */
static int tm_culled(const tri_t* t, v3 c) { return 0.0f > dot(c, t->n) - dot(t->v[0], t->n); }

static void tm_sphere(const pxo_world* w, const int* idx, int nc, v3 c, float r, float cd, tpatches_t* P)
{
	const float infl = r + cd;
	const float inflSq = infl * infl;
	static tm_cache_t cache;
	static tm_def_t def[TM_MAX_TRIS];
	cache.ne = cache.nv = 0;
	int nd = 0;
	for (int k = 0; k < nc; k++)
	{
		const tri_t* t = &w->tri[idx[k]];
		if (tm_culled(t, c))
			continue;
		v3 cp;
		int fc;
		const float sq = tm_point_tri(c, t->v[0], t->v[1], t->v[2], &cp, &fc);
		if (!(inflSq > sq))
			continue;
		if (fc == TF_FACE)
		{
			tm_cache_tri(&cache, t);
			const float dist = sqrtf(sq);
			tm_add(P, t->n, sub(c, scl(t->n, r)), dist - r);
		}
		else
		{
			def[nd].p = c;
			def[nd].n = normalize(sub(c, cp));
			def[nd].sq = sq;
			def[nd].fc = fc;
			def[nd].t = idx[k];
			nd++;
		}
	}
	/* deferred points are reported on the sphere: centre - n r */
	tm_flush_deferred(w, def, nd, &cache, r, P);
}

/* inward normals of the triangle's edge planes (the prism over it) */
static void tm_prism(const tri_t* t, v3* m)
{
	for (int k = 0; k < 3; k++)
		m[k] = cross(t->n, sub(t->v[(k + 1) % 3], t->v[k]));
}

static void tm_capsule(const pxo_world* w, const int* idx, int nc, v3 p0, v3 p1, v3 c, float r, float cd,
                       tpatches_t* P)
{
	const float infl = r + cd;
	static tm_cache_t cache;
	static tm_def_t def[TM_MAX_TRIS];
	cache.ne = cache.nv = 0;
	int nd = 0;
	for (int k = 0; k < nc; k++)
	{
		const tri_t* t = &w->tri[idx[k]];
		if (tm_culled(t, c))
			continue;
		v3 m[3];
		tm_prism(t, m);
		/* the part of the segment over the triangle */
		float t0 = 0.0f, t1 = 1.0f;
		int empty = 0;
		for (int e = 0; e < 3 && !empty; e++)
		{
			const float a = dot(m[e], sub(p0, t->v[e])), b = dot(m[e], sub(p1, t->v[e]));
			if (0.0f > a && 0.0f > b)
				empty = 1;
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
			const v3 d = sub(p1, p0);
			const v3 q[2] = {add(p0, scl(d, t0)), add(p0, scl(d, t1))};
			const int nq = t1 > t0 ? 2 : 1;
			int any = 0;
			for (int j = 0; j < nq; j++)
			{
				const float s = dot(sub(q[j], t->v[0]), t->n);
				if (infl > s)
				{
					tm_add(P, t->n, sub(q[j], scl(t->n, r)), s - r);
					any = 1;
				}
			}
			if (any)
				tm_cache_tri(&cache, t);
			continue;
		}
		/* beside the triangle: the closest point of its boundary */
		float best = FLT_MAX;
		v3 bs = p0, bt = p0;
		int bf = TF_E01;
		for (int e = 0; e < 3; e++)
		{
			v3 cs, ct;
			const float sq = seg_closest2(p0, p1, t->v[e], t->v[(e + 1) % 3], &cs, &ct);
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
		const v3 nn = normalize(sub(bs, bt));
		if (!(dot(nn, t->n) > 0.0f))
			continue;
		def[nd].p = bs;
		def[nd].n = nn;
		def[nd].sq = best;
		def[nd].fc = bf;
		def[nd].t = idx[k];
		nd++;
	}
	tm_flush_deferred(w, def, nd, &cache, r, P);
}

/* box or hull: the face most opposed to the triangle normal, clipped to the
 * triangle's prism (Sutherland-Hodgman), and the deepest vertex over it */
static void tm_poly(const pxo_world* w, const int* idx, int nc, const polyd* H, xf pose, float cd, tpatches_t* P)
{
	const m33 R = qmat(pose.q);
	for (int k = 0; k < nc; k++)
	{
		const tri_t* t = &w->tri[idx[k]];
		if (tm_culled(t, pose.p))
			continue;
		int f = 0;
		float best = FLT_MAX;
		for (int j = 0; j < H->np; j++)
		{
			const float d = dot(mmul(R, H->pn[j]), t->n);
			if (best > d) { best = d; f = j; }
		}
		v3 buf0[64 + 3], buf1[64 + 3];
		int n0 = H->pc[f] > 64 ? 64 : H->pc[f];
		for (int j = 0; j < n0; j++)
			buf0[j] = add(mmul(R, H->v[H->vr[H->pf[f] + j]]), pose.p);
		v3 m[3];
		tm_prism(t, m);
		v3* in = buf0;
		v3* out = buf1;
		for (int e = 0; e < 3 && n0 > 0; e++)
		{
			int n1 = 0;
			for (int j = 0; j < n0; j++)
			{
				const v3 a = in[j], b = in[(j + 1) % n0];
				const float da = dot(m[e], sub(a, t->v[e])), db = dot(m[e], sub(b, t->v[e]));
				if (da >= 0.0f)
					out[n1++] = a;
				if ((da >= 0.0f) != (db >= 0.0f))
					out[n1++] = add(a, scl(sub(b, a), da / (da - db)));
			}
			v3* tmp = in;
			in = out;
			out = tmp;
			n0 = n1;
		}
		for (int j = 0; j < n0; j++)
		{
			const float s = dot(sub(in[j], t->v[0]), t->n);
			if (cd > s)
				tm_add(P, t->n, sub(in[j], scl(t->n, s)), s);
		}
		/* the deepest vertex, when over the triangle */
		int vbest = 0;
		float dbest = FLT_MAX;
		for (int j = 0; j < H->nv; j++)
		{
			const float d = dot(mmul(R, H->v[j]), t->n);
			if (dbest > d) { dbest = d; vbest = j; }
		}
		const v3 sv = add(mmul(R, H->v[vbest]), pose.p);
		if (dot(m[0], sub(sv, t->v[0])) >= 0.0f && dot(m[1], sub(sv, t->v[1])) >= 0.0f &&
		    dot(m[2], sub(sv, t->v[2])) >= 0.0f)
		{
			const float s = dot(sub(sv, t->v[0]), t->n);
			if (cd > s)
				tm_add(P, t->n, sub(sv, scl(t->n, s)), s);
		}
	}
}

/* a body's contact patches against its world's mesh: constraints with the
 * static index -1 - (plane count + patch); returns -1 when more than
 * TM_MAX_TRIS triangles meet the body's bounds */
static int tm_body(pxo_world* w, int i, const tmesh_t* M, float cd)
{
	v3 mn, mx;
	shape_bounds(w->type[i], w->dims + 4 * i, w->pose[i], &mn, &mx);
	mn = sub(mn, V(cd, cd, cd));
	mx = add(mx, V(cd, cd, cd));
	int idx[TM_MAX_TRIS];
	int nc = 0;
	for (int k = M->first; k < M->first + M->count; k++)
	{
		const tri_t* t = &w->tri[k];
		if (t->lo.x > mx.x || mn.x > t->hi.x || t->lo.y > mx.y || mn.y > t->hi.y || t->lo.z > mx.z || mn.z > t->hi.z)
			continue;
		if (nc == TM_MAX_TRIS)
			return -1;
		idx[nc++] = k;
	}
	if (nc == 0)
		return 0;
	tpatches_t P;
	P.np = 0;
	const int type = w->type[i];
	if (type == PXO_SPHERE)
		tm_sphere(w, idx, nc, w->pose[i].p, w->dims[4 * i], cd, &P);
	else if (type == PXO_CAPSULE)
	{
		v3 s, e;
		capsule_segment(w->pose[i], w->dims[4 * i + 1], &s, &e);
		tm_capsule(w, idx, nc, s, e, w->pose[i].p, w->dims[4 * i], cd, &P);
	}
	else if (type == PXO_CONVEX)
		tm_poly(w, idx, nc, hull_of(w->dims[4 * i]), w->pose[i], cd, &P);
	else
	{
		boxpoly b;
		box_poly(V(w->dims[4 * i], w->dims[4 * i + 1], w->dims[4 * i + 2]), &b);
		tm_poly(w, idx, nc, &b.d, w->pose[i], cd, &P);
	}
	/* deepest patches first (stable) */
	int order[TM_NPATCH];
	for (int k = 0; k < P.np; k++)
	{
		int j = k - 1;
		while (j >= 0 && P.p[order[j]].maxPen > P.p[k].maxPen)
		{
			order[j + 1] = order[j];
			j--;
		}
		order[j + 1] = k;
	}
	const int keep = P.np < TM_MAXP ? P.np : TM_MAXP;
	for (int k = 0; k < keep; k++)
	{
		tpatch_t* q = &P.p[order[k]];
		int cnt = tm_reduce4(q->c, q->count, q->n);
		if (type == PXO_SPHERE)
		{
			int d = 0;
			for (int j = 1; j < cnt; j++)
				if (q->c[d].s > q->c[j].s) d = j;
			q->c[0] = q->c[d];
			cnt = 1;
		}
		cset c;
		c.n = q->n;
		c.count = cnt;
		for (int j = 0; j < cnt; j++)
		{
			c.p[j] = q->c[j].p;
			c.sep[j] = q->c[j].s;
		}
		push_con(w, i, -1 - (w->m + k), &c);
	}
	return keep;
}

int pxo_world_add_trimesh(void* wp, int nv, const float* verts, int nt, const int32_t* tris, const float* pose,
                          int32_t scene)
{
	pxo_world* w = (pxo_world*)wp;
	for (int k = 0; k < w->ntm; k++)
		if (w->tm[k].scene == scene)
			return -1;   /* one mesh per world */
	for (int k = 0; k < 3 * nt; k++)
		if (tris[k] < 0 || tris[k] >= nv)
			return -2;
	const xf t = load_pose(pose);
	v3* P = (v3*)malloc(sizeof(v3) * (size_t)(nv ? nv : 1));
	for (int k = 0; k < nv; k++)
		P[k] = add(qrot(t.q, V(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2])), t.p);
	w->tri = (tri_t*)realloc(w->tri, sizeof(tri_t) * (size_t)(w->ntri + nt + 1));
	const int first = w->ntri;
	for (int k = 0; k < nt; k++)
	{
		tri_t* r = &w->tri[w->ntri];
		for (int j = 0; j < 3; j++)
		{
			r->v[j] = P[tris[3 * k + j]];
			r->vi[j] = w->ntmverts + tris[3 * k + j];
		}
		const v3 n = cross(sub(r->v[1], r->v[0]), sub(r->v[2], r->v[0]));
		if (!(dot(n, n) > 0.0f))
			continue;   /* degenerate: not stored */
		r->n = normalize(n);
		r->lo = r->v[0];
		r->hi = r->v[0];
		for (int j = 1; j < 3; j++)
		{
			r->lo = V(fminf(r->lo.x, r->v[j].x), fminf(r->lo.y, r->v[j].y), fminf(r->lo.z, r->v[j].z));
			r->hi = V(fmaxf(r->hi.x, r->v[j].x), fmaxf(r->hi.y, r->v[j].y), fmaxf(r->hi.z, r->v[j].z));
		}
		w->ntri++;
	}
	free(P);
	w->ntmverts += nv;
	w->tm = (tmesh_t*)realloc(w->tm, sizeof(tmesh_t) * (size_t)(w->ntm + 1));
	tmesh_t* M = &w->tm[w->ntm];
	M->scene = scene;
	M->first = first;
	M->count = w->ntri - first;
	/* broadphase bounds: the local box about the pose (Gu::computeBounds of a
	 * triangle mesh shape, geomutils/src/GuBounds.cpp:354-395) */
	{
		v3 lmin = V(3.4e38f, 3.4e38f, 3.4e38f), lmax = V(-3.4e38f, -3.4e38f, -3.4e38f);
		for (int k = 0; k < nv; k++)
		{
			lmin = V(fminf(lmin.x, verts[3 * k]), fminf(lmin.y, verts[3 * k + 1]), fminf(lmin.z, verts[3 * k + 2]));
			lmax = V(fmaxf(lmax.x, verts[3 * k]), fmaxf(lmax.y, verts[3 * k + 1]), fmaxf(lmax.z, verts[3 * k + 2]));
		}
		const v3 c = V(0.5f * (lmin.x + lmax.x), 0.5f * (lmin.y + lmax.y), 0.5f * (lmin.z + lmax.z));
		const v3 e = V(0.5f * (lmax.x - lmin.x), 0.5f * (lmax.y - lmin.y), 0.5f * (lmax.z - lmin.z));
		const v3 wc = add(qrot(t.q, c), t.p);
		const m33 R = qmat(t.q);
		const v3 we = V(fabsf(R.c0.x) * e.x + fabsf(R.c1.x) * e.y + fabsf(R.c2.x) * e.z,
		                fabsf(R.c0.y) * e.x + fabsf(R.c1.y) * e.y + fabsf(R.c2.y) * e.z,
		                fabsf(R.c0.z) * e.x + fabsf(R.c1.z) * e.y + fabsf(R.c2.z) * e.z);
		M->blo = sub(wc, we);
		M->bhi = add(wc, we);
	}
	return w->ntm++;
}

/* ---- sleeping ----
 * Islands, wake-up and deactivation follow the reference CPU pipeline's
 * observable order (simulationcontroller/src/ScPipeline.cpp:1472-1776,
 * lowlevel/software/src/PxsIslandSim.cpp:1630-1720): the narrowphase runs on
 * every pair with an awake body; a touching pair of an awake and a sleeping
 * body wakes the sleeping body's whole island; then an island whose bodies
 * were all ready for sleep after the last step goes to sleep before the solve
 * (velocities zeroed, BodySim::deactivate, ScBodySim.cpp:469-490), and the
 * solve and integration skip sleeping bodies. */

/* A body counts a plane among its interactions when their broadphase bounds
 * overlap: the plane's bounds are the half-space behind it when its normal is
 * a coordinate axis, else the whole space (computePlaneBounds,
 * geomutils/src/GuBounds.cpp:210-250), each inflated by its contact offset. */
static int plane_counted(const pxo_world* w, int i, int q, float co)
{
	const v3 nr = w->pn[q];
	const float ax = fabsf(nr.x), ay = fabsf(nr.y), az = fabsf(nr.z);
	const float e = 1e-6f, one = 1.0f - e;
	int axis = -1;
	if (ax > one && ay < e && az < e) axis = 0;
	else if (ax < e && ay > one && az < e) axis = 1;
	else if (ax < e && ay < e && az > one) axis = 2;
	if (axis < 0)
		return 1;
	v3 mn, mx;
	shape_bounds(w->type[i], w->dims + 4 * i, w->pose[i], &mn, &mx);
	const float lo = (axis == 0 ? mn.x : axis == 1 ? mn.y : mn.z) - co;
	const float hi = (axis == 0 ? mx.x : axis == 1 ? mx.y : mx.z) + co;
	const float c = axis == 0 ? nr.x : axis == 1 ? nr.y : nr.z;
	return c > 0.0f ? lo <= -w->pd[q] + co : hi >= w->pd[q] - co;
}

/* drops the constraints of sleeping bodies (their contact managers are
 * inactive); returns the number kept */
static void drop_sleeping(pxo_world* w)
{
	long long k = 0;
	for (long long c = 0; c < w->ncon; c++)
	{
		const pxo_con* x = &w->con[c];
		const int off = x->b >= 0 ? (w->asleep[x->a] && w->asleep[x->b]) : w->asleep[x->a];
		if (!off)
			w->con[k++] = *x;
	}
	w->ncon = k;
}

static int isl_find(int* par, int x)
{
	while (par[x] != x)
	{
		par[x] = par[par[x]];
		x = par[x];
	}
	return x;
}

static void isl_union(int* par, int a, int b)
{
	a = isl_find(par, a);
	b = isl_find(par, b);
	if (a < b) par[b] = a;
	else if (b < a) par[a] = b;
}

static int joint_asleep(const pxo_world* w, const pxo_joint* j)
{
	return (j->a < 0 || w->asleep[j->a]) && (j->b < 0 || w->asleep[j->b]);
}

/* wake-up, islands and deactivation between the narrowphase and the solve */
static void sleep_islands(pxo_world* w)
{
	const int n = w->n;
	/* a touching pair of an awake and a sleeping body wakes the sleeping
	 * body's island (setEdgesConnected, secondPassIslandGen, wakeObjectsUp:
	 * ScPipeline.cpp:1726-1776); so do islands named by a wake-up. The woken
	 * bodies keep their wake counter and are not ready for sleep
	 * (IslandSim::activateNode clears the flag, PxsIslandSim.cpp:400-432) */
	for (long long c = 0; c < w->ncon; c++)
	{
		const pxo_con* x = &w->con[c];
		if (x->b >= 0 && w->asleep[x->a] != w->asleep[x->b])
			w->wakeIsl[w->isl[w->asleep[x->a] ? x->a : x->b]] = 1;
	}
	for (int i = 0; i < n; i++)
		if (w->asleep[i] && w->wakeIsl[w->isl[i]])
		{
			w->asleep[i] = 0;
			w->ready[i] = 0;
		}
	memset(w->wakeIsl, 0, (size_t)n);
	drop_sleeping(w);
	/* islands of the awake bodies over touching body pairs and joints,
	 * labelled by their smallest body index; sleeping bodies keep theirs */
	int* par = (int*)malloc(sizeof(int) * (size_t)(n ? n : 1));
	for (int i = 0; i < n; i++)
		par[i] = i;
	for (long long c = 0; c < w->ncon; c++)
		if (w->con[c].b >= 0)
			isl_union(par, w->con[c].a, w->con[c].b);
	for (int k = 0; k < w->nj; k++)
	{
		const pxo_joint* j = &w->joint[k];
		if (j->a >= 0 && j->b >= 0 && !w->asleep[j->a] && !w->asleep[j->b])
			isl_union(par, j->a, j->b);
	}
	unsigned char* notReady = (unsigned char*)calloc((size_t)(n ? n : 1), 1);
	for (int i = 0; i < n; i++)
		if (!w->asleep[i])
		{
			w->isl[i] = isl_find(par, i);
			if (!w->ready[i])
				notReady[w->isl[i]] = 1;
		}
	for (int i = 0; i < n; i++)
		if (!w->asleep[i] && !notReady[w->isl[i]])
		{
			w->asleep[i] = 1;
			w->lv[i] = V(0, 0, 0);
			w->av[i] = V(0, 0, 0);
		}
	free(notReady);
	free(par);
	drop_sleeping(w);
}

/* updateWakeCounter without stabilization and the readiness it reports
 * (lowleveldynamics/src/DySleep.cpp:173-235): motion velocities ml, ma
 * (world frame) of the step just taken, `counted` interactions */
static void sleep_check(pxo_world* w, int i, float dt, v3 ml, v3 ma, int counted)
{
	const float resetTime = 20.0f * 0.02f;
	float wc = w->wc[i];
	if (wc < resetTime * 0.5f || wc < dt)
	{
		const v3 t = w->invI[i];
		const v3 I = V(t.x > 0.0f ? 1.0f / t.x : 1.0f, t.y > 0.0f ? 1.0f / t.y : 1.0f, t.z > 0.0f ? 1.0f / t.z : 1.0f);
		w->accL[i] = add(w->accL[i], ml);
		w->accA[i] = add(w->accA[i], qrotinv(w->pose[i].q, ma));
		float invMass = w->invMass[i];
		if (invMass == 0.0f)
			invMass = 1.0f;
		const v3 a = w->accA[i], l = w->accL[i];
		const float angular = (a.x * a.x * I.x + a.y * a.y * I.y + a.z * a.z * I.z) * invMass;
		const float linear = l.x * l.x + l.y * l.y + l.z * l.z;
		const float energy = 0.5f * (angular + linear);
		const float cluster = (float)(1 + counted);
		const float threshold = cluster * w->d.sleep_threshold;
		if (energy >= threshold)
		{
			w->accL[i] = V(0, 0, 0);
			w->accA[i] = V(0, 0, 0);
			const float factor = threshold == 0.0f ? 2.0f : fminf(energy / threshold, 2.0f);
			const float old = wc;
			w->wc[i] = factor * 0.5f * resetTime + dt * (cluster - 1.0f);
			if (old == 0.0f)
				w->ready[i] = 0;   /* eACTIVATE_THIS_FRAME: not ready for sleep */
			return;
		}
	}
	wc = fmaxf(wc - dt, 0.0f);
	w->wc[i] = wc;
	if (wc == 0.0f)
	{
		/* eDEACTIVATE_THIS_FRAME: ready for sleep, filter reset */
		w->ready[i] = 1;
		w->accL[i] = V(0, 0, 0);
		w->accA[i] = V(0, 0, 0);
	}
}

int pxo_world_step(void* wp, float dt)
{
	pxo_world* w = (pxo_world*)wp;
	const pxo_desc* d = &w->d;
	const int n = w->n;
	const float co = d->contact_offset, cd = 2.0f * co, tol = d->tolerance_length;
	memset(w->stats, 0, sizeof(w->stats));
	w->ncon = 0;

	/* 1. bounds + broadphase */
	float* lo = (float*)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1));
	float* hi = (float*)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1));
	for (int i = 0; i < n; i++)
	{
		v3 mn, mx;
		shape_bounds(w->type[i], w->dims + 4 * i, w->pose[i], &mn, &mx);
		lo[3 * i] = mn.x - co; lo[3 * i + 1] = mn.y - co; lo[3 * i + 2] = mn.z - co;
		hi[3 * i] = mx.x + co; hi[3 * i + 1] = mx.y + co; hi[3 * i + 2] = mx.z + co;
	}
	pairvec pv = {0, 0, 0};
	bp_core(lo, hi, w->scene, n, &pv);
	free(lo);
	free(hi);
	if (w->agg)
	{
		long long k = 0;
		for (long long p = 0; p < pv.n; p++)
			if (!aggregate_filtered(w, (int)pv.p[2 * p], (int)pv.p[2 * p + 1]))
			{
				pv.p[2 * k] = pv.p[2 * p];
				pv.p[2 * k + 1] = pv.p[2 * p + 1];
				k++;
			}
		pv.n = k;
	}
	w->stats[0] = pv.n;

	/* interactions each body counts for its sleep threshold: broadphase
	 * pairs that are not filtered out (ShapeInteraction, ScShapeInteraction.cpp:84-93,
	 * ScPipeline.cpp:1222-1228), planes whose bounds it overlaps, and its
	 * joints (ScBodySim.h:159) */
	int* counted = NULL;
	if (d->enable_sleeping)
	{
		counted = (int*)calloc((size_t)(n ? n : 1), sizeof(int));
		for (long long p = 0; p < pv.n; p++)
		{
			const int a = (int)pv.p[2 * p], b = (int)pv.p[2 * p + 1];
			if (!jointed(w, a, b))
			{
				counted[a]++;
				counted[b]++;
			}
		}
		for (int i = 0; i < n; i++)
		{
			counted[i] += w->jcount[i];
			for (int q = 0; q < w->m; q++)
				if (w->pscene[q] == w->scene[i] && plane_counted(w, i, q, co))
					counted[i]++;
			/* a mesh whose broadphase bounds overlap the body's */
			for (int k = 0; k < w->ntm; k++)
			{
				if (w->tm[k].scene != w->scene[i])
					continue;
				v3 mn, mx;
				shape_bounds(w->type[i], w->dims + 4 * i, w->pose[i], &mn, &mx);
				const v3 a = V(mn.x - co, mn.y - co, mn.z - co), b = V(mx.x + co, mx.y + co, mx.z + co);
				const v3 lo = w->tm[k].blo, hi = w->tm[k].bhi;
				if (!(lo.x - co > b.x || a.x > hi.x + co || lo.y - co > b.y || a.y > hi.y + co || lo.z - co > b.z ||
				      a.z > hi.z + co))
					counted[i]++;
			}
		}
	}

	/* 2. narrowphase: body pairs in sorted pair order, then body-plane pairs
	 * (pairs of two sleeping bodies and the planes of sleeping bodies are
	 * dropped after the wake-up below; they never touch an awake body) */
	for (long long p = 0; p < pv.n; p++)
		add_pair(w, (int)pv.p[2 * p], (int)pv.p[2 * p + 1], cd, tol);
	free(pv.p);
	for (int i = 0; i < n; i++)
		for (int q = 0; q < w->m; q++)
		{
			if (w->pscene[q] != w->scene[i])
				continue;
			w->stats[1]++;
			cset c;
			c.count = 0;
			int k;
			if (w->type[i] == PXO_SPHERE)
				k = c_sphere_plane(w->dims[4 * i], w->pose[i].p, w->pn[q], w->pd[q], cd, &c);
			else if (w->type[i] == PXO_CAPSULE)
			{
				k = c_plane_capsule(w->pn[q], w->pd[q], w->dims[4 * i], w->dims[4 * i + 1], w->pose[i], cd, &c);
				c.n = neg(c.n); /* the capsule is body A: normal from the plane towards it */
			}
			else if (w->type[i] == PXO_CONVEX)
			{
				k = c_plane_convex(w->pn[q], w->pd[q], hull_of(w->dims[4 * i]), w->pose[i], cd, tol, &c);
				c.n = neg(c.n); /* the hull is body A: normal from the plane towards it */
			}
			else
			{
				k = c_plane_box(w->pn[q], w->pd[q], V(w->dims[4 * i], w->dims[4 * i + 1], w->dims[4 * i + 2]),
				                w->pose[i], cd, &c);
				c.n = neg(c.n); /* box is body A: normal from the plane towards the box */
			}
			if (k > 0)
				push_con(w, i, -1 - q, &c);
		}
	/* body-mesh contacts (one mesh per world) */
	for (int k = 0; k < w->ntm; k++)
		for (int i = 0; i < n; i++)
			if (w->scene[i] == w->tm[k].scene && tm_body(w, i, &w->tm[k], cd) < 0)
			{
				free(counted);
				return -6;   /* more than TM_MAX_TRIS triangles under one body */
			}
	if (d->enable_sleeping)
		sleep_islands(w);
	w->stats[2] = w->ncon;
	for (long long c = 0; c < w->ncon; c++)
		w->stats[3] += w->con[c].c.count;

	/* 3. colouring. A body-body constraint that existed in the previous step
	 * keeps its colour (below 64): those constraints were coloured together
	 * then, so they still never share a colour at a body. The others go
	 * through rounds of local minima of the ordering key at both bodies; each
	 * minimum takes the first colour free at both bodies (64-bit masks; past 63
	 * one colour per round, 64 + round). This replaces the reference's
	 * greedy partitioning (lowleveldynamics/src/DyConstraintPartition.cpp:858)
	 * with an order-free rule the GPU evaluates in parallel, kept from step to
	 * step like the reference GPU solver's incremental partitions
	 * (gpusolver/src/PxgConstraintPartition.cpp:1940). */
	uint64_t* best = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
	uint64_t* mask = (uint64_t*)calloc((size_t)(n ? n : 1), sizeof(uint64_t));
	long long left = 0;
	for (long long c = 0; c < w->ncon; c++)
	{
		pxo_con* k = &w->con[c];
		k->color = k->b >= 0 ? -1 : -2; /* -2: body-plane, placed after the rounds */
		if (k->b >= 0)
		{
			const fpatch* pp = find_patch(w, k->pkey);
			if (pp && pp->color >= 0 && pp->color < 64)
			{
				k->color = pp->color;
				mask[k->a] |= 1ull << pp->color;
				mask[k->b] |= 1ull << pp->color;
			}
			else
				left++;
		}
	}
	int round = 0, ncolors = 0;
	for (long long c = 0; c < w->ncon; c++)
		if (w->con[c].color + 1 > ncolors) ncolors = w->con[c].color + 1;
	while (left > 0)
	{
		for (int i = 0; i < n; i++)
			best[i] = ~0ull;
		for (long long c = 0; c < w->ncon; c++)
		{
			pxo_con* k = &w->con[c];
			if (k->color != -1) continue;
			if (k->key < best[k->a]) best[k->a] = k->key;
			if (k->b >= 0 && k->key < best[k->b]) best[k->b] = k->key;
		}
		for (long long c = 0; c < w->ncon; c++)
		{
			pxo_con* k = &w->con[c];
			if (k->color != -1) continue;
			if (best[k->a] == k->key && (k->b < 0 || best[k->b] == k->key))
			{
				const uint64_t used = mask[k->a] | (k->b >= 0 ? mask[k->b] : 0ull);
				int col = 64 + round;
				if (used != ~0ull)
					for (col = 0; (used >> col) & 1ull; col++)
						;
				k->color = col;
				if (col < 64)
				{
					mask[k->a] |= 1ull << col;
					if (k->b >= 0) mask[k->b] |= 1ull << col;
				}
				if (col + 1 > ncolors) ncolors = col + 1;
				left--;
			}
		}
		round++;
	}
	/* body-plane constraints go after every partition of their body's
	 * body-body constraints, one partition each, in constraint order
	 * (lowleveldynamics/src/DyConstraintPartition.cpp:168-179, 604-616) */
	int* next = (int*)calloc((size_t)(n ? n : 1), sizeof(int));
	for (long long c = 0; c < w->ncon; c++)
	{
		const pxo_con* k = &w->con[c];
		if (k->b < 0) continue;
		if (k->color + 1 > next[k->a]) next[k->a] = k->color + 1;
		if (k->color + 1 > next[k->b]) next[k->b] = k->color + 1;
	}
	for (long long c = 0; c < w->ncon; c++)
	{
		pxo_con* k = &w->con[c];
		if (k->b >= 0) continue;
		k->color = next[k->a]++;
		if (k->color + 1 > ncolors) ncolors = k->color + 1;
	}
	free(next);
	free(best);
	free(mask);
	w->stats[4] = ncolors;

	/* 4. pre-integration (DyTGSDynamics.cpp:992-1021, DyBodyCoreIntegrator.h:40-75) */
	sbody* sb = (sbody*)malloc(sizeof(sbody) * (size_t)(n ? n : 1));
	xf* frame = (xf*)malloc(sizeof(xf) * (size_t)(n ? n : 1));
	const v3 g = V(d->gravity[0], d->gravity[1], d->gravity[2]);
	for (int i = 0; i < n; i++)
	{
		v3 lv = w->lv[i], av = w->av[i];
		/* forces and torques change the velocities before gravity
		 * (BodySim::updateForces, ScBodySim.cpp:656-713) and last one step */
		const v3 el = w->extLin[i], ea = w->extAng[i];
		if (el.x != 0.0f || el.y != 0.0f || el.z != 0.0f) lv = add(lv, scl(el, dt));
		if (ea.x != 0.0f || ea.y != 0.0f || ea.z != 0.0f) av = add(av, scl(ea, dt));
		w->extLin[i] = V(0, 0, 0);
		w->extAng[i] = V(0, 0, 0);
		if (!w->asleep[i])   /* sleeping bodies are not in the solver: no gravity */
			lv = add(lv, scl(scl(g, dt), 1.0f));
		const float ld = 1.0f - d->linear_damping * dt, ad = 1.0f - d->angular_damping * dt;
		lv = scl(lv, ld >= 0.0f ? ld : 0.0f);
		av = scl(av, ad >= 0.0f ? ad : 0.0f);
		const float l2 = dot(lv, lv);
		if (l2 > d->max_lin_vel_sq) lv = scl(lv, sqrtf(d->max_lin_vel_sq / l2));
		const float a2 = dot(av, av);
		if (a2 > d->max_ang_vel_sq) av = scl(av, sqrtf(d->max_ang_vel_sq / a2));
		const m33 R = qmat(w->pose[i].q);
		const v3 si = V(sqrtf(w->invI[i].x), sqrtf(w->invI[i].y), sqrtf(w->invI[i].z));
		const v3 sI = V(si.x == 0.0f ? 0.0f : 1.0f / si.x, si.y == 0.0f ? 0.0f : 1.0f / si.y, si.z == 0.0f ? 0.0f : 1.0f / si.z);
		/* R diag(s) R^T (Cm::transformInertiaTensor) */
		m33 A, B;
		A.c0 = scl(R.c0, si.x); A.c1 = scl(R.c1, si.y); A.c2 = scl(R.c2, si.z);
		B.c0 = scl(R.c0, sI.x); B.c1 = scl(R.c1, sI.y); B.c2 = scl(R.c2, sI.z);
		const m33 Rt = mtrans(R);
		sbody* b = &sb[i];
		b->sqrtInvI = mmm(A, Rt);
		const m33 sqrtI = mmm(B, Rt);
		b->lin = lv;
		b->ang = mmul(sqrtI, av);
		b->dlin = V(0, 0, 0);
		b->dang = V(0, 0, 0);
		b->p = w->pose[i].p;
		b->dq.x = 0; b->dq.y = 0; b->dq.z = 0; b->dq.w = 1;
		b->origLin = lv;
		b->origAng = av;
		b->invMass = w->invMass[i];
		frame[i] = w->pose[i];
	}
	sbody worldBody;
	memset(&worldBody, 0, sizeof(worldBody));
	worldBody.dq.w = 1.0f;

	/* 5. contact prep (DyTGSDynamics.cpp:1898-1900 step dt and bias coefficient) */
	const int P = d->pos_iters > 0 ? d->pos_iters : 1, Vi = d->vel_iters;
	const float stepDt = dt / (float)P;
	const float invStepDt = 1.0f / stepDt;
	const float biasCoefficient = 2.0f * sqrtf(1.0f / (float)P);
	const float scale = fminf(0.8f, biasCoefficient);
	const int pgs = d->solver == 0;
	if (pgs && w->nj > 0)
	{
		free(sb);
		free(frame);
		free(counted);
		return -5;   /* joints: TGS only */
	}
	/* joint rows (DyTGSDynamics.cpp:1180-1256): poses at the start of the
	 * step, the world frame for a missing body; erp = half the bias
	 * coefficient (DyTGSContactPrep.cpp:2018) */
	{
		const xf id = {{0.0f, 0.0f, 0.0f, 1.0f}, {0.0f, 0.0f, 0.0f}};
		for (int i = 0; i < w->nj && !pgs; i++)
		{
			pxo_joint* j = &w->joint[i];
			if (joint_asleep(w, j))
				continue;
			prep_joint(j, j->a >= 0 ? w->pose[j->a] : id, j->b >= 0 ? w->pose[j->b] : id,
			           j->a >= 0 ? &sb[j->a] : &worldBody, j->b >= 0 ? &sb[j->b] : &worldBody, stepDt, invStepDt,
			           1.0f / dt,
			           0.5f * biasCoefficient, d->tolerance_length);
		}
	}
	if (pgs)
		for (int i = 0; i < n; i++)
		{
			sb[i].lin = V(0, 0, 0); /* PGS solves for velocity changes */
			sb[i].ang = V(0, 0, 0);
		}
	for (long long c = 0; c < w->ncon; c++)
	{
		if (pgs)
			prep_con_pgs(w, &w->con[c], sb, &worldBody, frame, 1.0f / dt, d->static_friction, d->dynamic_friction);
		else
			prep_con(w, &w->con[c], sb, &worldBody, frame, invStepDt, stepDt, dt, scale);
		w->stats[5] += w->con[c].nFric / 2;
	}

	/* 6. solve: position iterations by colour with substep integration, then
	 * velocity iterations (DyTGSDynamics.cpp:2515-2790) */
	long long* order = (long long*)malloc(sizeof(long long) * (size_t)(w->ncon ? w->ncon : 1));
	long long* start = (long long*)calloc((size_t)ncolors + 2, sizeof(long long));
	for (long long c = 0; c < w->ncon; c++) start[w->con[c].color + 1]++;
	for (int r = 0; r < ncolors; r++) start[r + 1] += start[r];
	{
		long long* fill = (long long*)malloc(sizeof(long long) * ((size_t)ncolors + 1));
		memcpy(fill, start, sizeof(long long) * ((size_t)ncolors + 1));
		for (long long c = 0; c < w->ncon; c++) order[fill[w->con[c].color]++] = c;
		free(fill);
	}
	v3* motionLin = NULL;
	v3* motionAng = NULL;
	float elapsed = 0.0f;
	if (pgs)
	{
		/* PGS (solveV_Blocks, lowleveldynamics/src/DySolverControl.cpp:163-330):
		 * position passes, friction only in the last three (no
		 * friction-every-iteration flag), the last one concluding; the
		 * velocity changes at that point move the bodies; then velocity
		 * passes */
		for (int it = P; it > 0; it--)
			for (int r = 0; r < ncolors; r++)
				for (long long o = start[r]; o < start[r + 1]; o++)
				{
					pxo_con* k = &w->con[order[o]];
					solve_con_pgs(k, &sb[k->a], k->b >= 0 ? &sb[k->b] : &worldBody, it <= 3, it == 1);
				}
		motionLin = (v3*)malloc(sizeof(v3) * (size_t)(n ? n : 1));
		motionAng = (v3*)malloc(sizeof(v3) * (size_t)(n ? n : 1));
		for (int i = 0; i < n; i++)
		{
			motionLin[i] = sb[i].lin;
			motionAng[i] = sb[i].ang;
		}
		/* at least one velocity pass (DySolverControl.cpp:343) */
		for (int it = 0; it < (Vi > 1 ? Vi : 1); it++)
			for (int r = 0; r < ncolors; r++)
				for (long long o = start[r]; o < start[r + 1]; o++)
				{
					pxo_con* k = &w->con[order[o]];
					solve_con_pgs(k, &sb[k->a], k->b >= 0 ? &sb[k->b] : &worldBody, 1, 0);
				}
	}
	for (int it = 0; it < P && !pgs; it++)
	{
		/* joints by colour, then contacts by colour; the last position
		 * iteration concludes the joint rows */
		for (int r = 0; r < w->njcolors; r++)
			for (int i = 0; i < w->nj; i++)
			{
				pxo_joint* j = &w->joint[i];
				if (j->color != r || joint_asleep(w, j)) continue;
				solve_joint(j, j->a >= 0 ? &sb[j->a] : &worldBody, j->b >= 0 ? &sb[j->b] : &worldBody, elapsed,
				            it == P - 1);
			}
		for (int r = 0; r < ncolors; r++)
			for (long long o = start[r]; o < start[r + 1]; o++)
			{
				pxo_con* k = &w->con[order[o]];
				solve_con(k, &sb[k->a], k->b >= 0 ? &sb[k->b] : &worldBody, -3.402823466e+38f, elapsed,
				          d->static_friction, d->dynamic_friction);
			}
		for (int i = 0; i < n; i++)
			integrate_step(&sb[i], stepDt);
		elapsed += stepDt;
	}
	for (int it = 0; it < Vi && !pgs; it++)
	{
		for (int r = 0; r < w->njcolors; r++)
			for (int i = 0; i < w->nj; i++)
			{
				pxo_joint* j = &w->joint[i];
				if (j->color != r || joint_asleep(w, j)) continue;
				solve_joint(j, j->a >= 0 ? &sb[j->a] : &worldBody, j->b >= 0 ? &sb[j->b] : &worldBody, elapsed, 0);
			}
		for (int r = 0; r < ncolors; r++)
			for (long long o = start[r]; o < start[r + 1]; o++)
			{
				pxo_con* k = &w->con[order[o]];
				solve_con(k, &sb[k->a], k->b >= 0 ? &sb[k->b] : &worldBody, 0.0f, elapsed, d->static_friction,
				          d->dynamic_friction);
			}
	}
	free(order);
	free(start);

	/* friction patches for the next step */
	w->prev = (fpatch*)realloc(w->prev, sizeof(fpatch) * (size_t)(w->ncon ? w->ncon : 1));
	for (long long c = 0; c < w->ncon; c++)
	{
		w->prev[c] = w->con[c].fp;
		w->prev[c].color = w->con[c].color;
	}
	w->nprev = w->ncon;
	qsort(w->prev, (size_t)w->nprev, sizeof(fpatch), cmp_patch);

	/* 7. PGS: integrate with the motion velocities, keep the final velocities
	 * (integrateCore, DyBodyCoreIntegrator.h:83-175) */
	if (pgs)
	{
		for (int i = 0; i < n; i++)
		{
			if (w->asleep[i])
				continue;   /* sleeping: pose and zero velocities kept */
			const sbody* b = &sb[i];
			const v3 lm = add(b->origLin, motionLin[i]);
			w->pose[i].p = add(w->pose[i].p, scl(lm, dt));
			v3 am = add(b->origAng, mmul(b->sqrtInvI, motionAng[i]));
			float ww = dot(am, am);
			if (ww != 0.0f)
			{
				ww = sqrtf(ww);
				if (ww > 1e7f)
				{
					am = scl(normalize(am), 1e7f);
					ww = 1e7f;
				}
				const float v = dt * ww * 0.5f;
				float sn = sinf(v);
				const float cs = cosf(v);
				sn /= ww;
				const v3 pqr = scl(am, sn);
				const qt qv = {pqr.x, pqr.y, pqr.z, 0.0f};
				const qt q0 = w->pose[i].q;
				qt res = qmul(qv, q0);
				res.x += q0.x * cs; res.y += q0.y * cs; res.z += q0.z * cs; res.w += q0.w * cs;
				w->pose[i].q = qnormalize(res);
			}
			w->lv[i] = add(b->origLin, b->lin);
			w->av[i] = add(b->origAng, mmul(b->sqrtInvI, b->ang));
			/* sleep check on the motion velocities (DyDynamics.cpp:1403-1420) */
			if (counted)
				sleep_check(w, i, dt, lm, am, counted[i]);
		}
		free(motionLin);
		free(motionAng);
		free(sb);
		free(frame);
		free(counted);
		return 0;
	}

	/* 7. copy back (DyTGSDynamics.cpp:1549-1580) */
	const float invDt = 1.0f / dt;
	for (int i = 0; i < n; i++)
	{
		if (w->asleep[i])
			continue;   /* sleeping: pose and zero velocities kept */
		const sbody* b = &sb[i];
		w->pose[i].q = qnormalize(qmul(b->dq, w->pose[i].q));
		w->pose[i].p = b->p;
		w->lv[i] = b->lin;
		w->av[i] = mmul(b->sqrtInvI, b->ang);
		/* sleep check on the step's mean velocities (DyTGSDynamics.cpp:1561, 1577) */
		if (counted)
			sleep_check(w, i, dt, scl(b->dlin, invDt), mmul(b->sqrtInvI, scl(b->dang, invDt)), counted[i]);
	}
	free(sb);
	free(frame);
	free(counted);
	return 0;
}
