// Joints on the device: the 1D constraint rows of fixed, spherical, revolute
// and prismatic joints, with the hard limits and the velocity drive of
// revolute and prismatic joints, for the TGS solver. One lane per joint: the
// joint's solverPrep with ConstraintHelper::prepareLockedAxes
// (physxextensions/src/ExtConstraintHelper.h:150-208, 243-297), the drive
// row and limit pair (RevoluteJointSolverPrep, ExtRevoluteJoint.cpp:195-258;
// PrismaticJointSolverPrep, ExtPrismaticJoint.cpp:119-153; addLimit,
// ExtConstraintHelper.h:316-393), then
// setupSolverConstraintStep with preprocessRows / orthogonalize
// (lowleveldynamics/src/DyTGSContactPrep.cpp:1939-2164,
// DyConstraintSetup.cpp:313-597); each iteration solve1DStep
// (DyTGSContactPrep.cpp:2418-2747) and, after the last position iteration,
// conclude1DStep (:2750-2780: springs give no impulse in the velocity
// iterations). The GPU reference's counterparts are
// constraintBlockPrePrep.cu:1168,1219 and solve1DBlockTGS
// (gpusolver/src/CUDA/solverBlockTGS.cuh:744).
//
// Layout per joint j (in colour order):
//   info[j]        int4  kind, body a, body b (-1: the world), colour
//   frame[4j..+3]  float4 frame0 p, frame0 q, frame1 p, frame1 q
//   param[3j..+2]  float4 (flags JP_*, lower, upper, restitution),
//                         (bounce threshold, drive velocity, drive force
//                         limit, drive gear ratio), (limit stiffness, limit
//                         damping, 0, 0)
//   head[9j..+8]   float4 (rA, invMass0), (rB, invMass1),
//                         (angOrthoAxis0[k], recipResponse[k]) k = 0..2,
//                         (angOrthoAxis1[k], error[k]) k = 0..2, (row count)
//   rows[48j + 6k ..+5] float4 per row k: (lin0, error), (lin1, biasScale),
//                         (ang0, velMultiplier), (ang1, velTarget),
//                         (minImpulse, maxImpulse, appliedForce, maxBias),
//                         (flags as bits, 0, 0, 0)
#pragma once
#include "pxh_math.h"

namespace pxh {

enum { JOINT_FIXED = 0, JOINT_SPHERICAL = 1, JOINT_REVOLUTE = 2, JOINT_PRISMATIC = 3, JOINT_DISTANCE = 4 };
enum { JF_ROT_EQ = 1, JF_ORTHO_TARGET = 2, JF_ANGULAR_ERROR = 4, JF_INEQUALITY = 8, JF_KEEP_BIAS = 16, JF_SPRING = 32 };
// PxConstraintSolveHint (physx/include/PxConstraintDesc.h:79-110)
enum { JHINT_NONE = 0, JHINT_ROT_EQ = 1024, JHINT_ROT_INEQ = 1025, JHINT_EQ = 2048, JHINT_INEQ = 2049 };
// Px1DConstraintFlag eKEEPBIAS, eRESTITUTION, eSPRING
enum { RF_KEEP_BIAS = 1, RF_RESTITUTION = 2, RF_SPRING = 4, RF_DRIVE_LIMIT = 8 };
// joint parameter flags: limit enabled, drive enabled, drive freespin
enum { JP_LIMIT = 1, JP_DRIVE = 2, JP_FREESPIN = 4, JP_COLLIDE = 8, JP_FORCES = 16 };
constexpr int JOINT_MAX_ROWS = 8;

// locked linear / angular axes (bit masks x y z) of each kind
PXH_D void jointAxes(int kind, int& lin, int& ang)
{
	lin = kind == JOINT_DISTANCE ? 0 : (kind == JOINT_PRISMATIC ? 6 : 7);
	ang = kind == JOINT_SPHERICAL || kind == JOINT_DISTANCE ? 0 : (kind == JOINT_REVOLUTE ? 6 : 7);
}

struct Row1D
{
	f3 lin0, ang0, lin1, ang1;
	float geomErr, velTarget, minImp, maxImp, rest, bthr, stiff, damp;
	int hint, angular, flags;
};

// a limit row (addLimit): soft (stiffness or damping positive), a spring row
// pushing back (no solve hint); hard, an inequality keeping its bias while
// the limit is not reached, bouncing when restitution is set; impulses at
// least zero. p2: (stiffness, damping, -, -).
PXH_D void limitRow(Row1D& c, float rest, float bthr, float4 p2)
{
	if (p2.x > 0.0f || p2.y > 0.0f)
	{
		c.flags |= RF_SPRING;
		c.stiff = p2.x;
		c.damp = p2.y;
		c.minImp = 0.0f;
		return;
	}
	c.hint = c.angular ? JHINT_ROT_INEQ : JHINT_INEQ;
	c.rest = rest;
	c.bthr = bthr;
	if (c.geomErr > 0.0f)
		c.flags |= RF_KEEP_BIAS;
	if (rest > 0.0f)
		c.flags |= RF_RESTITUTION;
	c.minImp = 0.0f;
}

// the twist angle of frame B about frame A's x axis (computePhi,
// ExtRevoluteJoint.cpp:154-179)
PXH_D float twistAngle(qt qa, qt qb)
{
	qt ca; ca.x = -qa.x; ca.y = -qa.y; ca.z = -qa.z; ca.w = qa.w;
	const qt r = qmul(ca, qb);
	float x = r.x, ww = r.w;
	const float y = 0.0f, z = 0.0f;
	const float mag = sqrtf(x * x + y * y + z * z + ww * ww);
	if (mag != 0.0f)
	{
		const float im = 1.0f / mag;
		x *= im;
		ww *= im;
	}
	float angle = acosf(fminf(fmaxf(ww, -1.0f), 1.0f)) * 2.0f;
	if (x < 0.0f)
		angle = -angle;
	return angle;
}

// computeJacobianAxes (ExtConstraintHelper.h:77-102)
PXH_D void jacobianAxes(f3 row[3], qt qa, qt qb)
{
	const float wa = qa.w, wb = qb.w;
	const f3 va = mk(qa.x, qa.y, qa.z), vb = mk(qb.x, qb.y, qb.z);
	const f3 c = vb * wa + va * wb;
	const float d0 = wa * wb;
	const float d1 = dot(va, vb);
	const float d = d0 - d1;
	row[0] = (va * vb.x + vb * va.x + mk(d, c.z, -c.y)) * 0.5f;
	row[1] = (va * vb.y + vb * va.y + mk(-c.z, d, c.x)) * 0.5f;
	row[2] = (va * vb.z + vb * va.z + mk(c.y, -c.x, d)) * 0.5f;
	if ((d0 + d1) == 0.0f)
	{
		row[0].x += 1.192092896e-07f;
		row[1].y += 1.192092896e-07f;
		row[2].z += 1.192092896e-07f;
	}
}

// the closest point of the ellipse with radii (ry, rz) in the y-z plane to
// point (PxEllipseClamp, physx/include/foundation/PxMathUtils.h:204-262)
PXH_D f3 ellipseClamp(f3 point, float ry, float rz)
{
	const float qy = fabsf(point.y), qz = fabsf(point.z);
	const float tinyEps = 1e-6f;
	if (ry >= rz)
	{
		if (qz < tinyEps)
			return mk(0.0f, point.y > 0.0f ? ry : -ry, 0.0f);
	}
	else if (qy < tinyEps)
		return mk(0.0f, 0.0f, point.z > 0.0f ? rz : -rz);
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
			return mk(0.0f, e2y * point.y * dy, e2z * point.z * dz);
		const float df = (0.0f * 0.0f + fy * dy + fz * dz) * -2.0f;
		t = t - f / df;
	}
	const float ry_ = e2y * point.y * dy, rz_ = e2z * point.z * dz;
	const float a = ry_ / ry, b = rz_ / rz;
	const float sc = 1.0f / sqrtf(a * a + b * b);
	return mk(0.0f * sc, ry_ * sc, rz_ * sc);
}

// the cone limit's axis and error (ConeLimitHelperTanLess::getLimit with
// computeAxisAndError, common/src/CmConeLimitHelper.h:50-69, 163-178)
PXH_D float coneLimit(qt swing, float ymax, float zmax, f3& axis)
{
	const float x2 = swing.x * 2.0f, w2 = swing.w * 2.0f;
	const f3 twistAxis = mk((swing.w * w2) - 1.0f + swing.x * x2, (swing.z * w2) + swing.y * x2,
	                        (-swing.y * w2) + swing.z * x2);
	const f3 swingAngle = mk(0.0f, 4.0f * atan2f(swing.y, 1.0f + swing.w), 4.0f * atan2f(swing.z, 1.0f + swing.w));
	const f3 p = ellipseClamp(swingAngle, ymax, zmax);
	const f3 normal = mk(0.0f, p.y / (ymax * ymax), p.z / (zmax * zmax));
	const f3 r = mk(0.0f, tanf(p.y / 4.0f), tanf(p.z / 4.0f)), d = mk(0.0f, normal.y, normal.z);
	const f3 px = mk(1.0f, 0.0f, 0.0f);
	const float r2 = dot(r, r), a = 1.0f - r2, b = 1.0f / (1.0f + r2), b2 = b * b;
	const float v1 = 2.0f * a * b2;
	const f3 v2 = mk(a, 2.0f * r.z, -2.0f * r.y);
	const f3 coneLine = v2 * v1 - px;
	const float rd = dot(r, d);
	const float dv1 = -4.0f * rd * (3.0f - r2) * b2 * b;
	const f3 dv2 = mk(-2.0f * rd, 2.0f * d.z, -2.0f * d.y);
	const f3 coneNormal = dv2 * v1 + v2 * dv1;
	const float inv = 1.0f / sqrtf(dot(coneNormal, coneNormal));
	axis = cross(coneLine, coneNormal) * inv;
	return dot(cross(coneLine, axis), twistAxis);
}

// the joint's rows (solverPrep); returns the count, ra / rb the anchor
// offsets. p0 / p1: the joint's parameters (flags, lower, upper,
// restitution), (bounce threshold, drive velocity, drive force limit, gear),
// p2 (limit stiffness, limit damping, -, -).
PXH_D int jointRows(int kind, float4 p0, float4 p1, float4 p2, const xf& bA, const xf& bB, const xf& f0,
                    const xf& f1, Row1D* rows, f3& raOut, f3& rbOut)
{
	xf cA, cB;
	cA.q = qmul(bA.q, f0.q);
	cA.p = qrot(bA.q, f0.p) + bA.p;
	cB.q = qmul(bB.q, f1.q);
	cB.p = qrot(bB.q, f1.p) + bB.p;
	if (cA.q.x * cB.q.x + cA.q.y * cB.q.y + cA.q.z * cB.q.z + cA.q.w * cB.q.w < 0.0f)
	{
		cB.q.x = -cB.q.x; cB.q.y = -cB.q.y; cB.q.z = -cB.q.z; cB.q.w = -cB.q.w;
	}
	const f3 ra0 = cB.p - bA.p;
	f3 ra = ra0;
	const f3 rb = cB.p - bB.p;
	int lin, ang;
	jointAxes(kind, lin, ang);
	const int pflags = int(p0.x);
	const float lo = p0.y, hi = p0.z;
	const bool limitOn = (kind == JOINT_REVOLUTE || kind == JOINT_PRISMATIC) && (pflags & JP_LIMIT);
	const bool limitLocked = limitOn && lo >= hi;
	if (limitLocked)
	{
		if (kind == JOINT_REVOLUTE)
			ang = 7;
		else
			lin = 7;
	}
	const f3 cB2cAp = qrotinv(cA.q, cB.p - cA.p);
	for (int i = 0; i < JOINT_MAX_ROWS; i++)
	{
		rows[i].lin0 = rows[i].ang0 = rows[i].lin1 = rows[i].ang1 = mk(0.0f, 0.0f, 0.0f);
		rows[i].geomErr = rows[i].velTarget = rows[i].rest = rows[i].bthr = rows[i].stiff = rows[i].damp = 0.0f;
		rows[i].minImp = -3.402823466e+38f;
		rows[i].maxImp = 3.402823466e+38f;
		rows[i].hint = 0;
		rows[i].angular = 0;
		rows[i].flags = 0;
	}
	int k = 0;
	if (kind == JOINT_DISTANCE)
	{
		// DistanceJointSolverPrep (ExtDistanceJoint.cpp:179-271): rows along
		// the anchors' direction through frame B's anchor; flags 1 max
		// distance, 2 min distance, 4 spring; lower / upper the min / max
		// distance, the restitution slot the tolerance
		const bool enforceMax = (pflags & 1) && hi >= 0.0f;
		const bool enforceMin = (pflags & 2) && lo >= 0.0f;
		if (enforceMax || enforceMin)
		{
			f3 dir = cA.p - cB.p;
			const float dist = sqrtf(dot(dir, dir));
			if (dist > 0.0f)
				dir = dir * (1.0f / dist);
			if (dist < 1.192092896e-07f)
				dir = mk(1.0f, 0.0f, 0.0f);
			const float tol = p0.w;
			const bool both = enforceMin && enforceMax;
			const int nrows = both && lo != hi ? 2 : 1;
			for (int r = 0; r < nrows; r++)
			{
				Row1D& e = rows[k + r];
				e.lin0 = dir; e.ang0 = cross(ra0, dir);
				e.lin1 = dir; e.ang1 = cross(rb, dir);
				if (pflags & 4)
				{
					e.flags |= RF_SPRING;
					e.stiff = p2.x;
					e.damp = p2.y;
				}
			}
			if (both && lo == hi)
			{
				const float error = dist - hi;
				rows[k].geomErr = error > tol ? error - tol : (error < -tol ? error + tol : 0.0f);
			}
			else
			{
				int c = k;
				if (enforceMin)
				{
					rows[c].geomErr = dist - lo + tol;
					rows[c].minImp = 0.0f;
					if (dist >= lo)
						rows[c].flags |= RF_KEEP_BIAS;
					c++;
				}
				if (enforceMax)
				{
					rows[c].geomErr = dist - hi - tol;
					rows[c].maxImp = 0.0f;
					if (dist <= hi)
						rows[c].flags |= RF_KEEP_BIAS;
				}
			}
			k += nrows;
		}
	}
	if (kind == JOINT_SPHERICAL && (pflags & JP_LIMIT))
	{
		// the cone limit row comes first (SphericalJointSolverPrep,
		// ExtSphericalJoint.cpp:157-190): the swing of B's frame in A's
		// (PxSeparateSwingTwist, PxMathUtils.h:272-276), its axis in the world
		qt ca; ca.x = -cA.q.x; ca.y = -cA.q.y; ca.z = -cA.q.z; ca.w = cA.q.w;
		const qt q = qmul(ca, cB.q);
		qt tw; tw.x = 0.0f; tw.y = 0.0f; tw.z = 0.0f; tw.w = 1.0f;
		if (q.x != 0.0f)
		{
			const float sc = 1.0f / sqrtf(q.x * q.x + 0.0f * 0.0f + 0.0f * 0.0f + q.w * q.w);
			tw.x = q.x * sc; tw.y = 0.0f * sc; tw.z = 0.0f * sc; tw.w = q.w * sc;
		}
		qt tc; tc.x = -tw.x; tc.y = -tw.y; tc.z = -tw.z; tc.w = tw.w;
		const qt swing = qmul(q, tc);
		f3 ax;
		const float err = coneLimit(swing, lo, hi, ax);
		Row1D& c = rows[k++];
		c.ang0 = qrot(cA.q, ax);
		c.ang1 = c.ang0;
		c.angular = 1;
		c.geomErr = err;
		limitRow(c, p0.w, p1.x, p2);
	}
	const m33 axes = qmat(cA.q);
	if (lin)
	{
		f3 ev = mk(0.0f, 0.0f, 0.0f);
		if (lin & 1) ev = ev - axes.c0 * cB2cAp.x;
		if (lin & 2) ev = ev - axes.c1 * cB2cAp.y;
		if (lin & 4) ev = ev - axes.c2 * cB2cAp.z;
		ra = ra + ev;
		for (int i = 0; i < 3; i++)
			if (lin & (1 << i))
			{
				const f3 ax = i == 0 ? axes.c0 : (i == 1 ? axes.c1 : axes.c2);
				Row1D& c = rows[k++];
				c.hint = JHINT_EQ;
				c.lin0 = ax;
				c.ang0 = cross(ra, ax);
				c.lin1 = ax;
				c.ang1 = cross(rb, ax);
				c.geomErr = -comp(cB2cAp, i);
			}
	}
	if (ang)
	{
		qt ca; ca.x = -cA.q.x; ca.y = -cA.q.y; ca.z = -cA.q.z; ca.w = cA.q.w;
		const qt qB2qA = qmul(ca, cB.q);
		f3 row[3];
		jacobianAxes(row, cA.q, cB.q);
		for (int i = 0; i < 3; i++)
			if (ang & (1 << i))
			{
				Row1D& c = rows[k++];
				c.hint = JHINT_EQ;
				c.ang0 = row[i];
				c.ang1 = row[i];
				c.geomErr = -(i == 0 ? qB2qA.x : (i == 1 ? qB2qA.y : qB2qA.z));
				c.angular = 1;
			}
	}
	const f3 axis = axes.c0;
	const f3 naxis = -axis;
	if (kind == JOINT_REVOLUTE && !limitLocked)
	{
		if (pflags & JP_DRIVE)
		{
			Row1D& c = rows[k++];
			c.hint = JHINT_NONE;
			c.ang0 = naxis;
			c.ang1 = naxis * p1.w;
			c.velTarget = p1.y;
			c.minImp = -p1.z;
			c.maxImp = p1.z;
			c.angular = 1;
			c.flags |= RF_DRIVE_LIMIT;
			if (pflags & JP_FREESPIN)
			{
				if (p1.y > 0.0f) c.minImp = 0.0f;
				if (p1.y < 0.0f) c.maxImp = 0.0f;
			}
		}
		if (limitOn)
		{
			// anglePair: a soft limit only beyond its stop
			const float phi = twistAngle(cA.q, cB.q);
			const bool soft = p2.x > 0.0f || p2.y > 0.0f;
			if (!soft || phi < lo)
			{
				Row1D& c0 = rows[k++];
				c0.ang0 = naxis; c0.ang1 = naxis; c0.angular = 1;
				c0.geomErr = -(lo - phi);
				limitRow(c0, p0.w, p1.x, p2);
			}
			if (!soft || phi > hi)
			{
				Row1D& c1 = rows[k++];
				c1.ang0 = axis; c1.ang1 = axis; c1.angular = 1;
				c1.geomErr = hi - phi;
				limitRow(c1, p0.w, p1.x, p2);
			}
		}
	}
	if (kind == JOINT_PRISMATIC && limitOn && !limitLocked)
	{
		// linearLimit: a soft limit only beyond its stop
		const float ordinate = cB2cAp.x;
		const bool soft = p2.x > 0.0f || p2.y > 0.0f;
		if (!soft || ordinate > hi)
		{
			Row1D& c0 = rows[k++];
			c0.lin0 = axis; c0.ang0 = cross(ra0, axis); c0.lin1 = axis; c0.ang1 = cross(rb, axis);
			c0.geomErr = hi - ordinate;
			limitRow(c0, p0.w, p1.x, p2);
		}
		if (!soft || -ordinate > -lo)
		{
			Row1D& c1 = rows[k++];
			c1.lin0 = naxis; c1.ang0 = cross(ra0, naxis); c1.lin1 = naxis; c1.ang1 = cross(rb, naxis);
			c1.geomErr = -lo - -ordinate;
			limitRow(c1, p0.w, p1.x, p2);
		}
	}
	raOut = ra;
	rbOut = rb;
	return k;
}

// orthogonalize over one group of rows (Gram-Schmidt in the inverse-mass
// metric; angular inverse-mass scales are 1), each row against the group's
// first eq (equality) rows
PXH_D void orthogonalizeRows(Row1D* const* row, f3* as0, f3* as1, int count, int eq, float im0, float im1)
{
	f3 lin0m[JOINT_MAX_ROWS], ang0m[JOINT_MAX_ROWS], lin1m[JOINT_MAX_ROWS], ang1m[JOINT_MAX_ROWS];
	f3 lin0[JOINT_MAX_ROWS], ang0[JOINT_MAX_ROWS], lin1[JOINT_MAX_ROWS], ang1[JOINT_MAX_ROWS];
	float g[JOINT_MAX_ROWS], t[JOINT_MAX_ROWS];
	for (int i = 0; i < count; i++)
	{
		f3 l0 = row[i]->lin0, a0 = row[i]->ang0, l1 = row[i]->lin1, a1 = row[i]->ang1;
		float gi = row[i]->geomErr, ti = row[i]->velTarget;
		f3 s0v = as0[i], s1v = as1[i];
		const int elim = i < eq ? i : eq;
		for (int k = 0; k < elim; k++)
		{
			const f3 sa = mulv(l1, lin1m[k]) + mulv(l0, lin0m[k]);
			const f3 sb = mulv(s1v, ang1m[k]) + mulv(s0v, ang0m[k]);
			const f3 ss = sa + sb;
			const float tt = ss.x + ss.y + ss.z;
			l0 = l0 - lin0[k] * tt;
			gi = gi - g[k] * tt;
			a0 = a0 - ang0[k] * tt;
			ti = ti - t[k] * tt;
			l1 = l1 - lin1[k] * tt;
			a1 = a1 - ang1[k] * tt;
			s0v = s0v - as0[k] * tt;
			s1v = s1v - as1[k] * tt;
		}
		row[i]->lin0 = l0; row[i]->geomErr = gi;
		row[i]->ang0 = a0; row[i]->velTarget = ti;
		row[i]->lin1 = l1;
		row[i]->ang1 = a1;
		as0[i] = s0v;
		as1[i] = s1v;
		lin0[i] = l0; g[i] = gi; ang0[i] = a0; t[i] = ti; lin1[i] = l1; ang1[i] = a1;
		const f3 l0m = l0 * im0, l1m = l1 * im1;
		const f3 q0 = mulv(l0, l0m) + mulv(l1, l1m);
		const f3 q1 = mulv(s0v, s0v) + mulv(s1v, s1v);
		const f3 qq = q0 + q1;
		const float sum = qq.x + qq.y + qq.z;
		const float a = sum > 0.0f ? 1.0f / sum : 0.0f;
		lin0m[i] = l0m * a;
		ang0m[i] = s0v * a;
		lin1m[i] = l1m * a;
		ang1m[i] = s1v * a;
	}
}

// prepare joint j's header and rows for this step. sqrtInvI0 / sqrtInvI1:
// world-space square-root inverse inertia of the two bodies (zero for the
// world); invMass0 / invMass1 likewise; v0 / w0 and v1 / w1 their velocities
// at the start of the solve (zero for the world).
PXH_D void prepJoint(int kind, float4 p0, float4 p1, float4 p2, const xf& bA, const xf& bB, const xf& f0,
                     const xf& f1, const m33& sI0, const m33& sI1, float invMass0, float invMass1, f3 v0, f3 w0, f3 v1,
                     f3 w1, float stepDt, float invStepDt, float invDt, float erp, float lengthScale, float4* head,
                     float4* rowsOut)
{
	Row1D rows[JOINT_MAX_ROWS];
	f3 ra, rb;
	const int n = jointRows(kind, p0, p1, p2, bA, bB, f0, f1, rows, ra, rb);
	for (int i = 0; i < n; i++)
		if (rows[i].angular)
		{
			if (rows[i].hint == JHINT_EQ)
				rows[i].hint = JHINT_ROT_EQ;
			else if (rows[i].hint == JHINT_INEQ)
				rows[i].hint = JHINT_ROT_INEQ;
		}
	Row1D* sorted[JOINT_MAX_ROWS];
	for (int i = 0; i < n; i++)
	{
		int k = i;
		for (; k > 0 && rows[i].hint < sorted[k - 1]->hint; k--)
			sorted[k] = sorted[k - 1];
		sorted[k] = &rows[i];
	}
	f3 as0[JOINT_MAX_ROWS], as1[JOINT_MAX_ROWS];
	for (int i = 0; i < n; i++)
	{
		as0[i] = mmul(sI0, sorted[i]->ang0);
		as1[i] = mmul(sI1, sorted[i]->ang1);
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
			orthogonalizeRows(sorted + start, as0 + start, as1 + start, i - start, eq - start, invMass0, invMass1);
		}
	}
	f3 o0[3], o1[3];
	float orecip[3], oerr[3];
	for (int k = 0; k < 3; k++)
	{
		o0[k] = o1[k] = mk(0.0f, 0.0f, 0.0f);
		orecip[k] = oerr[k] = 0.0f;
	}
	int ortho = 0;
	for (int i = 0; i < n; i++)
	{
		const Row1D& c = *sorted[i];
		int flags = JF_ANGULAR_ERROR;
		const float linSum = dot(c.lin0, c.lin0) * invMass0 + dot(c.lin1, c.lin1) * invMass1;
		const float resp0 = dot(as0[i], as0[i]), resp1 = dot(as1[i], as1[i]);
		const float unit = resp0 + resp1 + linSum;
		// projectVelocity of both bodies on the row as prepared
		const float vel0 = dot(v0, c.lin0) + dot(w0, c.ang0);
		const float vel1 = dot(v1, c.lin1) + dot(w1, c.ang1);
		const float jointSpeed = vel0 - vel1;
		f3 ang0 = c.ang0, ang1 = c.ang1;
		if (!c.angular)
		{
			ang0 = ang1 = mk(0.0f, 0.0f, 0.0f);
			flags &= ~JF_ANGULAR_ERROR;
		}
		const float recip = unit <= 0.0f ? 0.0f : 1.0f / unit;
		float bounce = jointSpeed * -c.rest;
		if (!((c.flags & RF_RESTITUTION) && -jointSpeed > c.bthr && bounce * c.geomErr <= 0.0f))
			bounce = 0.0f;
		float velMul = -1.0f;
		float maxBias, biasScale, error, velTarget;
		if (c.flags & RF_SPRING)
		{
			// a force spring: implicit in the step
			const float a = stepDt * (stepDt * c.stiff + c.damp);
			const float b = stepDt * (c.damp * c.velTarget);
			const float x = 1.0f / (1.0f + a * unit);
			maxBias = 3.402823466e+38f;
			biasScale = -x * c.stiff * stepDt;
			error = biasScale * c.geomErr;
			velMul = -x * a;
			velTarget = x * b - velMul * 0.0f;
			flags |= JF_SPRING;
		}
		else if (bounce != 0.0f)
		{
			maxBias = 0.0f;
			biasScale = 0.0f;
			error = 0.0f;
			velTarget = bounce - velMul * 0.0f;
		}
		else
		{
			maxBias = c.angular ? invDt * 0.75f : invDt * 15.0f * lengthScale;
			biasScale = -invStepDt * erp;
			error = c.geomErr * biasScale;
			velTarget = c.velTarget - velMul * 0.0f;
		}
		if (c.flags & RF_KEEP_BIAS)
			flags |= JF_KEEP_BIAS;
		if (c.hint & 1)
			flags |= JF_INEQUALITY;
		if (c.hint == JHINT_ROT_EQ)
		{
			flags |= JF_ROT_EQ;
			o0[ortho] = as0[i];
			orecip[ortho] = recip;
			o1[ortho] = as1[i];
			oerr[ortho] = c.geomErr;
			ortho++;
		}
		else if (c.hint & JHINT_EQ)
			flags |= JF_ORTHO_TARGET;
		float4* r = rowsOut + 6 * i;
		r[0] = f4(c.lin0, error);
		r[1] = f4(c.lin1, biasScale);
		r[2] = f4(ang0, velMul);
		r[3] = f4(ang1, velTarget);
		// a drive's limits are forces with eDRIVE_LIMITS_ARE_FORCES (flag 16)
		const float driveScale = ((c.flags & RF_DRIVE_LIMIT) && (int(p0.x) & JP_FORCES)) ? 1.0f / invDt : 1.0f;
		r[4] = make_float4(c.minImp * driveScale, c.maxImp * driveScale, 0.0f, maxBias);
		r[5] = make_float4(__int_as_float(flags), 0.0f, 0.0f, 0.0f);
	}
	head[0] = f4(ra, invMass0);
	head[1] = f4(rb, invMass1);
	for (int k = 0; k < 3; k++)
	{
		head[2 + k] = f4(o0[k], orecip[k]);
		head[5 + k] = f4(o1[k], oerr[k]);
	}
	head[8] = make_float4(__int_as_float(n), 0.0f, 0.0f, 0.0f);
}

// One pass over joint rows. v0 / v1: the bodies' velocity records (linear
// velocity, angular velocity in sqrt-inertia space, accumulated linear and
// angular motion); dq0 / dq1 their rotation since the step began.
PXH_D void solveJoint(int n, const float4* head, float4* rows, f3& l0, f3& a0, f3 dlin0, f3 dang0, qt dq0,
                      const m33& sI0, f3& l1, f3& a1, f3 dlin1, f3 dang1, qt dq1, const m33& sI1, float elapsed,
                      bool conclude)
{
	const f3 rA = mk(head[0]), rB = mk(head[1]);
	const float im0 = head[0].w, im1 = head[1].w;
	const f3 ra = qrot(dq0, rA);
	const f3 rb = qrot(dq1, rB);
	const f3 raMotion = (ra + dlin0) - rA;
	const f3 rbMotion = (rb + dlin1) - rB;
	f3 o0[3], o1[3];
	float orecip[3], oerr[3];
	for (int k = 0; k < 3; k++)
	{
		const float4 h0 = head[2 + k], h1 = head[5 + k];
		o0[k] = mk(h0);
		orecip[k] = h0.w;
		o1[k] = mk(h1);
		oerr[k] = h1.w + (dot(o0[k], dang0) - dot(o1[k], dang1));
	}
	for (int i = 0; i < n; i++)
	{
		float4* r = rows + 6 * i;
		const float4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3], r4 = r[4];
		const int flags = __float_as_int(r[5].x);
		const f3 lin0 = mk(r0), lin1 = mk(r1);
		const float biasScale = r1.w, velMul = r2.w, velTarget = r3.w;
		const float minImp = r4.x, maxImp = r4.y, applied = r4.z, maxBias = r4.w;
		const f3 cang0 = mk(r2) + cross(ra, lin0);
		const f3 cang1 = mk(r3) + cross(rb, lin1);
		float error = r0.w;
		const float minBias = (flags & JF_INEQUALITY) ? -3.402823466e+38f : -maxBias;
		f3 raXnI = mmul(sI0, cang0);
		f3 rbXnI = mmul(sI1, cang1);
		if (flags & JF_ORTHO_TARGET)
		{
			float proj[3];
			for (int k = 0; k < 3; k++)
			{
				const f3 pp = mulv(raXnI, o0[k]) + mulv(rbXnI, o1[k]);
				proj[k] = (pp.x + pp.y + pp.z) * orecip[k];
			}
			const f3 d0 = o0[0] * proj[0] + (o0[1] * proj[1] + o0[2] * proj[2]);
			const f3 d1 = o1[0] * proj[0] + (o1[1] * proj[1] + o1[2] * proj[2]);
			raXnI = raXnI - d0;
			rbXnI = rbXnI - d1;
			const float ob = biasScale * (oerr[0] * proj[0] + (oerr[1] * proj[1] + oerr[2] * proj[2]));
			error = error - ob;
		}
		const float angScale = (flags & JF_ANGULAR_ERROR) ? 1.0f : 0.0f;
		const float dAng = angScale * (dot(raXnI, dang0) - dot(rbXnI, dang1));
		const float dLin = dot(lin0, raMotion) - dot(lin1, rbMotion);
		const bool spring = (flags & JF_SPRING) != 0;
		const float motion = dLin + dAng;
		const float errorChange = spring ? motion : motion - velTarget * elapsed;
		const f3 q0 = mulv(raXnI, raXnI), q1 = mulv(rbXnI, rbXnI);
		const float resp0 = im0 * dot(lin0, lin0) + (q0.x + q0.y + q0.z);
		const float resp1 = im1 * dot(lin1, lin1) - -(q1.x + q1.y + q1.z);
		const float response = resp0 + resp1;
		const float recip = response > 0.0f ? 1.0f / response : 0.0f;
		const float vMul = spring ? velMul : recip * velMul;
		const float ub = errorChange * biasScale + error;
		const float bias = fminf(fmaxf(ub, minBias), maxBias);
		const float constant = spring ? bias + velTarget : recip * (bias + velTarget);
		const f3 v0 = mulv(l0, lin0) + mulv(a0, raXnI);
		const f3 v1 = mulv(l1, lin1) + mulv(a1, rbXnI);
		const f3 dv = v0 - v1;
		const float normalVel = dv.x + dv.y + dv.z;
		const float unclamped = applied + (vMul * normalVel + constant);
		const float clamped = fminf(fmaxf(unclamped, minImp), maxImp);
		const float df = clamped - applied;
		r[4].z = clamped;
		l0 = lin0 * (df * im0) + l0;
		l1 = l1 - lin1 * (df * im1);
		a0 = raXnI * df + a0;
		a1 = rbXnI * (df * -1.0f) + a1;
		if (conclude && !(flags & JF_KEEP_BIAS))
		{
			r[0].w = 0.0f;
			r[1].w = 0.0f;
		}
		if (conclude && spring)
		{
			// no spring impulse in the velocity iterations
			r[0].w = 0.0f;
			r[1].w = 0.0f;
			r[2].w = 0.0f;
			r[3].w = 0.0f;
		}
	}
}

}  // namespace pxh
