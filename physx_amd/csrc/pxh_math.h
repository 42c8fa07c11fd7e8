// Device math for the rigid-body step: float3/quaternion/3x3 helpers with the
// arithmetic order of the reference's PxVec3/PxQuat/PxMat33
// (physx/include/foundation/PxQuat.h, PxMat33.h, PxTransform.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define PXH_D __device__ __forceinline__

namespace pxh {

struct f3 { float x, y, z; };
struct qt { float x, y, z, w; };
struct m33 { f3 c0, c1, c2; };     // columns
struct xf { qt q; f3 p; };         // PxTransform

PXH_D f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
PXH_D f3 mk(float4 v) { return mk(v.x, v.y, v.z); }
PXH_D float4 f4(f3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
PXH_D f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
PXH_D f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
PXH_D f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
PXH_D f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
PXH_D f3 mulv(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
PXH_D float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PXH_D f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
PXH_D f3 vabs(f3 a) { return mk(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
PXH_D f3 normalize(f3 a) { return a * (1.0f / sqrtf(dot(a, a))); }
PXH_D float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

PXH_D qt mkq(float4 v) { qt q; q.x = v.x; q.y = v.y; q.z = v.z; q.w = v.w; return q; }
PXH_D float4 f4(qt q) { return make_float4(q.x, q.y, q.z, q.w); }

PXH_D f3 qrot(qt q, f3 v)
{
	const float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
	const float w2 = q.w * q.w - 0.5f;
	const float dot2 = q.x * vx + q.y * vy + q.z * vz;
	return mk(vx * w2 + (q.y * vz - q.z * vy) * q.w + q.x * dot2, vy * w2 + (q.z * vx - q.x * vz) * q.w + q.y * dot2,
	          vz * w2 + (q.x * vy - q.y * vx) * q.w + q.z * dot2);
}
PXH_D f3 qrotinv(qt q, f3 v)
{
	const float vx = 2.0f * v.x, vy = 2.0f * v.y, vz = 2.0f * v.z;
	const float w2 = q.w * q.w - 0.5f;
	const float dot2 = q.x * vx + q.y * vy + q.z * vz;
	return mk(vx * w2 - (q.y * vz - q.z * vy) * q.w + q.x * dot2, vy * w2 - (q.z * vx - q.x * vz) * q.w + q.y * dot2,
	          vz * w2 - (q.x * vy - q.y * vx) * q.w + q.z * dot2);
}
PXH_D qt qmul(qt a, qt b)
{
	qt r;
	r.x = a.w * b.x + b.w * a.x + a.y * b.z - b.y * a.z;
	r.y = a.w * b.y + b.w * a.y + a.z * b.x - b.z * a.x;
	r.z = a.w * b.z + b.w * a.z + a.x * b.y - b.x * a.y;
	r.w = a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z;
	return r;
}
PXH_D qt qnormalize(qt q)
{
	const float m = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
	const float s = m > 0.0f ? 1.0f / m : 0.0f;
	qt r; r.x = q.x * s; r.y = q.y * s; r.z = q.z * s; r.w = q.w * s;
	return r;
}
PXH_D m33 qmat(qt q)
{
	const float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
	const float xx = x2 * q.x, yy = y2 * q.y, zz = z2 * q.z;
	const float xy = x2 * q.y, xz = x2 * q.z, xw = x2 * q.w;
	const float yz = y2 * q.z, yw = y2 * q.w, zw = z2 * q.w;
	m33 m;
	m.c0 = mk(1.0f - yy - zz, xy + zw, xz - yw);
	m.c1 = mk(xy - zw, 1.0f - xx - zz, yz + xw);
	m.c2 = mk(xz + yw, yz - xw, 1.0f - xx - yy);
	return m;
}
PXH_D f3 mmul(const m33& m, f3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
PXH_D f3 mtmul(const m33& m, f3 v) { return mk(dot(m.c0, v), dot(m.c1, v), dot(m.c2, v)); }
PXH_D m33 mtrans(const m33& m)
{
	m33 t;
	t.c0 = mk(m.c0.x, m.c1.x, m.c2.x);
	t.c1 = mk(m.c0.y, m.c1.y, m.c2.y);
	t.c2 = mk(m.c0.z, m.c1.z, m.c2.z);
	return t;
}
PXH_D m33 mmm(const m33& a, const m33& b)
{
	m33 r;
	r.c0 = mmul(a, b.c0);
	r.c1 = mmul(a, b.c1);
	r.c2 = mmul(a, b.c2);
	return r;
}
PXH_D f3 xtrans(const xf& t, f3 v) { return qrot(t.q, v) + t.p; }
PXH_D f3 xtransinv(const xf& t, f3 v) { return qrotinv(t.q, v - t.p); }
// a^-1 * b (PxTransform::transformInv(const PxTransform&))
PXH_D xf xinvmul(const xf& a, const xf& b)
{
	qt qi; qi.x = -a.q.x; qi.y = -a.q.y; qi.z = -a.q.z; qi.w = a.q.w;
	xf r;
	r.q = qmul(qi, b.q);
	r.p = qrot(qi, b.p - a.p);
	return r;
}

struct mxf { m33 r; f3 p; };
PXH_D f3 mxtrans(const mxf& t, f3 v) { return mmul(t.r, v) + t.p; }
PXH_D mxf mxinvmul(const mxf& a, const mxf& b)
{
	mxf r;
	r.r = mmm(mtrans(a.r), b.r);
	r.p = mtmul(a.r, b.p - a.p);
	return r;
}

// order-preserving float -> u32 encoding (monotone: a <= b  <=>  enc(a) <= enc(b))
PXH_D uint32_t encode_float(float f)
{
	const uint32_t u = __float_as_uint(f);
	return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

PXH_D uint64_t mix64(uint64_t z)
{
	z += 0x9E3779B97F4A7C15ull;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
	return z ^ (z >> 31);
}

}  // namespace pxh
