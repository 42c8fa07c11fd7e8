// Convex hull cooking (host): incremental hull in double precision, coplanar
// triangles merged into polygons, then the per-hull data the device
// narrowphase reads. Replaces the convex cooking of PxCreateConvexMesh
// (physx/include/cooking/PxCooking.h) for this path.
#include "pxh_hull.h"
#include "../../include/pxhip.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <map>
#include <numeric>
#include <utility>

namespace pxh {
namespace {

struct D3
{
	double x, y, z;
};
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 cross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
double len(D3 a) { return std::sqrt(dot(a, a)); }

struct Tri
{
	int v[3];
	D3 n;
	double d;
	bool alive;
};

Tri makeTri(const std::vector<D3>& P, int a, int b, int c)
{
	Tri t;
	t.v[0] = a; t.v[1] = b; t.v[2] = c;
	D3 n = cross(P[b] - P[a], P[c] - P[a]);
	const double l = len(n);
	t.n = l > 0.0 ? n * (1.0 / l) : D3{0, 0, 0};
	t.d = -dot(t.n, P[a]);
	t.alive = true;
	return t;
}

double dist(const Tri& t, D3 p) { return dot(t.n, p) + t.d; }

struct FaceLoop
{
	std::vector<int> v;   // counter-clockwise point indices
	float n[3];
	float area;           // length of the unnormalised normal
	float offset;         // n . centroid
	bool alive;
};

float dotf(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// normal summed over the loop against its longest leading edge (the longest
// of the first three), normalised; offset through the vertex centroid
void faceNormal(const float* pts, FaceLoop& f)
{
	const size_t m = f.v.size();
	size_t start = 0;
	float best = -1.0f;
	for (size_t k = 0; k < 3; k++)
	{
		const float* a = pts + 3 * f.v[k];
		const float* b = pts + 3 * f.v[(k + 1) % m];
		const float dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
		const float d = dx * dx + dy * dy + dz * dz;
		if (d > best) { best = d; start = k; }
	}
	const float* p0 = pts + 3 * f.v[start];
	const float* p1 = pts + 3 * f.v[(start + 1) % m];
	const float d[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
	float n[3] = {0.0f, 0.0f, 0.0f};
	float c[3] = {p0[0], p0[1], p0[2]};
	int cnt = 1;
	for (size_t j = 1; j < m; j++)
	{
		const float* pj = pts + 3 * f.v[(start + j) % m];
		const float* pn = pts + 3 * f.v[(start + j + 1) % m];
		c[0] += pj[0]; c[1] += pj[1]; c[2] += pj[2];
		cnt++;
		const float e[3] = {pn[0] - p0[0], pn[1] - p0[1], pn[2] - p0[2]};
		n[0] += d[1] * e[2] - d[2] * e[1];
		n[1] += d[2] * e[0] - d[0] * e[2];
		n[2] += d[0] * e[1] - d[1] * e[0];
	}
	const float len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
	if (len > 0.0f)
	{
		const float inv = 1.0f / len;
		n[0] *= inv; n[1] *= inv; n[2] *= inv;
	}
	f.area = len;
	const float ic = 1.0f / float(cnt);
	c[0] *= ic; c[1] *= ic; c[2] *= ic;
	f.n[0] = n[0]; f.n[1] = n[1]; f.n[2] = n[2];
	f.offset = n[0] * c[0] + n[1] * c[1] + n[2] * c[2];
}

}  // namespace

void hull_edges_bounds(Hull& out)
{
	// edges in order of first appearance around the polygons: v0 -> v1 runs
	// counter-clockwise in the face on the left, the reverse in the other
	out.edgeVerts.clear();
	out.edgeFaces.clear();
	std::map<std::pair<int, int>, int> edgeIndex;
	for (size_t p = 0; p < out.polys.size(); p++)
	{
		const HullPolygon& hp = out.polys[p];
		for (int k = 0; k < hp.count; k++)
		{
			const int a = out.vref[hp.first + k], b = out.vref[hp.first + (k + 1) % hp.count];
			auto it = edgeIndex.find({b, a});
			if (it != edgeIndex.end())
				out.edgeFaces[2 * it->second + 1] = uint8_t(p);
			else
			{
				edgeIndex[{a, b}] = int(out.edgeVerts.size() / 2);
				out.edgeVerts.push_back(uint8_t(a));
				out.edgeVerts.push_back(uint8_t(b));
				out.edgeFaces.push_back(uint8_t(p));
				out.edgeFaces.push_back(uint8_t(p));
			}
		}
	}
	// local bounds from the float vertices
	float mn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, mx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (size_t k = 0; k < out.verts.size() / 3; k++)
		for (int c = 0; c < 3; c++)
		{
			mn[c] = std::min(mn[c], out.verts[3 * k + c]);
			mx[c] = std::max(mx[c], out.verts[3 * k + c]);
		}
	for (int c = 0; c < 3; c++)
	{
		out.aabbCenter[c] = (mx[c] + mn[c]) * 0.5f;
		out.aabbExtents[c] = (mx[c] - mn[c]) * 0.5f;
	}
}

void hull_internal_objects(Hull& out, const float com[3])
{
	// internal objects: the largest sphere about the centre of mass inside the
	// hull, and a box of that radius grown along the largest bounds axis, then
	// along the other two (ConvexMeshBuilder::computeInternalObjects,
	// geomutils/src/cooking/GuCookingConvexMeshBuilder.cpp:420-531)
	float ir = FLT_MAX;
	for (const HullPolygon& hp : out.polys)
	{
		const float d = std::fabs(hp.plane[0] * com[0] + hp.plane[1] * com[1] + hp.plane[2] * com[2] + hp.plane[3]);
		ir = std::min(ir, d);
	}
	out.internalRadius = ir;
	const float e[3] = {2.0f * out.aabbExtents[0], 2.0f * out.aabbExtents[1], 2.0f * out.aabbExtents[2]};
	const int L = (e[0] >= e[1] && e[0] >= e[2]) ? 0 : (e[1] >= e[2] ? 1 : 2);
	int e0 = (L + 1) % 3, e1 = (e0 + 1) % 3;
	if (e[e0] < e[e1])
		std::swap(e0, e1);
	const float r = ir / std::sqrt(3.0f);
	const float tiny = 1e-7f;
	float ext[3] = {FLT_MAX, FLT_MAX, FLT_MAX};
	for (const HullPolygon& hp : out.polys)
	{
		const float dl = hp.plane[L];
		if (-tiny < dl && dl < tiny)
			continue;
		const float base = -hp.plane[3] - (hp.plane[0] * com[0] + hp.plane[1] * com[1] + hp.plane[2] * com[2]);
		const float inv = 1.0f / dl;
		const float a0 = r * hp.plane[e0], a1 = r * hp.plane[e1];
		const float nums[4] = {base - a0 - a1, base - a0 + a1, base + a0 + a1, base + a0 - a1};
		for (float num : nums)
			ext[L] = std::min(ext[L], std::max(std::fabs(num * inv), r));
	}
	for (const HullPolygon& hp : out.polys)
	{
		const float add = hp.plane[e0] + hp.plane[e1];
		const float sub = hp.plane[e0] - hp.plane[e1];
		const float base = -hp.plane[3] - (hp.plane[0] * com[0] + hp.plane[1] * com[1] + hp.plane[2] * com[2]);
		const float b0 = ext[L] * hp.plane[L];
		for (float den : {add, sub})
		{
			if (-tiny < den && den < tiny)
				continue;
			ext[e0] = std::min(ext[e0], std::max(std::fabs((base - b0) / den), r));
			ext[e0] = std::min(ext[e0], std::max(std::fabs((base + b0) / den), r));
		}
	}
	ext[e1] = ext[e0];
	for (int c = 0; c < 3; c++)
		out.internalExtents[c] = ext[c];
}

int build_hull(const float* pts, int n, Hull& out)
{
	out = Hull();
	if (!pts || n < 4)
		return -1;
	std::vector<D3> P(static_cast<size_t>(n));
	double scale = 0.0;
	for (int i = 0; i < n; i++)
	{
		P[i] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
		scale = std::max(scale, std::fabs(P[i].x) + std::fabs(P[i].y) + std::fabs(P[i].z));
	}
	const double eps = std::max(scale, 1e-30) * 1e-9;

	// initial tetrahedron from extreme points
	int i0 = 0;
	for (int i = 1; i < n; i++)
		if (P[i].x < P[i0].x)
			i0 = i;
	int i1 = -1;
	double best = 0.0;
	for (int i = 0; i < n; i++)
	{
		const double d = len(P[i] - P[i0]);
		if (d > best) { best = d; i1 = i; }
	}
	if (i1 < 0 || best <= eps)
		return -2;
	int i2 = -1;
	best = 0.0;
	const D3 ax = (P[i1] - P[i0]) * (1.0 / len(P[i1] - P[i0]));
	for (int i = 0; i < n; i++)
	{
		const D3 w = P[i] - P[i0];
		const double d = len(w - ax * dot(w, ax));
		if (d > best) { best = d; i2 = i; }
	}
	if (i2 < 0 || best <= eps)
		return -2;
	Tri base = makeTri(P, i0, i1, i2);
	int i3 = -1;
	best = 0.0;
	for (int i = 0; i < n; i++)
	{
		const double d = std::fabs(dist(base, P[i]));
		if (d > best) { best = d; i3 = i; }
	}
	if (i3 < 0 || best <= eps)
		return -2;
	if (dist(base, P[i3]) > 0.0)
		std::swap(i1, i2);  // keep the fourth point below the base
	std::vector<Tri> T;
	T.push_back(makeTri(P, i0, i1, i2));
	T.push_back(makeTri(P, i0, i3, i1));
	T.push_back(makeTri(P, i1, i3, i2));
	T.push_back(makeTri(P, i2, i3, i0));
	std::vector<char> used(static_cast<size_t>(n), 0);
	used[i0] = used[i1] = used[i2] = used[i3] = 1;

	// add the farthest outside point until none is left
	std::vector<int> edgeFace(size_t(n) * n, -1);  // directed edge (a, b) -> face
	for (;;)
	{
		int pick = -1;
		double far = eps;
		for (int i = 0; i < n; i++)
		{
			if (used[i])
				continue;
			for (const Tri& t : T)
				if (t.alive)
				{
					const double d = dist(t, P[i]);
					if (d > far) { far = d; pick = i; }
				}
		}
		if (pick < 0)
			break;
		used[pick] = 1;
		std::fill(edgeFace.begin(), edgeFace.end(), -1);
		for (int f = 0; f < int(T.size()); f++)
			if (T[f].alive)
				for (int k = 0; k < 3; k++)
					edgeFace[size_t(T[f].v[k]) * n + T[f].v[(k + 1) % 3]] = f;
		std::vector<char> vis(T.size(), 0);
		for (int f = 0; f < int(T.size()); f++)
			vis[f] = T[f].alive && dist(T[f], P[pick]) > eps;
		std::vector<std::pair<int, int>> horizon;
		for (int f = 0; f < int(T.size()); f++)
			if (vis[f])
				for (int k = 0; k < 3; k++)
				{
					const int a = T[f].v[k], b = T[f].v[(k + 1) % 3];
					const int g = edgeFace[size_t(b) * n + a];
					if (g < 0 || !vis[g])
						horizon.push_back({a, b});
				}
		for (int f = 0; f < int(T.size()); f++)
			if (vis[f])
				T[f].alive = false;
		for (const auto& e : horizon)
			T.push_back(makeTri(P, pick, e.first, e.second));  // loops start at the added point
		// points left inside the hull are never picked again
	}
	std::vector<Tri> tris;
	for (const Tri& t : T)
		if (t.alive)
			tris.push_back(t);

	// Polygons: the hull's triangles, then adjacent faces whose normals lie
	// within 3 degrees merged while the merged face stays planar to the plane
	// tolerance and convex, the larger face absorbing the smaller (the post
	// merge of the reference's quickhull cooking, QuickHull::postMergeHull /
	// canMergeFaces, geomutils/src/cooking/GuCookingQuickHullConvexHullLib.cpp:1121-1595).
	// Face planes use the same float construction as the cooked hull's
	// (QuickHullFace::computeNormalAndCentroid, :296-336): normal summed over
	// the loop against the face's longest leading edge, plane through the
	// vertex centroid.
	std::vector<FaceLoop> faces;
	for (const Tri& t : tris)
		faces.push_back(FaceLoop{{t.v[0], t.v[1], t.v[2]}, {0, 0, 0}, 0.0f, 0.0f, true});
	for (FaceLoop& f : faces)
		faceNormal(pts, f);
	float bmn[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, bmx[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
	for (int i = 0; i < n; i++)
		for (int c = 0; c < 3; c++)
		{
			bmn[c] = std::min(bmn[c], pts[3 * i + c]);
			bmx[c] = std::max(bmx[c], pts[3 * i + c]);
		}
	const float sizeTol = (bmx[0] - bmn[0] + bmx[1] - bmn[1] + bmx[2] - bmn[2]) * 0.5f;
	const float thick = std::max(3.0f * FLT_EPSILON * sizeTol, 3.0f * FLT_EPSILON);
	const float planeTol = std::max(0.0007f * sizeTol, 0.0007f);
	const float maxDot = std::cos(3.0f * 3.14159265358979323846f / 180.0f);
	std::vector<int> hullVerts;
	for (const Tri& t : tris)
		for (int k = 0; k < 3; k++)
			hullVerts.push_back(t.v[k]);
	std::sort(hullVerts.begin(), hullVerts.end());
	hullVerts.erase(std::unique(hullVerts.begin(), hullVerts.end()), hullVerts.end());
	std::vector<int> owner(size_t(n) * n, -1);  // directed edge (a, b) -> face
	for (int f = 0; f < int(faces.size()); f++)
		for (size_t k = 0; k < faces[f].v.size(); k++)
			owner[size_t(faces[f].v[k]) * n + faces[f].v[(k + 1) % faces[f].v.size()]] = f;
	auto facesAt = [&](int v) {
		int c = 0;
		for (const FaceLoop& f : faces)
			if (f.alive && std::find(f.v.begin(), f.v.end(), v) != f.v.end())
				c++;
		return c;
	};
	for (int fi = 0; fi < int(faces.size()); fi++)
	{
		bool merged = true;
		while (merged && faces[fi].alive)
		{
			merged = false;
			FaceLoop& f = faces[fi];
			const size_t m = f.v.size();
			for (size_t k = 0; k < m && !merged; k++)
			{
				const int a = f.v[k], b = f.v[(k + 1) % m];
				const int gi = owner[size_t(b) * n + a];
				FaceLoop& g = faces[gi];
				if (!(dotf(f.n, g.n) > maxDot) || !(f.area >= g.area))
					continue;
				// a vertex shared only by the two faces and one other would be left
				// with two faces
				if (facesAt(a) <= 3 || facesAt(b) <= 3)
					continue;
				// merged loop: f from its first vertex (or the next one when the
				// shared edge leads the loop), g spliced in at the shared edge
				FaceLoop mf;
				mf.alive = true;
				const size_t s0 = (k == 0) ? 1 : 0;
				for (size_t j = 0; j < m; j++)
				{
					const size_t idx = (s0 + j) % m;
					mf.v.push_back(f.v[idx]);
					if (f.v[idx] == a)
					{
						// g runs b -> a; insert its vertices after a up to b (exclusive)
						const size_t gm = g.v.size();
						size_t ga = 0;
						while (g.v[ga] != a)
							ga++;
						for (size_t t = 1; t + 1 < gm; t++)
							mf.v.push_back(g.v[(ga + t) % gm]);
					}
				}
				faceNormal(pts, mf);
				bool ok = true;
				for (int v : hullVerts)
					if (dotf(mf.n, pts + 3 * v) - mf.offset > planeTol)
					{
						ok = false;
						break;
					}
				for (size_t e = 0; e < mf.v.size() && ok; e++)
				{
					const float* p0 = pts + 3 * mf.v[e];
					const float* p1 = pts + 3 * mf.v[(e + 1) % mf.v.size()];
					float ed[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
					const float el = std::sqrt(ed[0] * ed[0] + ed[1] * ed[1] + ed[2] * ed[2]);
					if (el > 0.0f)
						for (float& x : ed)
							x *= 1.0f / el;
					const float out[3] = {-(mf.n[1] * ed[2] - mf.n[2] * ed[1]), -(mf.n[2] * ed[0] - mf.n[0] * ed[2]),
					                      -(mf.n[0] * ed[1] - mf.n[1] * ed[0])};
					for (int v : mf.v)
					{
						const float* q = pts + 3 * v;
						const float dd = (q[0] - p0[0]) * out[0] + (q[1] - p0[1]) * out[1] + (q[2] - p0[2]) * out[2];
						if (dd > thick)
						{
							ok = false;
							break;
						}
					}
				}
				if (!ok)
					continue;
				g.alive = false;
				f = mf;
				for (size_t j = 0; j < f.v.size(); j++)
					owner[size_t(f.v[j]) * n + f.v[(j + 1) % f.v.size()]] = fi;
				merged = true;
			}
		}
	}
	std::vector<std::vector<int>> loops;
	std::vector<const FaceLoop*> kept;
	for (const FaceLoop& f : faces)
		if (f.alive)
		{
			loops.push_back(f.v);
			kept.push_back(&f);
		}
	if (loops.size() > 255)
		return -4;

	// vertex numbering in order of first use by the polygons
	std::vector<int> local(static_cast<size_t>(n), -1);
	std::vector<int> order;
	for (const auto& loop : loops)
		for (int v : loop)
			if (local[v] < 0)
			{
				local[v] = int(order.size());
				order.push_back(v);
			}
	if (order.size() > 255)
		return -4;
	out.verts.resize(order.size() * 3);
	for (size_t k = 0; k < order.size(); k++)
		for (int c = 0; c < 3; c++)
			out.verts[3 * k + c] = pts[3 * order[k] + c];
	for (size_t p = 0; p < loops.size(); p++)
	{
		const auto& loop = loops[p];
		HullPolygon hp;
		hp.plane[0] = kept[p]->n[0];
		hp.plane[1] = kept[p]->n[1];
		hp.plane[2] = kept[p]->n[2];
		hp.plane[3] = -kept[p]->offset;
		hp.first = int(out.vref.size());
		hp.count = int(loop.size());
		for (int v : loop)
			out.vref.push_back(uint8_t(local[v]));
		out.polys.push_back(hp);
	}
	hull_edges_bounds(out);

	// mass properties at unit density of the hull of the points: tetrahedra
	// (origin, a, b, c) over the hull's triangles; second moments through the
	// covariance of the canonical tetrahedron (diagonal 1/60, off-diagonal 1/120)
	double V = 0.0, m1[3] = {0, 0, 0}, C[9] = {0};
	for (const Tri& tri : tris)
		{
			const D3 a = P[tri.v[0]], b = P[tri.v[1]], c = P[tri.v[2]];
			const double det = dot(a, cross(b, c));
			V += det / 6.0;
			const D3 s = a + b + c;
			m1[0] += det / 24.0 * s.x;
			m1[1] += det / 24.0 * s.y;
			m1[2] += det / 24.0 * s.z;
			const double A[3][3] = {{a.x, b.x, c.x}, {a.y, b.y, c.y}, {a.z, b.z, c.z}};
			for (int r = 0; r < 3; r++)
				for (int q = 0; q < 3; q++)
				{
					double acc = 0.0;
					for (int u = 0; u < 3; u++)
						for (int w = 0; w < 3; w++)
							acc += A[r][u] * (u == w ? 1.0 / 60.0 : 1.0 / 120.0) * A[q][w];
					C[3 * r + q] += det * acc;
				}
		}
	out.volume = V;
	for (int c = 0; c < 3; c++)
		out.com[c] = V != 0.0 ? m1[c] / V : 0.0;
	double Cc[9];
	for (int r = 0; r < 3; r++)
		for (int q = 0; q < 3; q++)
			Cc[3 * r + q] = C[3 * r + q] - V * out.com[r] * out.com[q];
	const double tr = Cc[0] + Cc[4] + Cc[8];
	for (int r = 0; r < 3; r++)
		for (int q = 0; q < 3; q++)
			out.inertia[3 * r + q] = (r == q ? tr : 0.0) - Cc[3 * r + q];

	const float com[3] = {float(out.com[0]), float(out.com[1]), float(out.com[2])};
	hull_internal_objects(out, com);
	return 0;
}

float hull_margin(const Hull& h, float toleranceLength)
{
	const float m = std::min(h.internalExtents[0], std::min(h.internalExtents[1], h.internalExtents[2]));
	return std::min(m * 0.25f, toleranceLength * 0.05f);
}

}  // namespace pxh

extern "C" int32_t pxh_convex_hull(const float* pts, int32_t n, float* verts, float* planes, int32_t* polys,
                                   int32_t* vrefs, int32_t* edges, double* mass, float* shape, int32_t* counts)
{
	pxh::Hull h;
	const int r = pxh::build_hull(pts, n, h);
	if (r)
		return -1;  // PXH_ERR_INVALID: degenerate or oversized point cloud
	const int nv = int(h.verts.size() / 3), np = int(h.polys.size()), nr = int(h.vref.size());
	const int ne = int(h.edgeVerts.size() / 2);
	if (verts)
		std::copy(h.verts.begin(), h.verts.end(), verts);
	for (int p = 0; p < np; p++)
	{
		if (planes)
			std::copy(h.polys[p].plane, h.polys[p].plane + 4, planes + 4 * p);
		if (polys)
		{
			polys[2 * p] = h.polys[p].first;
			polys[2 * p + 1] = h.polys[p].count;
		}
	}
	if (vrefs)
		for (int k = 0; k < nr; k++)
			vrefs[k] = h.vref[k];
	if (edges)
		for (int e = 0; e < ne; e++)
		{
			edges[4 * e] = h.edgeVerts[2 * e];
			edges[4 * e + 1] = h.edgeVerts[2 * e + 1];
			edges[4 * e + 2] = h.edgeFaces[2 * e];
			edges[4 * e + 3] = h.edgeFaces[2 * e + 1];
		}
	if (mass)
	{
		mass[0] = h.volume;
		for (int c = 0; c < 3; c++)
			mass[1 + c] = h.com[c];
		for (int c = 0; c < 9; c++)
			mass[4 + c] = h.inertia[c];
	}
	if (shape)
	{
		for (int c = 0; c < 3; c++)
		{
			shape[c] = h.aabbCenter[c];
			shape[3 + c] = h.aabbExtents[c];
			shape[7 + c] = h.internalExtents[c];
		}
		shape[6] = h.internalRadius;
	}
	if (counts)
	{
		counts[0] = nv;
		counts[1] = np;
		counts[2] = nr;
		counts[3] = ne;
	}
	return 0;
}
