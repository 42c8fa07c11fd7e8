// Convex hull cooking on the host: the hull of a point cloud as vertices,
// polygons (outward plane + counter-clockwise vertex loop), edges with their
// two faces, the local bounds, mass properties at unit density and the
// internal extents that set the contact-generation margin.
#pragma once

#include <cstdint>
#include <vector>

namespace pxh {

struct HullPolygon
{
	float plane[4];        // n xyz (unit, outward), d with n.x + d = 0
	int first, count;      // range in Hull::vref
};

struct Hull
{
	std::vector<float> verts;          // 3 per vertex, bit-identical to input points
	std::vector<HullPolygon> polys;
	std::vector<uint8_t> vref;         // polygon vertex loops
	std::vector<uint8_t> edgeFaces;    // 2 per edge: face on the left, face on the right
	std::vector<uint8_t> edgeVerts;    // 2 per edge: v0, v1 (v0 -> v1 runs counter-clockwise in face 0)
	float aabbCenter[3], aabbExtents[3];
	double volume;                     // mass at density 1
	double com[3];
	double inertia[9];                 // about the centre of mass, unit density, row-major
	float internalRadius;
	float internalExtents[3];
};

// Builds the hull of n points (xyz). Returns 0, or a negative code when the
// points are degenerate (fewer than 4 non-coplanar) or the hull exceeds 255
// vertices or polygons.
int build_hull(const float* pts, int n, Hull& out);

// Edges (with their two faces) and local bounds of a hull whose verts, polys
// and vref are set.
void hull_edges_bounds(Hull& out);

// Internal radius and extents about the centre of mass com of a hull whose
// polygons and bounds are set (ConvexMeshBuilder::computeInternalObjects,
// geomutils/src/cooking/GuCookingConvexMeshBuilder.cpp:420-531).
void hull_internal_objects(Hull& out, const float com[3]);

// PCM margin of a hull: a quarter of the smallest internal extent, capped at
// 5% of the tolerance length.
float hull_margin(const Hull& h, float toleranceLength);

}  // namespace pxh
