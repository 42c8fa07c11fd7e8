// C-ABI harness over the reference PhysX 5.6.1 CPU SDK, built from the
// reference sources by oracle/ref.mk into oracle/_ref/libpxref.so.
//
// Test infrastructure only: tests/ and bench.py's CPU-baseline leg load it as
// the checker and the timed CPU baseline; the product library never links it.
//
// Entry points:
//   pxref_bp_pairs          standalone ABP broadphase (PxCreateBroadPhase,
//                           physx/include/PxBroadPhase.h:629) on given bounds
//   pxref_generate_contacts immediate-mode PCM contact generation on a fresh
//                           cache (PxGenerateContacts, physx/include/PxImmediateMode.h:267)
//   pxref_scene_*           a CPU PxScene (ABP/PABP, PCM, PGS/TGS) built from
//                           flat arrays, stepped with simulate()+fetchResults()
#include "PxPhysicsAPI.h"
#include "PxImmediateMode.h"
#include "cooking/PxCooking.h"
#include "GuConvexMesh.h"
#include <chrono>
#include <cstdint>
#include <cstring>
#include <vector>

using namespace physx;
using namespace physx::immediate;

namespace {

PxDefaultAllocator gAllocator;
PxDefaultErrorCallback gErrorCallback;
PxFoundation* gFoundation = nullptr;
PxPhysics* gPhysics = nullptr;

bool ensureSdk()
{
	if (gPhysics)
		return true;
	gFoundation = PxCreateFoundation(PX_PHYSICS_VERSION, gAllocator, gErrorCallback);
	if (!gFoundation)
		return false;
	gPhysics = PxCreatePhysics(PX_PHYSICS_VERSION, *gFoundation, PxTolerancesScale(), false, nullptr);
	return gPhysics != nullptr;
}

// convex meshes cooked by pxref_convex_create, named by index
std::vector<PxConvexMesh*> gConvex;
std::vector<PxTriangleMesh*> gTriMesh;
std::vector<PxHeightField*> gHeightField;

// geometry codes shared with the product C-ABI: 0 sphere, 1 plane, 2 capsule,
// 3 box, 5 convex mesh (g[0] = index of a mesh from pxref_convex_create),
// 8 triangle mesh (g[0] = index of a mesh from pxref_trimesh_create),
// 9 heightfield (g[0] = index from pxref_heightfield_create, g[1] height
// scale, g[2] row scale, g[3] column scale)
PxGeometryHolder makeGeom(int type, const float* g)
{
	switch (type)
	{
	case 0: return PxGeometryHolder(PxSphereGeometry(g[0]));
	case 1: return PxGeometryHolder(PxPlaneGeometry());
	case 2: return PxGeometryHolder(PxCapsuleGeometry(g[0], g[1]));
	case 5: return PxGeometryHolder(PxConvexMeshGeometry(gConvex[size_t(g[0])]));
	case 8: return PxGeometryHolder(PxTriangleMeshGeometry(gTriMesh[size_t(g[0])]));
	case 9: return PxGeometryHolder(PxHeightFieldGeometry(gHeightField[size_t(g[0])], PxMeshGeometryFlags(), g[1], g[2], g[3]));
	default: return PxGeometryHolder(PxBoxGeometry(g[0], g[1], g[2]));
	}
}

PxTransform makePose(const float* p)  // px py pz qx qy qz qw
{
	return PxTransform(PxVec3(p[0], p[1], p[2]), PxQuat(p[3], p[4], p[5], p[6]));
}

struct Recorder : PxContactRecorder
{
	float* out;
	int maxContacts;
	int count = 0;
	bool recordContacts(const PxContactPoint* cp, PxU32 n, PxU32) override
	{
		for (PxU32 i = 0; i < n && count < maxContacts; i++, count++)
		{
			float* o = out + 8 * count;
			o[0] = cp[i].point.x; o[1] = cp[i].point.y; o[2] = cp[i].point.z; o[3] = cp[i].separation;
			o[4] = cp[i].normal.x; o[5] = cp[i].normal.y; o[6] = cp[i].normal.z; o[7] = 0.f;
		}
		return true;
	}
};

struct CacheAlloc : PxCacheAllocator
{
	std::vector<std::vector<PxU8>> blocks;
	PxU8* allocateCacheData(const PxU32 byteSize) override
	{
		blocks.emplace_back(byteSize + 16);
		PxU8* p = blocks.back().data();
		return reinterpret_cast<PxU8*>((reinterpret_cast<uintptr_t>(p) + 15) & ~uintptr_t(15));
	}
};

// Contact points the scene reports for every touching pair of a step
// (actor ids: dynamic body i -> i + 1, static plane j -> -(j + 1)).
struct ContactCollector : public PxSimulationEventCallback
{
	std::vector<float> rows;  // a, b, point xyz, separation, normal xyz, impulse
	void onContact(const PxContactPairHeader& h, const PxContactPair* pairs, PxU32 nb) override
	{
		const float a = float(reinterpret_cast<intptr_t>(h.actors[0]->userData));
		const float b = float(reinterpret_cast<intptr_t>(h.actors[1]->userData));
		PxContactPairPoint pts[64];
		for (PxU32 i = 0; i < nb; i++)
		{
			const PxU32 k = pairs[i].extractContacts(pts, 64);
			for (PxU32 j = 0; j < k; j++)
			{
				const float r[10] = {a, b, pts[j].position.x, pts[j].position.y, pts[j].position.z, pts[j].separation,
				                     pts[j].normal.x, pts[j].normal.y, pts[j].normal.z, pts[j].impulse.magnitude()};
				rows.insert(rows.end(), r, r + 10);
			}
		}
	}
	void onConstraintBreak(PxConstraintInfo*, PxU32) override {}
	void onWake(PxActor**, PxU32) override {}
	void onSleep(PxActor**, PxU32) override {}
	void onTrigger(PxTriggerPair*, PxU32) override {}
	void onAdvance(const PxRigidBody* const*, const PxTransform*, const PxU32) override {}
};

PxFilterFlags reportingFilterShader(PxFilterObjectAttributes a0, PxFilterData d0, PxFilterObjectAttributes a1,
                                    PxFilterData d1, PxPairFlags& pairFlags, const void* block, PxU32 size)
{
	const PxFilterFlags f = PxDefaultSimulationFilterShader(a0, d0, a1, d1, pairFlags, block, size);
	pairFlags |= PxPairFlag::eNOTIFY_TOUCH_FOUND | PxPairFlag::eNOTIFY_TOUCH_PERSISTS |
	             PxPairFlag::eNOTIFY_CONTACT_POINTS;
	return f;
}

struct RefScene
{
	PxDefaultCpuDispatcher* dispatcher = nullptr;
	std::vector<PxDefaultCpuDispatcher*> retired;   // dispatchers replaced by pxref_scene_set_threads
	PxScene* scene = nullptr;
	PxMaterial* material = nullptr;
	std::vector<PxRigidDynamic*> bodies;
	int nPlanes = 0;
	int joints = 0;
	std::vector<PxJoint*> jointList;
	ContactCollector* contacts = nullptr;
};

}  // namespace

extern "C" {

int pxref_version() { return PX_PHYSICS_VERSION; }

// Standalone ABP over n boxes. bounds: n x (minx miny minz maxx maxy maxz);
// dist: n inflation distances; groups: n PxBpFilterGroup values (use 0 for
// statics, 2*i+... for dynamics, as PxGetBroadPhase*FilterGroup returns).
// Writes up to maxPairs (id0,id1) pairs with id0<id1; returns the pair count
// (which may exceed maxPairs) or -1 on failure.
long long pxref_bp_pairs(const float* bounds, const float* dist, const uint32_t* groups, int n,
                         uint32_t* outPairs, long long maxPairs, int usePabp)
{
	if (!ensureSdk())
		return -1;
	PxBroadPhaseDesc desc(usePabp ? PxBroadPhaseType::ePABP : PxBroadPhaseType::eABP);
	PxBroadPhase* bp = PxCreateBroadPhase(desc);
	if (!bp)
		return -1;
	// the API may read 4 bytes past the last bound: copy into a padded array
	std::vector<PxBounds3> b(size_t(n) + 1);
	for (int i = 0; i < n; i++)
		b[i] = PxBounds3(PxVec3(bounds[6 * i], bounds[6 * i + 1], bounds[6 * i + 2]),
		                 PxVec3(bounds[6 * i + 3], bounds[6 * i + 4], bounds[6 * i + 5]));
	std::vector<PxBpIndex> created(n);
	for (int i = 0; i < n; i++)
		created[i] = PxBpIndex(i);
	PxBroadPhaseUpdateData ud(created.data(), PxU32(n), nullptr, 0, nullptr, 0, b.data(),
	                          reinterpret_cast<const PxBpFilterGroup*>(groups), dist, PxU32(n));
	bp->update(ud);
	PxBroadPhaseResults res;
	bp->fetchResults(res);
	long long count = res.mNbCreatedPairs;
	for (long long i = 0; i < count && i < maxPairs; i++)
	{
		PxU32 a = res.mCreatedPairs[i].mID0, c = res.mCreatedPairs[i].mID1;
		outPairs[2 * i] = a < c ? a : c;
		outPairs[2 * i + 1] = a < c ? c : a;
	}
	bp->release();
	return count;
}

// Cooks the hull of n points (PxCreateConvexMesh, eCOMPUTE_CONVEX, default
// cooking parameters) and returns its mesh index, or -1.
int pxref_convex_create(const float* pts, int n)
{
	if (!ensureSdk())
		return -1;
	PxConvexMeshDesc desc;
	desc.points.count = PxU32(n);
	desc.points.stride = sizeof(PxVec3);
	desc.points.data = pts;
	desc.flags = PxConvexFlag::eCOMPUTE_CONVEX;
	PxCookingParams params{PxTolerancesScale()};
	PxConvexMesh* m = PxCreateConvexMesh(params, desc, gPhysics->getPhysicsInsertionCallback());
	if (!m)
		return -1;
	gConvex.push_back(m);
	return int(gConvex.size()) - 1;
}

// A cooked hull read back: verts[nv*3], planes[np*4], polys[np*2] (first,
// count into vrefs), vrefs[], mass[13] (mass at unit density, centre of mass,
// inertia 3x3 row-major), shape[10] (bounds centre, extents, internal radius,
// internal extents), counts[3] = nv, np, vrefs.
int pxref_convex_data(int id, float* verts, float* planes, int* polys, int* vrefs, double* mass, float* shape,
                      int* counts)
{
	if (id < 0 || size_t(id) >= gConvex.size())
		return -1;
	PxConvexMesh* m = gConvex[size_t(id)];
	const PxU32 nv = m->getNbVertices(), np = m->getNbPolygons();
	const PxVec3* v = m->getVertices();
	for (PxU32 i = 0; i < nv; i++)
	{
		verts[3 * i] = v[i].x; verts[3 * i + 1] = v[i].y; verts[3 * i + 2] = v[i].z;
	}
	const PxU8* ib = m->getIndexBuffer();
	int nr = 0;
	for (PxU32 p = 0; p < np; p++)
	{
		PxHullPolygon hp;
		m->getPolygonData(p, hp);
		for (int c = 0; c < 4; c++)
			planes[4 * p + c] = hp.mPlane[c];
		polys[2 * p] = hp.mIndexBase;
		polys[2 * p + 1] = hp.mNbVerts;
		for (int k = 0; k < hp.mNbVerts; k++)
			vrefs[hp.mIndexBase + k] = ib[hp.mIndexBase + k];
		nr = PxMax(nr, int(hp.mIndexBase + hp.mNbVerts));
	}
	PxReal ms;
	PxMat33 I;
	PxVec3 com;
	m->getMassInformation(ms, I, com);
	mass[0] = ms;
	mass[1] = com.x; mass[2] = com.y; mass[3] = com.z;
	for (int r = 0; r < 3; r++)
		for (int c = 0; c < 3; c++)
			mass[4 + 3 * r + c] = I[c][r];
	const Gu::ConvexHullData& h = static_cast<const Gu::ConvexMesh*>(m)->getHull();
	const PxVec3 bc = h.mAABB.mCenter, be = h.mAABB.mExtents;
	shape[0] = bc.x; shape[1] = bc.y; shape[2] = bc.z;
	shape[3] = be.x; shape[4] = be.y; shape[5] = be.z;
	shape[6] = h.mInternal.mInternalRadius;
	shape[7] = h.mInternal.mInternalExtents.x;
	shape[8] = h.mInternal.mInternalExtents.y;
	shape[9] = h.mInternal.mInternalExtents.z;
	counts[0] = int(nv);
	counts[1] = int(np);
	counts[2] = nr;
	return 0;
}

// Cooks a triangle mesh (PxCreateTriangleMesh, default cooking parameters:
// BV4 midphase, active-edge data) from nv vertices and nt index triples;
// returns its mesh index (geometry code 8 with dims[0] = index), or -1.
int pxref_trimesh_create(const float* verts, int nv, const int* tris, int nt)
{
	if (!ensureSdk())
		return -1;
	PxTriangleMeshDesc desc;
	desc.points.count = PxU32(nv);
	desc.points.stride = sizeof(PxVec3);
	desc.points.data = verts;
	desc.triangles.count = PxU32(nt);
	desc.triangles.stride = 3 * sizeof(PxU32);
	desc.triangles.data = tris;
	PxCookingParams params{PxTolerancesScale()};
	PxTriangleMesh* m = PxCreateTriangleMesh(params, desc, gPhysics->getPhysicsInsertionCallback());
	if (!m)
		return -1;
	gTriMesh.push_back(m);
	return int(gTriMesh.size()) - 1;
}

// Creates a heightfield of rows x cols samples (PxCreateHeightField):
// heights[rows*cols] row-major, tess[rows*cols] (may be null) the sample's
// tessellation flag (the cell's diagonal from its zeroth vertex); returns its
// index (geometry code 9), or -1.
int pxref_heightfield_create(int rows, int cols, const int16_t* heights, const uint8_t* tess)
{
	if (!ensureSdk())
		return -1;
	std::vector<PxHeightFieldSample> samples(size_t(rows) * cols);
	for (size_t i = 0; i < samples.size(); i++)
	{
		samples[i].height = heights[i];
		samples[i].materialIndex0 = 0;
		samples[i].materialIndex1 = 0;
		if (tess && tess[i])
			samples[i].setTessFlag();
	}
	PxHeightFieldDesc desc;
	desc.nbRows = PxU32(rows);
	desc.nbColumns = PxU32(cols);
	desc.samples.data = samples.data();
	desc.samples.stride = sizeof(PxHeightFieldSample);
	PxHeightField* hf = PxCreateHeightField(desc, gPhysics->getPhysicsInsertionCallback());
	if (!hf)
		return -1;
	gHeightField.push_back(hf);
	return int(gHeightField.size()) - 1;
}

// Adds a static actor with one heightfield shape (geom as for code 9) at pose.
int pxref_scene_add_heightfield(void* h, const float* geom, const float* pose, float contactOffset)
{
	RefScene* rs = static_cast<RefScene*>(h);
	if (geom[0] < 0.0f || size_t(geom[0]) >= gHeightField.size())
		return -1;
	PxGeometryHolder g = makeGeom(9, geom);
	PxRigidStatic* s = gPhysics->createRigidStatic(makePose(pose));
	PxShape* shape = PxRigidActorExt::createExclusiveShape(*s, g.any(), *rs->material);
	shape->setContactOffset(contactOffset);
	shape->setRestOffset(0.0f);
	s->userData = reinterpret_cast<void*>(intptr_t(-(++rs->nPlanes)));
	rs->scene->addActor(*s);
	return 0;
}

// Adds a static actor with one triangle mesh shape (identity scale) at pose;
// the shape's contact offset as given, rest offset 0.
int pxref_scene_add_trimesh(void* h, int mesh, const float* pose, float contactOffset)
{
	RefScene* rs = static_cast<RefScene*>(h);
	if (mesh < 0 || size_t(mesh) >= gTriMesh.size())
		return -1;
	PxRigidStatic* s = gPhysics->createRigidStatic(makePose(pose));
	PxShape* shape = PxRigidActorExt::createExclusiveShape(*s, PxTriangleMeshGeometry(gTriMesh[size_t(mesh)]),
	                                                       *rs->material);
	shape->setContactOffset(contactOffset);
	shape->setRestOffset(0.0f);
	s->userData = reinterpret_cast<void*>(intptr_t(-(++rs->nPlanes)));
	rs->scene->addActor(*s);
	return 0;
}

uint32_t pxref_static_group() { return PxGetBroadPhaseStaticFilterGroup(); }
uint32_t pxref_dynamic_group(uint32_t id) { return PxGetBroadPhaseDynamicFilterGroup(id); }

// One pair through PxGenerateContacts with a fresh PxCache. geom: sphere (r),
// capsule (r, halfHeight), box (hx hy hz), plane (unused). pose: p xyz + q xyzw.
// out: per contact (point xyz, separation, normal xyz, 0). Returns the count.
int pxref_generate_contacts(int type0, const float* geom0, const float* pose0, int type1, const float* geom1,
                            const float* pose1, float contactDistance, float toleranceLength, float* out, int maxContacts)
{
	if (!ensureSdk())
		return -1;
	PxGeometryHolder g0 = makeGeom(type0, geom0), g1 = makeGeom(type1, geom1);
	const PxGeometry* gp0 = &g0.any();
	const PxGeometry* gp1 = &g1.any();
	PxTransform p0 = makePose(pose0), p1 = makePose(pose1);
	PxCache cache;
	Recorder rec;
	rec.out = out;
	rec.maxContacts = maxContacts;
	CacheAlloc alloc;
	if (!PxGenerateContacts(&gp0, &gp1, &p0, &p1, &cache, 1, rec, contactDistance, 0.01f, toleranceLength, alloc))
		return -1;
	return rec.count;
}

// Scene: gravity xyz; solverType 0 PGS / 1 TGS; bpType 0 ABP / 1 PABP.
// reportContacts != 0: the scene reports every touching pair's contact points
// each step (pxref_scene_contacts); the simulation itself is unchanged
void* pxref_scene_create2(const float* gravity, int solverType, int bpType, int nThreads, int reportContacts)
{
	if (!ensureSdk())
		return nullptr;
	RefScene* rs = new RefScene();
	PxSceneDesc sd(gPhysics->getTolerancesScale());
	sd.gravity = PxVec3(gravity[0], gravity[1], gravity[2]);
	rs->dispatcher = PxDefaultCpuDispatcherCreate(PxU32(nThreads));
	sd.cpuDispatcher = rs->dispatcher;
	sd.filterShader = PxDefaultSimulationFilterShader;
	if (reportContacts)
	{
		rs->contacts = new ContactCollector();
		sd.filterShader = reportingFilterShader;
		sd.simulationEventCallback = rs->contacts;
	}
	sd.solverType = solverType ? PxSolverType::eTGS : PxSolverType::ePGS;
	sd.broadPhaseType = bpType ? PxBroadPhaseType::ePABP : PxBroadPhaseType::eABP;
	rs->scene = gPhysics->createScene(sd);
	if (!rs->scene)
	{
		delete rs;
		return nullptr;
	}
	rs->material = gPhysics->createMaterial(0.5f, 0.5f, 0.0f);
	return rs;
}

void* pxref_scene_create(const float* gravity, int solverType, int bpType, int nThreads)
{
	return pxref_scene_create2(gravity, solverType, bpType, nThreads, 0);
}

// Contact rows of the last step (10 floats each, see ContactCollector);
// returns the row count (rows beyond max are not copied).
int pxref_scene_contacts(void* h, float* out, int max)
{
	RefScene* rs = static_cast<RefScene*>(h);
	if (!rs->contacts)
		return -1;
	const int k = int(rs->contacts->rows.size() / 10);
	std::memcpy(out, rs->contacts->rows.data(), sizeof(float) * 10 * size_t(k < max ? k : max));
	return k;
}

int pxref_scene_set_material(void* h, float staticFriction, float dynamicFriction, float restitution)
{
	RefScene* rs = static_cast<RefScene*>(h);
	rs->material->setStaticFriction(staticFriction);
	rs->material->setDynamicFriction(dynamicFriction);
	rs->material->setRestitution(restitution);
	return 0;
}

// Adds n dynamic bodies, one exclusive shape each. type: 0 sphere / 3 box;
// dims: n x 4 (r or hx hy hz); pose: n x 7; linVel, angVel: n x 3; mass: n;
// inertia: n x 3 (mass-space diagonal). Sleeping is disabled (threshold 0).
int pxref_scene_add_bodies(void* h, int n, const int* type, const float* dims, const float* pose, const float* linVel,
                           const float* angVel, const float* mass, const float* inertia, int posIters, int velIters,
                           float contactOffset, float restOffset, float linearDamping, float angularDamping)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (int i = 0; i < n; i++)
	{
		PxRigidDynamic* body = gPhysics->createRigidDynamic(makePose(pose + 7 * i));
		PxGeometryHolder g = makeGeom(type[i], dims + 4 * i);
		PxShape* shape = PxRigidActorExt::createExclusiveShape(*body, g.any(), *rs->material);
		shape->setContactOffset(contactOffset);
		shape->setRestOffset(restOffset);
		body->setMass(mass[i]);
		body->setMassSpaceInertiaTensor(PxVec3(inertia[3 * i], inertia[3 * i + 1], inertia[3 * i + 2]));
		body->setLinearVelocity(PxVec3(linVel[3 * i], linVel[3 * i + 1], linVel[3 * i + 2]));
		body->setAngularVelocity(PxVec3(angVel[3 * i], angVel[3 * i + 1], angVel[3 * i + 2]));
		body->setSolverIterationCounts(PxU32(posIters), PxU32(velIters));
		body->setLinearDamping(linearDamping);
		body->setAngularDamping(angularDamping);
		body->setSleepThreshold(0.0f);
		body->setStabilizationThreshold(0.0f);
		body->userData = reinterpret_cast<void*>(intptr_t(rs->bodies.size() + 1));
		rs->scene->addActor(*body);
		rs->bodies.push_back(body);
	}
	return int(rs->bodies.size());
}

// Adds n static planes n.x + d = 0 (plane: n x 4 = nx ny nz d).
int pxref_scene_add_planes(void* h, int n, const float* plane, float contactOffset)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (int i = 0; i < n; i++)
	{
		PxRigidStatic* s = PxCreatePlane(*gPhysics, PxPlane(plane[4 * i], plane[4 * i + 1], plane[4 * i + 2], plane[4 * i + 3]),
		                                 *rs->material);
		if (!s)
			return -1;
		PxShape* shape = nullptr;
		s->getShapes(&shape, 1);
		shape->setContactOffset(contactOffset);
		shape->setRestOffset(0.0f);
		s->userData = reinterpret_cast<void*>(intptr_t(-(++rs->nPlanes)));
		rs->scene->addActor(*s);
	}
	return 0;
}

// Steps nsteps times; writes wall ms per step (simulate + fetchResults) into
// stepMs (may be null). Returns total wall seconds.
double pxref_scene_step(void* h, float dt, int nsteps, double* stepMs)
{
	RefScene* rs = static_cast<RefScene*>(h);
	auto t0 = std::chrono::steady_clock::now();
	for (int s = 0; s < nsteps; s++)
	{
		if (rs->contacts)
			rs->contacts->rows.clear();
		auto a = std::chrono::steady_clock::now();
		rs->scene->simulate(dt);
		rs->scene->fetchResults(true);
		auto b = std::chrono::steady_clock::now();
		if (stepMs)
			stepMs[s] = std::chrono::duration<double, std::milli>(b - a).count();
	}
	return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// PxRigidDynamic::addForce / addTorque (PxForceMode::eFORCE) on every body
// for the next step; force / torque: n x 3 (either may be null)
void pxref_scene_add_forces(void* h, const float* force, const float* torque)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (size_t i = 0; i < rs->bodies.size(); i++)
	{
		if (force)
			rs->bodies[i]->addForce(PxVec3(force[3 * i], force[3 * i + 1], force[3 * i + 2]), PxForceMode::eFORCE);
		if (torque)
			rs->bodies[i]->addTorque(PxVec3(torque[3 * i], torque[3 * i + 1], torque[3 * i + 2]), PxForceMode::eFORCE);
	}
}

// An aggregate of k bodies already added (PxPhysics::createAggregate with the
// eGENERIC filter hint, with or without self-collisions; the bodies leave the
// scene, join the aggregate, and the aggregate is added). Returns 0 or -1.
int pxref_scene_add_aggregate(void* h, int k, const int* bodies, int selfCollisions)
{
	RefScene* rs = static_cast<RefScene*>(h);
	PxAggregate* agg = gPhysics->createAggregate(PxU32(k), PxU32(k),
	                                             PxGetAggregateFilterHint(PxAggregateType::eGENERIC, selfCollisions != 0));
	if (!agg)
		return -1;
	for (int i = 0; i < k; i++)
	{
		PxRigidDynamic* b = rs->bodies[size_t(bodies[i])];
		rs->scene->removeActor(*b);
		if (!agg->addActor(*b))
			return -1;
	}
	return rs->scene->addAggregate(*agg) ? 0 : -1;
}

// Adds k joints without limits or drives: type 0 fixed, 1 spherical,
// 2 revolute, 3 prismatic, 4 distance (PxFixedJointCreate,
// PxSphericalJointCreate, PxRevoluteJointCreate, PxPrismaticJointCreate,
// PxDistanceJointCreate with its defaults); body0/body1: indices of
// bodies added by pxref_scene_add_bodies or -1 for the world; frame0/frame1:
// k x 7 local frames (px py pz qx qy qz qw). Collision between the jointed
// bodies stays off (the joints' default). Returns the joint count or -1.
int pxref_scene_add_joints(void* h, int k, const int* type, const int* body0, const int* body1, const float* frame0,
                           const float* frame1)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (int i = 0; i < k; i++)
	{
		PxRigidActor* a0 = body0[i] >= 0 ? rs->bodies[size_t(body0[i])] : nullptr;
		PxRigidActor* a1 = body1[i] >= 0 ? rs->bodies[size_t(body1[i])] : nullptr;
		const PxTransform f0 = makePose(frame0 + 7 * i), f1 = makePose(frame1 + 7 * i);
		PxJoint* j = nullptr;
		switch (type[i])
		{
		case 0: j = PxFixedJointCreate(*gPhysics, a0, f0, a1, f1); break;
		case 1: j = PxSphericalJointCreate(*gPhysics, a0, f0, a1, f1); break;
		case 2: j = PxRevoluteJointCreate(*gPhysics, a0, f0, a1, f1); break;
		case 3: j = PxPrismaticJointCreate(*gPhysics, a0, f0, a1, f1); break;
		case 4: j = PxDistanceJointCreate(*gPhysics, a0, f0, a1, f1); break;
		default: return -1;
		}
		if (!j)
			return -1;
		rs->jointList.push_back(j);
		rs->joints++;
	}
	return rs->joints;
}

// Limit and drive parameters of k joints (indices in creation order):
// params[10*i] = flags (1 limit, 2 drive, 4 drive freespin), lower, upper,
// restitution, bounce threshold, drive velocity, drive force limit, drive
// gear ratio, limit stiffness, limit damping (a soft limit when either is
// positive). Revolute: PxRevoluteJoint::setLimit(PxJointAngularLimitPair),
// setRevoluteJointFlag(eLIMIT_ENABLED / eDRIVE_ENABLED / eDRIVE_FREESPIN),
// setDriveVelocity, setDriveForceLimit, setDriveGearRatio; prismatic:
// PxPrismaticJoint::setLimit(PxJointLinearLimitPair), eLIMIT_ENABLED;
// spherical: PxSphericalJoint::setLimitCone(PxJointLimitCone(lower, upper)),
// eLIMIT_ENABLED.
// Returns 0 or -1.
int pxref_scene_set_joint_params(void* h, int k, const int* joints, const float* params)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (int i = 0; i < k; i++)
	{
		if (joints[i] < 0 || size_t(joints[i]) >= rs->jointList.size())
			return -1;
		PxJoint* j = rs->jointList[size_t(joints[i])];
		const float* q = params + 10 * i;
		const int f = int(q[0]);
		j->setConstraintFlag(PxConstraintFlag::eCOLLISION_ENABLED, (f & 8) != 0);
		j->setConstraintFlag(PxConstraintFlag::eDRIVE_LIMITS_ARE_FORCES, (f & 16) != 0);
		if (PxRevoluteJoint* r = j->is<PxRevoluteJoint>())
		{
			PxJointAngularLimitPair lim(q[1], q[2]);
			lim.restitution = q[3];
			lim.bounceThreshold = q[4];
			lim.stiffness = q[8];
			lim.damping = q[9];
			r->setLimit(lim);
			r->setRevoluteJointFlag(PxRevoluteJointFlag::eLIMIT_ENABLED, (f & 1) != 0);
			r->setDriveVelocity(q[5], false);
			r->setDriveForceLimit(q[6]);
			r->setDriveGearRatio(q[7]);
			r->setRevoluteJointFlag(PxRevoluteJointFlag::eDRIVE_ENABLED, (f & 2) != 0);
			r->setRevoluteJointFlag(PxRevoluteJointFlag::eDRIVE_FREESPIN, (f & 4) != 0);
		}
		else if (PxPrismaticJoint* p = j->is<PxPrismaticJoint>())
		{
			PxJointLinearLimitPair lim(gPhysics->getTolerancesScale(), q[1], q[2]);
			lim.restitution = q[3];
			lim.bounceThreshold = q[4];
			lim.stiffness = q[8];
			lim.damping = q[9];
			p->setLimit(lim);
			p->setPrismaticJointFlag(PxPrismaticJointFlag::eLIMIT_ENABLED, (f & 1) != 0);
		}
		else if (PxSphericalJoint* sj = j->is<PxSphericalJoint>())
		{
			// lower / upper: the cone's y and z angles
			PxJointLimitCone lim(q[1], q[2]);
			lim.restitution = q[3];
			lim.bounceThreshold = q[4];
			lim.stiffness = q[8];
			lim.damping = q[9];
			sj->setLimitCone(lim);
			sj->setSphericalJointFlag(PxSphericalJointFlag::eLIMIT_ENABLED, (f & 1) != 0);
		}
		else if (PxDistanceJoint* dj = j->is<PxDistanceJoint>())
		{
			// flags 1 max, 2 min distance, 4 spring; lower / upper the min /
			// max distance; the restitution slot the tolerance
			dj->setMinDistance(q[1]);
			dj->setMaxDistance(q[2]);
			dj->setTolerance(q[3]);
			dj->setStiffness(q[8]);
			dj->setDamping(q[9]);
			dj->setDistanceJointFlag(PxDistanceJointFlag::eMAX_DISTANCE_ENABLED, (f & 1) != 0);
			dj->setDistanceJointFlag(PxDistanceJointFlag::eMIN_DISTANCE_ENABLED, (f & 2) != 0);
			dj->setDistanceJointFlag(PxDistanceJointFlag::eSPRING_ENABLED, (f & 4) != 0);
		}
		else if (f & 23)
			return -1;
	}
	return 0;
}

// pose: n x 7, linVel/angVel: n x 3 (any may be null)
void pxref_scene_get_bodies(void* h, float* pose, float* linVel, float* angVel)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (size_t i = 0; i < rs->bodies.size(); i++)
	{
		PxRigidDynamic* b = rs->bodies[i];
		if (pose)
		{
			PxTransform t = b->getGlobalPose();
			float* o = pose + 7 * i;
			o[0] = t.p.x; o[1] = t.p.y; o[2] = t.p.z; o[3] = t.q.x; o[4] = t.q.y; o[5] = t.q.z; o[6] = t.q.w;
		}
		if (linVel)
		{
			PxVec3 v = b->getLinearVelocity();
			linVel[3 * i] = v.x; linVel[3 * i + 1] = v.y; linVel[3 * i + 2] = v.z;
		}
		if (angVel)
		{
			PxVec3 v = b->getAngularVelocity();
			angVel[3 * i] = v.x; angVel[3 * i + 1] = v.y; angVel[3 * i + 2] = v.z;
		}
	}
}

// out[0] discrete contact pairs (all geometry pairs), out[1] BP pairs created
// in the last step, out[2] active dynamic bodies, out[3] nb patches
void pxref_scene_stats(void* h, long long* out)
{
	RefScene* rs = static_cast<RefScene*>(h);
	PxSimulationStatistics st;
	rs->scene->getSimulationStatistics(st);
	out[0] = st.nbDiscreteContactPairsTotal;
	out[1] = st.getNbBroadPhaseAdds();
	out[2] = st.nbActiveDynamicBodies;
	out[3] = st.nbDiscreteContactPairsWithContacts;
}

// Turns sleeping on for every body added so far: PxRigidDynamic::
// setSleepThreshold(threshold) (threshold 0 keeps them awake, the default of
// pxref_scene_add_bodies). Returns 0.
int pxref_scene_set_sleep_threshold(void* h, float threshold)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (PxRigidDynamic* b : rs->bodies)
		b->setSleepThreshold(threshold);
	return 0;
}

// Per body: PxRigidDynamic::getWakeCounter() into wc and isSleeping() into
// asleep (either may be null).
void pxref_scene_sleep_state(void* h, float* wc, int* asleep)
{
	RefScene* rs = static_cast<RefScene*>(h);
	for (size_t i = 0; i < rs->bodies.size(); i++)
	{
		if (wc)
			wc[i] = rs->bodies[i]->getWakeCounter();
		if (asleep)
			asleep[i] = rs->bodies[i]->isSleeping() ? 1 : 0;
	}
}

// Replaces the scene's CPU dispatcher by one of nThreads workers between
// steps (PxTaskManager::setCpuDispatcher, physx/include/task/PxTaskManager.h:88);
// every stage reads the worker count from the task manager when it runs
// (e.g. ScPipeline.cpp:350, DyTGSDynamics.cpp:2409), so the next step uses
// nThreads. Returns 0, or -1 for a bad argument.
int pxref_scene_set_threads(void* h, int nThreads)
{
	RefScene* rs = static_cast<RefScene*>(h);
	if (!rs || !rs->scene || nThreads < 0)
		return -1;
	PxDefaultCpuDispatcher* d = PxDefaultCpuDispatcherCreate(PxU32(nThreads));
	rs->scene->getTaskManager()->setCpuDispatcher(*d);
	rs->retired.push_back(rs->dispatcher);
	rs->dispatcher = d;
	return 0;
}

void pxref_scene_release(void* h)
{
	RefScene* rs = static_cast<RefScene*>(h);
	if (!rs)
		return;
	if (rs->scene)
		rs->scene->release();
	if (rs->dispatcher)
		rs->dispatcher->release();
	for (PxDefaultCpuDispatcher* d : rs->retired)
		d->release();
	if (rs->material)
		rs->material->release();
	delete rs->contacts;
	delete rs;
}

}  // extern "C"
