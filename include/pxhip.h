/* pxhip — MI355X (gfx950) GPU rigid-body step behind the PhysX GPU plug-in
 * boundary. C ABI: plain pointers and sizes, no C++ or framework types.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the PhysX 5.6.1 tree, physx/). Device pointers are HIP device addresses;
 * `stream` is a hipStream_t (NULL = the scene's own stream).
 *
 * Return codes: 0 on success, negative on error (pxh_last_error() describes
 * the last error of the calling thread).
 */
#ifndef PXHIP_H
#define PXHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PXH_VERSION 0x00010000

enum {
	PXH_OK = 0,
	PXH_ERR_INVALID = -1,        /* bad argument (PxErrorCode::eINVALID_PARAMETER) */
	PXH_ERR_DEVICE = -2,         /* HIP runtime error (PxCudaContext::getLastError) */
	PXH_ERR_STATE = -3,          /* call not allowed now, e.g. during simulate (NpDirectGPUAPI.cpp:65-72) */
	PXH_ERR_CAPACITY = -4,       /* buffer capacity exceeded (PxGpuDynamicsMemoryConfig, PxSceneDesc.h:457-493) */
	PXH_ERR_UNSUPPORTED = -5     /* feature outside the implemented path */
};

/* geometry codes = PxGeometryType::Enum (physx/include/geometry/PxGeometry.h) */
enum { PXH_GEOM_SPHERE = 0, PXH_GEOM_PLANE = 1, PXH_GEOM_CAPSULE = 2, PXH_GEOM_BOX = 3, PXH_GEOM_CONVEXMESH = 5 };

/* solver codes = PxSolverType::Enum (physx/include/PxSceneDesc.h) */
enum { PXH_SOLVER_PGS = 0, PXH_SOLVER_TGS = 1 };

/* Direct GPU API data kinds = PxRigidDynamicGPUAPIReadType/WriteType
 * (physx/include/PxDirectGPUAPI.h:45-71): pose is a PxTransform (q xyzw,
 * p xyz), the others one PxVec3 per body. Reads 3/4 are the accelerations of
 * the last step (velocity change over its dt; need enable_body_accelerations),
 * writes 3/4 a force / torque at the centre of mass for the next step
 * (PxForceMode::eFORCE, cleared after that step). */
enum { PXH_DATA_GLOBAL_POSE = 0, PXH_DATA_LINEAR_VELOCITY = 1, PXH_DATA_ANGULAR_VELOCITY = 2 };
enum { PXH_READ_LINEAR_ACCELERATION = 3, PXH_READ_ANGULAR_ACCELERATION = 4 };
enum { PXH_WRITE_FORCE = 3, PXH_WRITE_TORQUE = 4 };

/* Scene description: the PxSceneDesc / PxRigidDynamic / PxShape / PxMaterial
 * parameters the step reads (PxSceneDesc.h:1035-1075 defaults,
 * PxvDynamics.h:119-160 body defaults). */
typedef struct pxh_scene_desc {
	float gravity[3];
	int32_t solver_type;                 /* PXH_SOLVER_TGS */
	int32_t position_iterations;         /* PxRigidDynamic::setSolverIterationCounts */
	int32_t velocity_iterations;
	float contact_offset;                /* per shape (PxShape::setContactOffset) */
	float rest_offset;                   /* per shape (PxShape::setRestOffset) */
	float static_friction, dynamic_friction, restitution;  /* the scene's one material */
	float linear_damping, angular_damping;
	float max_linear_velocity_sq, max_angular_velocity_sq;
	float bounce_threshold_velocity;     /* PxSceneDesc::bounceThresholdVelocity */
	float friction_offset_threshold;     /* PxSceneDesc::frictionOffsetThreshold */
	float friction_correlation_distance; /* PxSceneDesc::frictionCorrelationDistance */
	float tolerance_length;              /* PxTolerancesScale::length */
	int32_t max_pairs_per_body;          /* initial pair buffer per body (default 4, >= 1); grows on demand */
	int32_t enable_body_accelerations;   /* PxSceneFlag::eENABLE_BODY_ACCELERATIONS: keep the velocities
	                                        before each step for the acceleration reads */
	int32_t report_touch_events;         /* PxPairFlag::eNOTIFY_TOUCH_FOUND | eNOTIFY_TOUCH_LOST on every
	                                        pair: compute the touch events of each step */
	int32_t enable_sleeping;             /* bodies sleep (PxRigidDynamic::setSleepThreshold with a positive
	                                        threshold; 0 = never, as with threshold 0) */
	float sleep_threshold;               /* PxRigidDynamic::setSleepThreshold of every body (default
	                                        5e-5 * speed^2 = 0.005 for PxTolerancesScale speed 10) */
} pxh_scene_desc;

typedef struct pxh_scene_t* pxh_scene;
typedef struct pxh_broadphase_t* pxh_broadphase;

int32_t pxh_get_version(void);
const char* pxh_last_error(void);
void pxh_default_scene_desc(pxh_scene_desc* desc);

/* Scene lifecycle. Replaces the GPU pipeline objects the plug-in factory
 * creates for a scene: PxPhysXGpu::createGpuBroadPhase / createGpuAABBManager /
 * createGpuNphaseImplementationContext / createGpuSimulationController /
 * createGpuDynamicsContext (physxgpu/include/PxPhysXGpu.h:101-200), constructed
 * by Sc::Scene::Scene (simulationcontroller/src/ScScene.cpp:797-968). */
int32_t pxh_scene_create(const pxh_scene_desc* desc, pxh_scene* out_scene);
int32_t pxh_scene_release(pxh_scene scene);

/* Adds n dynamic rigid bodies with one shape each (PxPhysics::createRigidDynamic
 * + PxRigidActorExt::createExclusiveShape + PxScene::addActor). Host arrays:
 * type[n] (sphere, capsule, box or convex mesh), dims[n*4] (sphere: radius;
 * capsule: radius, half height along the local x axis (PxCapsuleGeometry);
 * box: half extents; convex mesh: index from pxh_scene_add_convex_meshes), pose[n*7] (p xyz,
 * q xyzw), lin_vel[n*3], ang_vel[n*3], mass[n], inertia[n*3] (mass-space
 * diagonal), world[n] (independent-world id, non-decreasing; NULL = world 0).
 * Returns the index of the first added body (its PxRigidDynamicGPUIndex). */
int32_t pxh_scene_add_bodies(pxh_scene scene, int32_t n, const int32_t* type, const float* dims, const float* pose,
                             const float* lin_vel, const float* ang_vel, const float* mass, const float* inertia,
                             const int32_t* world);

/* Adds m static planes n.x + d = 0 (plane[m*4] = nx ny nz d; PxCreatePlane,
 * physx/include/extensions/PxSimpleFactory.h:272), world[m] or NULL. */
int32_t pxh_scene_add_planes(pxh_scene scene, int32_t m, const float* plane, const int32_t* world);

/* One step of dt seconds: bounds, broadphase, narrowphase, partitioning,
 * contact prep, TGS solve and integration, enqueued on the scene's stream.
 * Replaces PxScene::simulate (physx/source/physx/src/NpScene.cpp:2907) and
 * PxScene::fetchResults (NpSceneFetchResults.cpp:230) for the GPU pipeline. */
int32_t pxh_scene_simulate(pxh_scene scene, float dt);
int32_t pxh_scene_fetch_results(pxh_scene scene, int32_t block);

/* Aggregates: PxPhysics::createAggregate + PxAggregate::addActor +
 * PxScene::addAggregate (physx/include/PxAggregate.h; filter hint of
 * PxAggregateType::eGENERIC with or without self-collisions). The k bodies
 * (host array of indices, none already in an aggregate) form one aggregate;
 * without self-collisions its members never pair with each other in the
 * broadphase, while pairs with every other body are found as before (the
 * reference's aggregate bounds are a broadphase acceleration: the pair set
 * is the same). Returns the aggregate's id. */
int32_t pxh_scene_add_aggregate(pxh_scene scene, int32_t k, const int32_t* bodies, int32_t self_collisions);

/* Joints (no break forces; limits and drives through
 * pxh_scene_set_joint_params): PxFixedJointCreate,
 * PxSphericalJointCreate, PxRevoluteJointCreate (free twist about the joint
 * frame's x axis), PxPrismaticJointCreate (free translation along x)
 * (physx/include/extensions/PxFixedJoint.h, PxSphericalJoint.h,
 * PxRevoluteJoint.h, PxPrismaticJoint.h), PxDistanceJointCreate
 * (PxDistanceJoint.h), solved as 1D constraint rows
 * (the GPU reference's constraintBlockPrePrep.cu:1168,1219 and
 * solve1DBlockTGS, gpusolver/src/CUDA/solverBlockTGS.cuh:744). Host arrays:
 * type[k], body0[k] / body1[k] (body indices, -1 = the world frame, never
 * both; bodies of one world), frame0[k*7] / frame1[k*7] local joint frames
 * (p xyz, q xyzw) relative to each body (for the world: world coordinates).
 * Jointed bodies do not collide (PxConstraintFlag::eCOLLISION_ENABLED off,
 * the joints' default). TGS only (a PGS step with joints returns
 * PXH_ERR_UNSUPPORTED). Returns the index of the first joint. */
enum { PXH_JOINT_FIXED = 0, PXH_JOINT_SPHERICAL = 1, PXH_JOINT_REVOLUTE = 2, PXH_JOINT_PRISMATIC = 3,
       PXH_JOINT_DISTANCE = 4 };
int32_t pxh_scene_add_joints(pxh_scene scene, int32_t k, const int32_t* type, const int32_t* body0,
                             const int32_t* body1, const float* frame0, const float* frame1);

/* Limits and drives of revolute, prismatic and spherical joints (joint indices as
 * pxh_scene_add_joints returned them), params[10*i]: flags (1 limit, 2 drive,
 * 4 drive freespin, 8 the jointed bodies collide: PxConstraintFlag::
 * eCOLLISION_ENABLED, any joint type; 16 the drive force limit is a force,
 * PxConstraintFlag::eDRIVE_LIMITS_ARE_FORCES), lower, upper (radians about / metres along the joint
 * frame's x axis; lower >= upper locks the axis), restitution, bounce
 * threshold, drive velocity, drive force limit (an impulse per step; may be
 * +inf), drive gear ratio, limit stiffness, limit damping (either positive:
 * a soft limit, a spring acting only beyond the stops); for a spherical
 * joint lower / upper are the cone's y and z angles
 * (PxSphericalJoint::setLimitCone(PxJointLimitCone), eLIMIT_ENABLED,
 * PxSphericalJoint.h:97-120; SphericalJointSolverPrep,
 * ExtSphericalJoint.cpp:157-190); for a distance joint flags are 1 max
 * distance, 2 min distance, 4 spring, lower / upper the min / max distance
 * and the restitution slot the tolerance (PxDistanceJoint::setMinDistance /
 * setMaxDistance / setTolerance / setStiffness / setDamping /
 * setDistanceJointFlag, PxDistanceJoint.h:80-170; DistanceJointSolverPrep,
 * ExtDistanceJoint.cpp:179-271; a new distance joint keeps a max distance of
 * 0 with tolerance 0.025, PxDistanceJointCreate's defaults). Replaces
 * PxRevoluteJoint::setLimit (PxJointAngularLimitPair), setRevoluteJointFlag(eLIMIT_ENABLED /
 * eDRIVE_ENABLED / eDRIVE_FREESPIN), setDriveVelocity, setDriveForceLimit,
 * setDriveGearRatio (physx/include/extensions/PxRevoluteJoint.h:109-230)
 * and PxPrismaticJoint::setLimit (PxJointLinearLimitPair),
 * setPrismaticJointFlag(eLIMIT_ENABLED) (PxPrismaticJoint.h:97-140). Rows: RevoluteJointSolverPrep
 * (physxextensions/src/ExtRevoluteJoint.cpp:195-258), PrismaticJointSolverPrep
 * (ExtPrismaticJoint.cpp:119-153). */
int32_t pxh_scene_set_joint_params(pxh_scene scene, int32_t k, const int32_t* joints, const float* params);

/* PxDirectGPUAPI::getRigidDynamicData / setRigidDynamicData
 * (physx/include/PxDirectGPUAPI.h:311,325; NpDirectGPUAPI.cpp:63-92).
 * data: device buffer of n PxTransform (7 floats) or PxVec3 (3 floats);
 * index: device array of n body indices (NULL = bodies 0..n-1); the copy is
 * ordered after the scene's work and, when stream is given, before any work
 * enqueued on that stream afterwards. */
int32_t pxh_scene_get_rigid_dynamic_data(pxh_scene scene, void* data, const uint32_t* index, int32_t kind, int32_t n,
                                         void* stream);
int32_t pxh_scene_set_rigid_dynamic_data(pxh_scene scene, const void* data, const uint32_t* index, int32_t kind,
                                         int32_t n, void* stream);

/* Static triangle mesh (PxPhysics::createTriangleMesh + a static actor with
 * PxTriangleMeshGeometry, identity scale; physx/include/geometry/
 * PxTriangleMeshGeometry.h) in world `world`, at most one per world. Host
 * arrays: verts[nv*3] (mesh frame), tris[nt*3] vertex indices wound
 * counter-clockwise about the outward normal (the mesh is one-sided:
 * bodies whose centre is behind a triangle do not touch it), pose[7] (p xyz,
 * q xyzw). Degenerate triangles are dropped. Bodies meet the mesh through a
 * bounding-volume tree on the device (the reference's midphase, BV4 in
 * geomutils/src/mesh/GuBV4*.cpp) and a contact generation restated after
 * the reference's PCM mesh contacts (geomutils/src/pcm/
 * GuPCMContactSphereMesh.cpp, GuPCMTriangleContactGen.cpp): up to 4 contact
 * patches per body, one constraint each; in contact data and constraint
 * tables the static side of patch k reads -1 - (plane count + k). A step
 * returns PXH_ERR_CAPACITY when more than 128 triangles meet one body's
 * bounds. Returns the mesh index. */
int32_t pxh_scene_add_triangle_mesh(pxh_scene scene, int32_t nv, const float* verts, int32_t nt, const int32_t* tris,
                                    const float* pose, int32_t world);

/* Static heightfield (PxCreateHeightField + a static actor with
 * PxHeightFieldGeometry(hf, no flags, height_scale, row_scale, column_scale);
 * physx/include/geometry/PxHeightFieldGeometry.h, PxHeightFieldDesc.h) in
 * world `world`, added as the triangle mesh of its cells: rows x cols int16
 * samples (host, row-major), tess[rows*cols] (host, may be NULL) each
 * sample's tessellation flag (the cell's diagonal runs from its zeroth vertex,
 * PxHeightFieldSample::tessFlag), vertex (row * row_scale, height *
 * height_scale, col * column_scale) in the heightfield's frame at pose[7].
 * Counts as the world's one mesh. Returns the mesh index. */
int32_t pxh_scene_add_heightfield(pxh_scene scene, int32_t rows, int32_t cols, const int16_t* heights,
                                  const uint8_t* tess, float height_scale, float row_scale, float column_scale,
                                  const float* pose, int32_t world);

/* Sleeping (enable_sleeping): bodies and islands fall asleep and wake as in
 * the reference's CPU pipeline. Each body keeps a wake counter (0.4 s when
 * added or woken) that counts down while the body's mean motion over the
 * step, accumulated in the last 0.2 s, stays below (1 + interactions) times
 * the sleep threshold, and restarts when it rises above
 * (Dy::sleepCheck, lowleveldynamics/src/DySleep.cpp:173-235); an island of
 * bodies joined by touching contacts and joints whose wake counters all
 * reached zero by the end of a step sleeps from the next step on: zero
 * velocities, no integration, no contacts among its bodies (the island
 * manager's deactivation, lowlevel/software/src/PxsIslandSim.cpp:1630-1720,
 * ScBodySim.cpp:469-490). An awake body that touches a sleeping one wakes the
 * sleeping body's island (ScPipeline.cpp:1726-1776).
 * pxh_scene_get_sleep_state: PxRigidDynamic::getWakeCounter / isSleeping
 * (physx/include/PxRigidDynamic.h) of every body into host arrays
 * wake_counter[n] / asleep[n] (either may be NULL); n at most the body count.
 * pxh_scene_wake_bodies: PxRigidDynamic::wakeUp of k bodies (host array of
 * indices): wake counter 0.4 s; the rest of a sleeping body's island wakes
 * at the next step. Writing poses or velocities through
 * pxh_scene_set_rigid_dynamic_data wakes the written bodies the same way,
 * forces and torques when nonzero (autowake). Sleeping does not combine with
 * report_touch_events (PXH_ERR_UNSUPPORTED). */
int32_t pxh_scene_get_sleep_state(pxh_scene scene, float* wake_counter, int32_t* asleep, int32_t n);
int32_t pxh_scene_wake_bodies(pxh_scene scene, const int32_t* bodies, int32_t k);

/* PxScene::getSimulationStatistics (physx/include/PxSimulationStatistics.h):
 * out[0] bodies, [1] broadphase pairs (body-body), [2] body-plane pairs,
 * [3] touching pairs (constraints), [4] contact points, [5] colours,
 * [6] steps taken, [7] friction rows, [8] body-body constraints, [9] colouring rounds,
 * [10] friction rows of body-body constraints, [11] device bytes the scene
 * holds (PxSimulationStatistics::gpuMemHeap, PxSimulationStatistics.h:333),
 * [12] solve launches of the last step.
 * Returns the number of values written (<= n). */
int32_t pxh_scene_get_stats(pxh_scene scene, int64_t* out, int32_t n);

/* Touch events of the last step: pairs that touch (have contacts) now and did
 * not in the step before (found), and pairs that touched then and do not now
 * (lost). What PxSimulationEventCallback::onContact reports for
 * PxPairFlag::eNOTIFY_TOUCH_FOUND / eNOTIFY_TOUCH_LOST (physx/include/
 * PxSimulationEventCallback.h), found on the device like the GPU reference's
 * prepareLostFoundPairs (gpunarrowphase/src/CUDA/cudaGJKEPA.cu:1468,1532; CPU
 * PxsNphaseImplementationContext::appendNewLostPairs,
 * lowlevel/software/src/PxsNphaseImplementationContext.cpp:928). Needs
 * report_touch_events. Host arrays of (body a, body b or -1 - plane index)
 * pairs in the current body numbering, at most max_found / max_lost of each;
 * counts[0..1] receive the full counts. Lost pairs of worlds removed since are
 * not reported. The first step reports every touching pair as found. */
int32_t pxh_scene_get_touch_events(pxh_scene scene, int32_t* found, int64_t max_found, int32_t* lost,
                                   int64_t max_lost, int64_t* counts);

/* Last step's constraint table, host copy for inspection:
 * per constraint (body a, body b or -1-plane, colour, contact count). */
int64_t pxh_scene_get_constraints(pxh_scene scene, int32_t* out, int64_t max_constraints);

/* Per-stage GPU time from HIP events recorded on the scene's stream around
 * each launch, the device-side counterpart of the reference's profiler zones
 * (PX_PROFILE_ZONE, physx/include/common/PxProfileZone.h:34-50). Stages:
 * 0 bounds, 1 broadphase (untimed), 2 body-pair narrowphase, 3 body-plane
 * narrowphase, 4 contact prep, 5 solve (one colour per launch), 6 integration.
 * stage_mask: bit i times stage i (0 = off, PXH_PROFILE_ALL = every stage);
 * solve launches carry their events in the dispatch itself, so timing the
 * solve alone adds no marker packets to the stream.
 * ms[i] accumulates milliseconds and launches[i] the timed launches since
 * profiling was (re)set. */
#define PXH_PROFILE_ALL 0xff
int32_t pxh_scene_set_profiling(pxh_scene scene, int32_t stage_mask);
/* Sampling of the stages whose launches carry their own events (the solve):
 * launch j of the k-th step since profiling was set is timed when
 * (j + k) % stride == 0 (1 = every launch), so over `stride` consecutive steps
 * each launch position of a step is timed once, and a caller can keep
 * per-launch timing on while it measures throughput (the events of each timed
 * launch cost host work at submission).
 * Profiler-side statistics, no reference counterpart (PxProfilerCallback
 * zones are the reference's instrumentation). */
int32_t pxh_scene_set_profiling_stride(pxh_scene scene, int32_t stride);
int32_t pxh_scene_get_profile(pxh_scene scene, double* ms, int64_t* launches, int32_t n);

/* PxDirectGPUAPI::copyContactData (physx/include/PxDirectGPUAPI.h:401): the
 * last step's touching pairs into a device buffer, one 128-byte record per
 * pair (in place of PxGpuContactPair, physx/include/PxContact.h:818, whose
 * pointers lead into internal buffers): int32 body0, body1 (-1 - plane index
 * for a plane), contact count, 0 | normal xyz (from shape1 towards shape0),
 * 0 | 4 x (point xyz, separation) | 4 applied normal impulses | 0.
 * nb_pairs (device, may be NULL) receives the number of records written. */
int32_t pxh_scene_copy_contact_data(pxh_scene scene, void* data, uint32_t* nb_pairs, int32_t max_pairs, void* stream);

/* Launch shapes of the step. No counterpart in the reference, whose GPU
 * kernels have fixed shapes; results are the same bit for bit under every
 * setting. tuning[0] the TGS solve layout: PXH_SOLVE_LAYOUT_AUTO (the
 * per-world solve when every world has at most PXH_WORLD_SOLVE_MAX_BODIES
 * bodies, else one launch per colour and iteration), _COLORS or _WORLDS (one
 * persistent launch of every iteration of every world, each world in one
 * workgroup); tuning[1] workgroups of the per-world solve (0 = 4 per compute
 * unit); tuning[2] bodies per world whose velocity records the per-world
 * solve keeps in LDS (at most 1000, 64 B each; larger worlds work in HBM);
 * tuning[3] autotuning: when positive, the next steps try every candidate
 * shape of the solve (the colour launches; where every world fits one
 * workgroup, the per-world solve at 8, 4, 2 and 1 workgroups per compute unit
 * with records in LDS or HBM) for tuning[3] steps each (one step when that
 * step takes more than 1.25 times the best mean so far), timing the solve
 * with events, then keep the fastest mean (tuning[3] reads 0 again once
 * done).
 * The getter returns the count written; the log of the last autotuning
 * (each candidate's shape tuning[0..2] and its mean solve milliseconds per
 * step) comes from pxh_scene_get_tuning_log, which returns the candidates'
 * count. */
enum { PXH_SOLVE_LAYOUT_AUTO = 0, PXH_SOLVE_LAYOUT_COLORS = 1, PXH_SOLVE_LAYOUT_WORLDS = 2 };
#define PXH_WORLD_SOLVE_MAX_BODIES 1024
int32_t pxh_scene_set_tuning(pxh_scene scene, const int32_t* tuning, int32_t n);
int32_t pxh_scene_get_tuning(pxh_scene scene, int32_t* tuning, int32_t n);
int32_t pxh_scene_get_tuning_log(pxh_scene scene, int32_t* shapes, double* ms, int32_t max);

/* The scene's HIP stream (hipStream_t), for ordering caller work. */
void* pxh_scene_get_stream(pxh_scene scene);

/* Convex hull cooking on the host: PxCreateConvexMesh with
 * PxConvexFlag::eCOMPUTE_CONVEX (physx/include/cooking/PxCooking.h:528) read
 * back through PxConvexMesh::getVertices / getPolygonData / getIndexBuffer /
 * getMassInformation (physx/include/geometry/PxConvexMesh.h:97-145).
 * pts[n*3] in; out (host arrays, capacities PXH_HULL_MAX_*): verts[nv*3]
 * (bit-identical input points), planes[np*4] (n xyz outward, d), polys[np*2]
 * (first, count into vrefs), vrefs[] (counter-clockwise loops), edges[ne*4]
 * (v0, v1, face left, face right), mass[13] (volume, centre of mass xyz,
 * inertia about it at unit density, row-major 3x3), shape[10] (local bounds
 * centre xyz, extents xyz, internal radius, internal extents xyz), counts[4]
 * = nv, np, number of vrefs, ne. Returns 0 or a negative code. */
enum { PXH_HULL_MAX_VERTS = 255, PXH_HULL_MAX_POLYS = 255, PXH_HULL_MAX_VREFS = 1530, PXH_HULL_MAX_EDGES = 765 };
int32_t pxh_convex_hull(const float* pts, int32_t n, float* verts, float* planes, int32_t* polys, int32_t* vrefs,
                        int32_t* edges, double* mass, float* shape, int32_t* counts);

/* Registers k convex meshes with a scene (PxCreateConvexMesh for each; host
 * arrays): counts[k] points each, points[sum*3]. Bodies of type
 * PXH_GEOM_CONVEXMESH name one by dims[0] = its mesh index (the index of the
 * first registered mesh comes back in *first). */
int32_t pxh_scene_add_convex_meshes(pxh_scene scene, int32_t k, const int32_t* counts, const float* points,
                                    int32_t* first);

/* Registers one convex mesh already cooked, from the arrays a PxConvexMesh
 * gives (physx/include/geometry/PxConvexMesh.h:97-145): verts[nv*3]
 * (getVertices), planes[np*4] and polys[np*2] = (mIndexBase, mNbVerts) of each
 * PxHullPolygon (getPolygonData), indices (getIndexBuffer: counter-clockwise
 * vertex loops), com[3] (getMassInformation's local centre of mass), so the
 * step runs on the very hulls the reference cooked. The edges and the
 * internal extents (the contact-generation margin) follow from them as in
 * the reference cooker (GuCookingConvexMeshBuilder.cpp:420-531). *index
 * receives the mesh index bodies name in dims[0]. At most 255 vertices and
 * polygons, polygons of at most 16 vertices. */
int32_t pxh_scene_add_cooked_convex(pxh_scene scene, int32_t nv, const float* verts, int32_t np, const float* planes,
                                    const int32_t* polys, const uint8_t* indices, const float* com, int32_t* index);

/* Removes whole worlds (every body and plane whose world id is listed in
 * worlds[k], host array) between steps: PxScene::removeActors
 * (physx/include/PxScene.h:437, NpScene.cpp:752) applied to all actors of
 * those worlds. The remaining worlds keep their order and are renumbered
 * 0..W-k-1; their bodies are compacted, so body indices change. Friction
 * patches and colours of the remaining worlds carry over unchanged. */
int32_t pxh_scene_remove_worlds(pxh_scene scene, const int32_t* worlds, int32_t k);

/* Number of worlds, and the body range [first, first + count) of one world
 * (PxScene::getNbActors / getActors, PxScene.h:497). */
int32_t pxh_scene_world_count(pxh_scene scene);
int32_t pxh_scene_world_range(pxh_scene scene, int32_t world, int32_t* first, int32_t* count);

/* Serving loop over one batched scene (no counterpart in the reference, which
 * steps one scene per simulate call): independent step requests are queued,
 * admitted as worlds while max_worlds and the HBM budget allow (each body
 * charged bytes_per_body; no admission while the batched scene's measured
 * device bytes exceed the budget, unless nothing is admitted), stepped
 * together, and retired when their steps are done. */
typedef struct pxh_server_t* pxh_server;
int32_t pxh_server_create(const pxh_scene_desc* desc, int64_t hbm_budget_bytes, int32_t max_worlds,
                          int64_t bytes_per_body, pxh_server* out_server);
int32_t pxh_server_release(pxh_server server);
const char* pxh_server_last_error(void);
/* Queues one scene (arrays as pxh_scene_add_bodies / add_planes; lin_vel and
 * ang_vel may be NULL) to be stepped `steps` times; returns its id. */
int32_t pxh_server_submit(pxh_server server, int32_t n, const int32_t* type, const float* dims, const float* pose,
                          const float* lin_vel, const float* ang_vel, const float* mass, const float* inertia,
                          int32_t m, const float* planes, int32_t steps, int64_t* out_id);
/* Admits queued scenes in arrival order, steps all admitted scenes once by
 * dt, retires the finished ones. */
int32_t pxh_server_step(pxh_server server, float dt, int32_t* admitted, int32_t* completed);
/* A finished scene's bodies (host arrays pose[n*7] p xyz q xyzw, lin_vel[n*3],
 * ang_vel[n*3]) and times[3] = submit, admit, finish (seconds, steady clock).
 * Each result is read once. */
int32_t pxh_server_result(pxh_server server, int64_t id, float* pose, float* lin_vel, float* ang_vel,
                          double* times);
/* Ids of finished scenes not yet read (ascending, at most max_ids); returns the count. */
int32_t pxh_server_finished(pxh_server server, int64_t* ids, int32_t max_ids);
/* out: queued scenes, admitted scenes, admitted bodies, admitted HBM bytes
 * (at the admission price), finished scenes, serving steps, body-steps, device
 * bytes the batched scene held after the last step. Returns the count written. */
int32_t pxh_server_stats(pxh_server server, int64_t* out, int32_t n);

/* Standalone broadphase: PxCreateBroadPhase (physx/include/PxBroadPhase.h:629),
 * PxBroadPhase::update (:576) and fetchResults (:587). Bounds are the
 * non-inflated PxBounds3 (n*6 floats), inflated by distances[n]; pairs whose
 * groups are equal are filtered (lowlevelaabb/include/BpFiltering.h:99-109).
 * All arrays are device pointers. fetch returns device arrays of (id0,id1)
 * uint32 pairs, id0 < id1, sorted, valid until the next update. */
int32_t pxh_bp_create(int32_t capacity, pxh_broadphase* out_bp);
int32_t pxh_bp_update(pxh_broadphase bp, const float* bounds, const uint32_t* groups, const float* distances,
                      int32_t n, void* stream);
/* The same update with the caller's object handles and the reference's
 * group filter: what a Bp::BroadPhase over persistent handles needs
 * (lowlevelaabb/include/BpBroadPhase.h:98-206, PxBroadPhaseUpdateData,
 * PxBroadPhase.h:351-420). ids[n] (device): the handle of each box, every
 * one below id_bound; pairs are reported as handle pairs (id0 < id1), and
 * created / deleted are diffed against the previous update in handle space.
 * filter_lut: boxes of equal groups never pair; otherwise bit
 * (type0 * 8 + type1) of filter_lut, type = group & 7, says whether the two
 * may (Bp::groupFiltering, BpFiltering.h:99-109, with BpFilter's 7 x 7 table,
 * BpFiltering.cpp); PXH_BP_FILTER_ALL lets any two different groups pair, as
 * pxh_bp_update does. */
#define PXH_BP_FILTER_ALL 0xffffffffffffffffull
int32_t pxh_bp_update_handles(pxh_broadphase bp, const float* bounds, const uint32_t* groups, const float* distances,
                              const uint32_t* ids, uint32_t id_bound, uint64_t filter_lut, int32_t n, void* stream);
int32_t pxh_bp_fetch_results(pxh_broadphase bp, const uint32_t** created, int32_t* n_created,
                             const uint32_t** deleted, int32_t* n_deleted);
int32_t pxh_bp_release(pxh_broadphase bp);

/* Allocation-free broadphase update. No counterpart in the reference: for
 * callers that cannot allocate device memory (serving loops, harnesses that
 * forbid hipMalloc). The pxh_bp_update computation (same kernels, same pair
 * sets bit for bit) with every intermediate in one fixed device array of
 * PXH_BP_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends; no
 * broadphase object. Inputs as pxh_bp_update (device arrays, n at most
 * PXH_BP_NOALLOC_MAX_BOXES, at most PXH_BP_NOALLOC_MAX_PAIRS overlapping
 * pairs); outputs: created[max_created*2] and deleted[max_deleted*2]
 * (device, sorted (id0, id1) pairs) and their counts (host). created /
 * deleted are diffed against the last successful call on the same GPU: the
 * device array keeps that call's pairs, and calls on one GPU are serialised
 * within the process, so interleaved callers see each other's updates. Sizes
 * past the maxima fail with PXH_ERR_CAPACITY (the state is unchanged); so do
 * outputs too small for the result (*n_created / *n_deleted then hold the
 * sizes needed, and the update is dropped). Returns when the outputs are
 * written. */
#define PXH_BP_NOALLOC_MAX_BOXES 65536
#define PXH_BP_NOALLOC_MAX_PAIRS 262144
#define PXH_BP_NOALLOC_SCRATCH_BYTES (32u << 20)
int32_t pxh_bp_update_noalloc(const float* bounds, const uint32_t* groups, const float* distances, int32_t n,
                              uint32_t* created, int32_t max_created, int32_t* n_created, uint32_t* deleted,
                              int32_t max_deleted, int32_t* n_deleted, void* stream);

/* Batched immediate-mode contact generation on a fresh cache:
 * PxGenerateContacts (physx/include/PxImmediateMode.h:267). Device arrays:
 * type0/type1[n], geom0/geom1[n*4], pose0/pose1[n*7] (p xyz, q xyzw); out
 * [n*4*8]: per pair up to 4 contacts of (point xyz, separation, normal xyz, 0);
 * counts[n] (-1 for a pair kind without a generator here). Pairs: sphere-
 * sphere, sphere-plane, sphere-capsule, sphere-box, plane-capsule, plane-box,
 * capsule-capsule, capsule-box, box-box, in that shape order (lower geometry
 * type first). Normals point from shape1 to shape0. */
int32_t pxh_generate_contacts(int32_t n, const int32_t* type0, const float* geom0, const float* pose0,
                              const int32_t* type1, const float* geom1, const float* pose1, float contact_distance,
                              float tolerance_length, float* out, int32_t* counts, void* stream);

#ifdef __cplusplus
}
#endif
#endif
