/* CPU oracle for the rigid-body step: a plain-C restatement of the reference
 * PhysX 5.6.1 CPU pipeline for the GPU rigid-body path (ABP broadphase
 * predicate, PCM contact generation on a fresh cache, TGS contact prep/solve,
 * body integration).
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg may load it, and only as the checker; the
 * product library (physx_amd) never includes, links or calls it.
 *
 * Parity pinning (each function is checked in tests/ against the reference
 * CPU SDK built from source in oracle/_ref, plus closed forms):
 *   pxo_bp_pairs             pinned: bit-exact vs PxCreateBroadPhase(eABP)
 *   pxo_generate_contacts    pinned: vs PxGenerateContacts (fresh PxCache)
 *   pxo_world_* (TGS step)   pinned: one-step velocities on order-free scenes
 *                            and long-run statistics vs the reference PxScene
 *   box-box edge-edge case   partial: exact SAT/closest-feature restatement of
 *                            the GJK/EPA fallback (GuPCMContactBoxBox.cpp:922-958)
 *   convex hull contacts     pinned: vs PxGenerateContacts on hulls cooked by the
 *                            reference (exact closest features / minimum
 *                            translation stand in for GJK / EPA)
 *   capsule contacts         pinned: sphere-, capsule-, plane-capsule pair for
 *                            pair vs PxGenerateContacts and closed forms;
 *                            capsule-box / -convex vs PxGenerateContacts as
 *                            point sets (exact segment-polytope separation
 *                            stands in for GJK / EPA on the core segment)
 */
#ifndef PXO_H
#define PXO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { PXO_SPHERE = 0, PXO_PLANE = 1, PXO_CAPSULE = 2, PXO_BOX = 3, PXO_CONVEX = 5 };

/* Convex meshes: registers a cooked hull (verts nv x 3, planes np x 4 with
 * outward unit normals, polys np x 2 = first/count into vrefs, counter-
 * clockwise loops, internal extents xyz) and returns its index; a convex
 * shape's geometry (geom / dims) is that index in element 0. */
int  pxo_hull_add(int nv, const float* verts, int np, const float* planes, const int32_t* polys,
                  const int32_t* vrefs, const float* internal_extents);
void pxo_hull_clear(void);

typedef struct {
	float gravity[3];
	int   solver;             /* 1 = TGS, 0 = PGS */
	int   pos_iters, vel_iters;
	float contact_offset;     /* per shape; pair contact distance = sum */
	float rest_offset;
	float static_friction, dynamic_friction, restitution;
	float linear_damping, angular_damping;
	float max_lin_vel_sq, max_ang_vel_sq;
	float bounce_threshold, friction_offset_threshold, correlation_distance;
	float tolerance_length;
	/* test knob: xored into the colouring order keys (0 = the order the GPU
	 * path uses); a different value only reorders the Gauss-Seidel passes */
	unsigned int order_salt;
	/* sleeping (PxRigidDynamic::setSleepThreshold; 0 = bodies never sleep,
	 * as with a zero threshold) and the threshold of every body */
	int   enable_sleeping;
	float sleep_threshold;
} pxo_desc;

void pxo_default_desc(pxo_desc* d);

/* Contact generation for one pair, PhysX argument order (shape0, shape1):
 * geometry codes as PxGeometryType, poses p xyz + q xyzw. out: 8 floats per
 * contact (point xyz, separation, normal xyz, 0); normal points from shape1
 * towards shape0. Returns the contact count. */
int pxo_generate_contacts(int type0, const float* geom0, const float* pose0, int type1, const float* geom1,
                          const float* pose1, float contact_distance, float tolerance_length, float* out);

/* Inflated-AABB inclusive overlap pairs (ABP predicate) among n boxes, only
 * between boxes of equal world id. bounds n x 6, dist n. Pairs (i<j),
 * lexicographically sorted. Returns the count (may exceed max_pairs). */
long long pxo_bp_pairs(const float* bounds, const float* dist, const int32_t* world, int n, uint32_t* out,
                       long long max_pairs);

/* World: bodies of possibly many independent scenes, stepped together. */
void* pxo_world_create(const pxo_desc* d);
void  pxo_world_release(void* w);
int   pxo_world_add_bodies(void* w, int n, const int32_t* type, const float* dims, const float* pose,
                           const float* lin_vel, const float* ang_vel, const float* mass, const float* inertia,
                           const int32_t* scene);
int   pxo_world_add_planes(void* w, int m, const float* planes, const int32_t* scene);
int   pxo_world_step(void* w, float dt);
void  pxo_world_get_bodies(void* w, float* pose, float* lin_vel, float* ang_vel);
void  pxo_world_set_bodies(void* w, const float* pose, const float* lin_vel, const float* ang_vel);
void  pxo_world_set_forces(void* w, const float* force, const float* torque);
/* Static triangle mesh (PxTriangleMeshGeometry, identity scale, on a static
 * actor at pose p xyz q xyzw) in a world: nv vertices (verts[nv*3]), nt
 * triangles (tris[nt*3], counter-clockwise about the outward normal; the mesh
 * is one-sided). At most one per world. Returns the mesh index, -1 when the
 * world has one, -2 for an index out of range. A step returns -6 when more
 * than 128 triangles meet one body's bounds. */
int   pxo_world_add_trimesh(void* w, int nv, const float* verts, int nt, const int32_t* tris, const float* pose,
                            int32_t scene);
/* Sleep state per body (PxRigidDynamic::getWakeCounter, isSleeping). */
void  pxo_world_get_sleep(void* w, float* wake_counter, int32_t* asleep);
/* PxRigidDynamic::wakeUp on k bodies: wake counter 0.4 s, the rest of each
 * body's island wakes at the next step. */
void  pxo_world_wake(void* w, int k, const int32_t* bodies);
/* Joints without limits or drives (PxFixedJointCreate, PxSphericalJointCreate,
 * PxRevoluteJointCreate, PxPrismaticJointCreate): body0/body1 body indices
 * or -1 for the world frame; frame0/frame1 k x 7 local frames (p xyz, q xyzw)
 * in each body's frame (the world's: world coordinates). Jointed bodies do
 * not collide. TGS only. Returns the first joint's index or -1. */
enum { PXO_JOINT_FIXED = 0, PXO_JOINT_SPHERICAL = 1, PXO_JOINT_REVOLUTE = 2, PXO_JOINT_PRISMATIC = 3,
       PXO_JOINT_DISTANCE = 4 };
int   pxo_world_add_joints(void* w, int k, const int32_t* type, const int32_t* body0, const int32_t* body1,
                           const float* frame0, const float* frame1);
/* limit / drive parameters of revolute and prismatic joints: params[10*i] =
 * flags (1 limit, 2 drive, 4 drive freespin), lower, upper, restitution,
 * bounce threshold, drive velocity, drive force limit, drive gear ratio,
 * limit stiffness, limit damping */
int   pxo_world_set_joint_params(void* w, int k, const int32_t* joints, const float* params);
/* Aggregate of k bodies (PxAggregate): without self-collisions its members
 * never pair with each other. Returns the aggregate id or -1. */
int   pxo_world_add_aggregate(void* w, int k, const int32_t* bodies, int self_collisions);
/* out: [0] BP pairs, [1] body-plane pairs tested, [2] touching pairs (constraints),
 * [3] contact points, [4] colours, [5] friction anchors */
void  pxo_world_stats(void* w, long long* out);
/* last step's constraints: per constraint (bodyA, bodyB or -1-planeIndex,
 * colour, nContacts); returns count */
long long pxo_world_constraints(void* w, int32_t* out, long long max);
/* last step's contacts: 32 floats per constraint in the layout of the
 * library's contact-data records (int32 a, b, count; normal; 4 x (point,
 * separation); 4 applied normal impulses); returns count */
long long pxo_world_contacts(void* w, float* out, long long max);

#ifdef __cplusplus
}
#endif
#endif
