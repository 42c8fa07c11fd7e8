// Launchers of the rigid-body step kernels (pxh_kernels.hip) and the device
// layouts they share with the host scene (pxh_scene.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pxh {

// broadphase pair classes, spheres and capsules counting as round shapes:
// round-round, round-box, box-box, round-convex, box-convex, convex-convex
// (pairs are written class by class)
constexpr int PAIR_CLASSES = 6;
constexpr int PAIR_CLASS_FIRST_CONVEX = 3;

// Scene-wide convex hull storage (structure of arrays; per hull ranges in info)
struct HullSet
{
	const float4* v;         // vertices (xyz)
	const float4* pl;        // polygon planes (n xyz outward, d)
	const uint32_t* poly;    // first vref (relative to the hull's vref base) | count << 16
	const uint8_t* vr;       // polygon vertex loops (hull-local vertex indices)
	const uint32_t* edge;    // v0 | v1 << 8 | face left << 16 | face right << 24
	const int4* info;        // vertex first, vertex count, polygon first, polygon count
	const int4* info2;       // vref first, edge first, edge count, 0
	const float4* bc;        // local bounds centre, GJK margin
	const float4* be;        // local bounds extents, radius about the shape origin
};

// TGS solver body state, one entry per dynamic body
// (PxTGSSolverBodyVel + PxTGSSolverBodyTxInertia, physx/include/solver/PxSolverDefs.h)
struct SolverBodies
{
	// velocity record, 4 float4 per body at [4 * i] (one 64-byte line for the
	// solver's gathers): linear velocity | inverse mass, angular velocity in
	// sqrt-inertia space, accumulated linear motion this step, accumulated
	// angular motion (sqrt-inertia space)
	float4* vel;
	float4* p;        // body position during the step
	float4* dq;       // rotation accumulated this step
	float4* i0;       // world sqrt inverse inertia (xx yy zz xy)
	float4* i1;       // (xz yz - -)
	float4* origLin;  // pre-integrated linear velocity (friction directions)
	// what the contact preparation reads of a body, packed into one 128-byte
	// line at [8 * i]: position | inverse mass, rotation, sqrt inverse inertia
	// (xx yy zz xy), (xz yz - -), pre-integrated linear velocity, and for PGS
	// the pre-integrated angular velocity
	float4* rec;
};
constexpr int PREP_REC = 8;

// Contact rows in colour order; slot k of constraint o is at [k * C + o]
struct Rows
{
	int4* hI;     // bodyA, bodyB (<0: static), nContacts | nFrictionRows << 8
	float4* hF;   // normal xyz, -
	float4* nA;   // [4] raXnI xyz, separation
	float4* nB;   // [4] rbXnI xyz, reciprocal response (negative: separated point)
	float* nF;    // [4] applied normal impulse
	float4* fT;   // [4] tangent xyz, initial error
	float4* fA;   // [4] raXnI xyz, velocity multiplier
	float4* fB;   // [4] rbXnI xyz (dynamic body B only)
	float* fF;    // [4] applied friction impulse
};

// Friction patches kept from one step to the next, one per constraint row
// (FrictionPatch, lowleveldynamics/src/DyFrictionPatch.h). A patch is six
// float4 at [6 * o]: body-0 normal | anchor count, body-1 normal, the two
// body-0 anchors, the two body-1 anchors. Rows of the previous step are found
// by pair key through an open-addressing table (keys ~0 = empty).
struct PatchIO
{
	const uint64_t* htKey;     // previous step's table (null: no previous step)
	const int* htVal;          // row of the previous step holding the key
	uint32_t htMask;           // table size - 1 (power of two)
	const float4* prev;        // previous step's patches
	const uint32_t* prevBroken;
	const int* prevColor;      // previous step's colour of each row
	int* prevRow;              // per constraint: its row in the previous step, or -1
	const uint64_t* conPKey;   // per constraint: world-relative pair key
	const int* conColor;       // per constraint: this step's colour
	float4* cur;               // this step's patches
	uint64_t* curKey;          // this step's pair keys
	int* curColor;             // this step's colour of each row
	float corrDist;            // friction correlation distance
	// per constraint c, what the preparation gathers, packed into 32 bytes at
	// [2 * c]: body a, body b, contact slot, previous row | pair key (lo, hi)
	int4* conRec;
};

// Static triangle meshes: triangles in world space (vertices with the vertex
// id in w, unit normal, bounds) and each mesh's bounding-volume tree
constexpr int MESH_MAX_TRIS = 128;   // candidate triangles per body
constexpr int MESH_MAXP = 4;         // contact patches (constraints) per body and mesh
struct MeshSet
{
	const float4* v0;
	const float4* v1;
	const float4* v2;
	const float4* n;
	const float4* lo;
	const float4* hi;
	const float4* nodeLo;     // tree nodes, depth-first
	const float4* nodeHi;
	const int4* nodeInfo;     // first leaf entry, leaf count (0: inner node), next node after the subtree
	const int* leafTri;       // triangles of the leaves
};

// Sleeping, one entry per body: the sleep filter of PxsRigidBody
// (lowlevel/software/include/PxsRigidBody.h) and the island manager's node
// state. flags == nullptr: sleeping is off.
constexpr uint32_t SLEEP_ASLEEP = 1u, SLEEP_READY = 2u;
struct SleepState
{
	uint32_t* flags;     // SLEEP_ASLEEP | SLEEP_READY (ready for sleep)
	float* wc;           // wake counter (s)
	float4* accL;        // accumulated motion velocity: linear (world frame)
	float4* accA;        // angular (body frame)
	int* isl;            // island label: smallest body index of the island
	int* counted;        // interactions counted this step (numCountedInteractions)
	const int* jcount;   // joints per body
	uint32_t* wakeIsl;   // per label: wake the island at the next step
	uint32_t* notReady;  // per label: the island has a body not ready for sleep
	int* parent;         // island union-find
	float threshold;     // PxRigidDynamic::setSleepThreshold
};

int64_t scan_tmp_size(int64_t n);
void launch_scan(uint32_t* data, int64_t n, uint32_t* tmp, hipStream_t st);
// keys per radix-sort workgroup (its histogram holds 256 counts per workgroup)
int radix_tile();
void launch_radix_sort(uint64_t* keys, uint32_t* vals, uint64_t* keys2, uint32_t* vals2, int n, int bitBegin,
                       int bitEnd, uint32_t* hist, uint32_t* scanTmp, hipStream_t st, bool& resultInAlt);

void launch_bounds(int n, const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, float co,
                   float4* bmin, float4* bmax, hipStream_t st);
// rowOff: per-world row table (nScenes + 1), each world's slice capping its rows
void launch_scene_grid(int n, int nScenes, const int* scene, const float4* bmin, const float4* bmax, uint32_t* acc,
                       float4* gridF, int2* gridI, const uint32_t* rowOff, hipStream_t st);
void launch_row_off_single(uint32_t* rowOff, uint32_t cap, hipStream_t st);
void launch_keys(int n, const float4* bmin, const float4* bmax, const int* scene, const float4* gridF,
                 const int2* gridI, const uint32_t* rowOff, uint64_t* keys, uint32_t* vals, hipStream_t st);
void launch_gather_sorted(int n, const uint32_t* sv, const float4* bmin, const float4* bmax, float4* smin,
                          float4* smax, hipStream_t st);
void launch_row_start(int n, const uint64_t* sk, uint32_t* rowStart, uint32_t totalRows, hipStream_t st);
// write pass: pairs beyond pairCap are counted but not written; group (may be
// null: no filter) with groupLut, the filter table of BpFilter as a bit mask
// (bit type0 * 8 + type1, type = group & 7), equal groups never pairing
constexpr uint64_t GROUP_LUT_ALL = ~0ull;
void launch_sweep(bool write, int n, const uint64_t* sk, const float4* smin, const float4* smax, const int* scene,
                  const int* group, uint64_t groupLut, const int2* gridI, const uint32_t* rowOff,
                  const uint32_t* rowStart, uint32_t* counts, const uint32_t* offsets, uint2* pairs, uint32_t pairCap,
                  hipStream_t st);
// pair p < min(cap, *dP) fills slot slotBase + p; sphere/box pairs are those
// below *dConvex (the first convex pair), convex pairs run in launch_np_convex
void launch_np_pairs(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                     const float4* pos, const float4* rot, const float4* geom, float cd, float tol, float4* ctN,
                     float4* ctP, hipStream_t st);
// convex pairs; scratch: wit, 3 float4 per pair (separation records), and
// list, 2 + 2 cap entries (counts of touching and searched pairs, then the
// two lists), whose first two entries the caller zeroes
void launch_np_convex(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                      const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, float cd, float tol,
                      float4* wit, uint32_t* list, float4* ctN, float4* ctP, hipStream_t st);
// body-plane pairs: launch_np_planes skips convex bodies, launch_np_planes_convex takes only them
void launch_np_planes(int Q, int slotBase, const int* qBody, const int* qPlane, const float4* pos, const float4* rot,
                      const float4* geom, const float4* planes, float cd, float4* ctN, float4* ctP, hipStream_t st);
void launch_np_planes_convex(int Q, int slotBase, const int* qBody, const int* qPlane, const float4* pos,
                             const float4* rot, const float4* geom, const float4* planes, const HullSet& hs, float cd,
                             float tol, float4* ctN, float4* ctP, hipStream_t st);
// contact records of the last step (128 bytes per constraint, constraint order)
void launch_copy_contacts(int C, int maxPairs, const int* order, const int* conA, const int* conB, const int* conSlot,
                          const float4* ctN, const float4* ctP, const float* nF, float4* out, uint32_t* nbPairs,
                          hipStream_t st);
// slots [0, Q) are body-plane pairs, body pairs follow up to the device count *dP
void launch_touch_flags(int S, int Q, int pairEnd, const uint32_t* dP, const float4* ctN, uint32_t* flags,
                        hipStream_t st);
// constraint list; conPKey: world-relative pair identity (world tag 16 | local a
// 24 | local b or 0x800000 | plane 24)
void launch_con_build(int S, int Q, int pairEnd, const uint32_t* dP, const float4* ctN, const uint32_t* scanned,
                      const uint2* pairs, const int* qBody, const int* qPlane, const int* meshBody, int nPlanes,
                      const int* local, const int* planeLocal, const int* world, const int* worldTag, int* conA,
                      int* conB, int* conSlot, uint64_t* conKey, uint64_t* conPKey, int* conColor, hipStream_t st);
// out[i] = in[idx[i]], elements of 4 or 16 bytes
void launch_gather(int n, const int* idx, const void* in, void* out, int bytes, hipStream_t st);
// colouring over list[i] (identity when null) for i below *mDev (or m when
// mDev is null); m bounds the count
void launch_color_round(int m, const uint32_t* mDev, const int* list, int round, const int* conA, const int* conB,
                        const uint64_t* conKey, int* conColor, uint64_t* best, uint64_t* mask, hipStream_t st);
void launch_color_uncolored(int m, const uint32_t* mDev, const int* list, const int* conColor, uint32_t* flags,
                            hipStream_t st);
void launch_color_compact(int m, const int* list, const uint32_t* flags, const uint32_t* scanned, int* out,
                          hipStream_t st);
void launch_fill_u64(uint64_t* p, int n, uint64_t v, hipStream_t st);
void launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
void launch_color_keys(int C, const int* conColor, uint32_t limit, uint64_t* keys, uint32_t* vals, hipStream_t st);
void launch_static_colors(int n, int C, const int* conA, const int* conB, int* conColor, uint32_t* bodyNext,
                          hipStream_t st);
void launch_patch_table(int m, const uint64_t* keys, uint64_t* htKey, int* htVal, uint32_t htMask, hipStream_t st);
// finds each constraint's row of the previous step; body-body constraints keep
// their previous colour (marked in the per-body colour masks)
void launch_con_prev(int C, const int* conA, const int* conB, const int* conSlot, PatchIO io, int* conColor,
                     uint64_t* mask, hipStream_t st);
void launch_max_color(int C, const int* conColor, uint32_t* maxColor, hipStream_t st);
// extLin / extAng (may be null: none set): accelerations set through the
// Direct GPU API, applied for this step and cleared; prevLin / prevAng (may be
// null) receive the velocities before the step
void launch_preintegrate(int n, const float4* pos, const float4* rot, const float4* lin, const float4* ang,
                         const float4* invI, float3 gdt, float linDampDt, float angDampDt, float maxLinSq,
                         float maxAngSq, bool pgs, SolverBodies sb, float dt, float4* extLin, float4* extAng,
                         float4* prevLin, float4* prevAng, const uint32_t* sleepFlags, hipStream_t st);
void launch_prep(int C, const int* order, const int* conA, const int* conB, const int* conSlot, const float4* ctN,
                 const float4* ctP, const float4* pos, const float4* rot, SolverBodies sb, float invStepDt,
                 float scale, float restDist, float fot, Rows rows, PatchIO patches, unsigned long long* counters,
                 hipStream_t st);
// broken: when not null, each row's friction-clamp flag of this pass is stored there
// evStart/evStop (optional): recorded by the dispatch itself at the kernel's start and end
void launch_solve(int begin, int end, int C, Rows rows, SolverBodies sb, float minPen, float muS, float muD,
                  float invStepDt, float biasScale, float scale, uint32_t* broken, hipStream_t st,
                  hipEvent_t evStart = nullptr, hipEvent_t evStop = nullptr);
// rows in (world, colour) order for the per-world solve
void launch_color_keys_world(int C, const int* conColor, const int* conA, const int* bodyWorld, uint32_t limit,
                             uint64_t* keys, uint32_t* vals, hipStream_t st);
int solve_world_threads();
// every TGS iteration of every world in one launch of `grid` persistent
// workgroups; queueHead: one device counter (zeroed by the launch);
// segStart: nWorlds * CB + 1 row starts of (world, colour); worlds of at most
// ldsBodies bodies keep their velocity records in LDS
void launch_solve_worlds(int grid, int nWorlds, const int* worldOrder, uint32_t* queueHead, const int* bodyStart,
                         const uint32_t* segStart, int CB, int C, Rows rows, SolverBodies sb, int posIters,
                         int velIters, float stepDt, float muS, float muD, float invStepDt, float biasScale,
                         float scale, uint32_t* broken, int ldsBodies, hipStream_t st, hipEvent_t evStart = nullptr,
                         hipEvent_t evStop = nullptr);
// PGS: rows with the unbiased normal errors in rowsU; position passes
// (velocityPass false) then velocity passes; SolverBodies p/dq hold the motion
// velocities between the two
void launch_prep_pgs(int C, const int* order, const int* conA, const int* conB, const int* conSlot,
                     const float4* ctN, const float4* ctP, const float4* pos, const float4* rot, SolverBodies sb,
                     float invDt, float muS, float muD, float restDist, float fot, Rows rows, float* rowsU,
                     PatchIO patches, unsigned long long* counters, hipStream_t st);
void launch_solve_pgs(int begin, int end, int C, Rows rows, const float* rowsU, SolverBodies sb, bool doFriction,
                      bool velocityPass, uint32_t* broken, hipStream_t st, hipEvent_t evStart = nullptr,
                      hipEvent_t evStop = nullptr);
void launch_pgs_save_motion(int n, SolverBodies sb, hipStream_t st);
// write-back; with sleeping on, sleeping bodies keep their state and the
// others take the sleep check (ss.flags null: off)
void launch_pgs_finalize(int n, float dt, float4* pos, float4* rot, float4* lin, float4* ang, const float4* invI,
                         SolverBodies sb, SleepState ss, hipStream_t st);
void launch_integrate(int n, float dt, SolverBodies sb, hipStream_t st);
void launch_finalize(int n, float dt, float4* pos, float4* rot, float4* lin, float4* ang, const float4* invI,
                     SolverBodies sb, SleepState ss, hipStream_t st);
void launch_get_data(int n, const uint32_t* idx, int what, const float4* pos, const float4* rot, const float4* lin,
                     const float4* ang, const float4* prevLin, const float4* prevAng, float invDt, float* out,
                     hipStream_t st);
// writes wake the written bodies when sleeping is on (forces and torques when
// nonzero)
void launch_set_data(int n, const uint32_t* idx, int what, float4* pos, float4* rot, float4* lin, float4* ang,
                     const float4* invI, float4* extLin, float4* extAng, const float* in, SleepState ss,
                     hipStream_t st);
// keys of pairs mapped through ids (device, may be null: box indices)
void launch_pair_keys(int P, const uint2* pairs, const uint32_t* ids, uint64_t* keys, uint32_t* vals, hipStream_t st);
void launch_set_diff(int na, const uint64_t* a, int nb, const uint64_t* b, uint32_t* flag, hipStream_t st);
// touch events: found[c] for the C constraints without a previous row
// (prevRow null: all), lost[r] for the R previous rows none of them found;
// matched: R + 1 scratch flags
void launch_touch_events(int C, const int* prevRow, int R, uint32_t* found, uint32_t* matched, uint32_t* lost,
                         hipStream_t st);
void launch_compact_keys(int na, const uint64_t* a, const uint32_t* flag, const uint32_t* scanned, uint2* out,
                         hipStream_t st);

void launch_np_immediate(int n, const int* type0, const float* geom0, const float* pose0, const int* type1,
                         const float* geom1, const float* pose1, float cd, float tol, float* out, int* counts,
                         hipStream_t st);
void launch_bp_inflate(int n, const float* bounds, const float* dist, const uint32_t* groups, float4* bmin,
                       float4* bmax, int* scene, int* group, int* bodyStart, hipStream_t st);

// joints (pxh_joint.h): float4 per joint of the prepared header and rows
constexpr int JOINT_HEAD = 9, JOINT_ROWS = 48, JOINT_PARAM = 3;
// body pairs of slots [Q, Q + min(*dP, cap))
// whose (lower, higher) body key is in the sorted jointKeys[nKeys] lose their
// contacts (jointed bodies do not collide)
void launch_joint_filter(int cap, int Q, const uint32_t* dP, const uint2* pairs, const uint64_t* jointKeys, int nKeys,
                         float4* ctN, hipStream_t st);
// rows of J joints from the poses at the step's start and the solver bodies
// (joints whose bodies all sleep are skipped; sleepFlags may be null)
void launch_joint_prep(int J, const int4* info, const float4* frame, const float4* param, const float4* pos,
                       const float4* rot, SolverBodies sb, float stepDt, float invStepDt, float invDt, float erp,
                       float lengthScale, float4* head, float4* rows, const uint32_t* sleepFlags, hipStream_t st);
// one pass over joints [begin, end) (one colour)
void launch_joint_solve(int begin, int end, const int4* info, const float4* head, float4* rows, SolverBodies sb,
                        float elapsed, bool conclude, const uint32_t* sleepFlags, hipStream_t st);

// sleeping stages between the narrowphase and the constraint list (slots:
// Q body-plane pairs, then the body pairs [Q, Q + min(*dP, cap))):
// interactions each body counts (broadphase pairs that are not jointed,
// planes whose bounds it overlaps, joints)
void launch_sleep_count(int n, int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody,
                        const int* qPlane, const float4* planes, const float4* bmin, const float4* bmax, float co,
                        const uint64_t* jointKeys, int nKeys, int nMB, const int* meshBody, const int* world,
                        const float4* meshLo, const float4* meshHi, SleepState ss, hipStream_t st);
// touching pairs of an awake and a sleeping body wake the sleeping body's
// island, as do islands a wake-up named; then the pairs of sleeping bodies
// lose their contacts
void launch_sleep_wake(int n, int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody, int nMB,
                       const int* meshBody, float4* ctN, SleepState ss, hipStream_t st);
// islands of the awake bodies over touching pairs and joints; islands whose
// bodies were all ready for sleep go to sleep (velocities zeroed) and their
// pairs lose their contacts
void launch_sleep_islands(int n, int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody, int nMB,
                          const int* meshBody, float4* ctN, int J, const int4* jInfo, float4* lin, float4* ang,
                          SleepState ss, hipStream_t st);
// body-mesh contacts of the nMB bodies of meshBody (worlds with a mesh;
// worldMesh[w] = first tree node, node count, or -1) into slots
// slotBase + k * MESH_MAXP + patch; *overflow set when a body meets more than
// MESH_MAX_TRIS triangles
void launch_np_mesh(int nMB, const int* meshBody, const int* world, const int4* worldMesh, const MeshSet& M,
                    const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, float cd,
                    int slotBase, float4* ctN, float4* ctP, uint32_t* overflow, hipStream_t st);
// PxRigidDynamic::wakeUp of k bodies (device indices)
void launch_sleep_wake_bodies(int k, const int* bodies, SleepState ss, hipStream_t st);
// sleep state readback: wake counter and asleep flag per body
void launch_sleep_read(int n, SleepState ss, float* wc, int* asleep, hipStream_t st);

}  // namespace pxh
