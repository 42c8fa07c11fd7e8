// Host side of the GPU rigid-body step: the C ABI of include/pxhip.h.
//
// A pxh_scene owns one device-resident structure-of-arrays world (bodies of
// one or more independent worlds, their static planes) and the per-step
// buffers; pxh_scene_simulate enqueues the whole step on the scene's stream.
// The reference splits this work across the GPU pipeline objects created by
// PxPhysXGpu (physxgpu/src/PxgPhysXGpu.cpp:153-248) and CPU-side islands and
// partitioning (PxgConstraintPartition.cpp:1940); here every stage runs on
// the device and the host only reads back counts that size the next launches.
#include "pxh_bp_core.h"
#include "pxh_hull.h"
#include "pxh_kernels.h"
#include "../../include/pxhip.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

thread_local std::string gLastError;

int fail(int code, const char* what)
{
	gLastError = what;
	return code;
}

}  // namespace

void pxh::set_last_error(const std::string& what) { gLastError = what; }

namespace {
using pxh::bitsFor;
using pxh::bp_write_pairs;
using pxh::broadphase;
using pxh::readCounts;

// growable device buffer
template <typename T>
struct DBuf
{
	T* p = nullptr;
	size_t cap = 0;
	DBuf() = default;
	DBuf(const DBuf&) = delete;
	DBuf& operator=(const DBuf&) = delete;
	DBuf(DBuf&& o) noexcept : p(o.p), cap(o.cap)
	{
		o.p = nullptr;
		o.cap = 0;
	}
	DBuf& operator=(DBuf&& o) noexcept
	{
		std::swap(p, o.p);
		std::swap(cap, o.cap);
		return *this;
	}
	~DBuf() { release(); }
	hipError_t ensure(size_t n)
	{
		if (n <= cap && p)
			return hipSuccess;
		// doubling: a growing scene reallocates (and drains the device) at
		// most once per doubling of its counts
		size_t nc = std::max<size_t>(2 * n, 64);
		T* q = nullptr;
		hipError_t e = hipMalloc(reinterpret_cast<void**>(&q), nc * sizeof(T));
		if (e != hipSuccess)
			return e;
		if (p)
		{
			// buffers may still be read by enqueued work: drain the device first
			(void)hipDeviceSynchronize();
			(void)hipFree(p);
		}
		p = q;
		cap = nc;
		return hipSuccess;
	}
	// grow keeping the first n_keep elements
	hipError_t grow_keep(size_t n, size_t n_keep, hipStream_t st)
	{
		if (n <= cap && p)
			return hipSuccess;
		// doubling: a growing scene reallocates (and drains the device) at
		// most once per doubling of its counts
		size_t nc = std::max<size_t>(2 * n, 64);
		T* q = nullptr;
		hipError_t e = hipMalloc(reinterpret_cast<void**>(&q), nc * sizeof(T));
		if (e != hipSuccess)
			return e;
		if (p && n_keep)
		{
			e = hipMemcpyAsync(q, p, n_keep * sizeof(T), hipMemcpyDeviceToDevice, st);
			if (e != hipSuccess)
				return e;
			e = hipStreamSynchronize(st);
			if (e != hipSuccess)
				return e;
		}
		if (p)
			(void)hipFree(p);
		p = q;
		cap = nc;
		return hipSuccess;
	}
	void release()
	{
		if (p)
			(void)hipFree(p);
		p = nullptr;
		cap = 0;
	}
	size_t bytes() const { return p ? cap * sizeof(T) : 0; }
};

}  // namespace

struct pxh_scene_t
{
	pxh_scene_desc desc;
	hipStream_t stream = nullptr;
	bool simulating = false;

	// convex meshes (scene-wide, shared by any number of bodies)
	std::vector<float4> hHullV, hHullP, hHullBC, hHullBE;
	std::vector<uint32_t> hHullPoly, hHullEdge;
	std::vector<uint8_t> hHullVR;
	std::vector<int4> hHullInfo, hHullInfo2;
	DBuf<float4> hullV, hullP, hullBC, hullBE;
	DBuf<uint32_t> hullPoly, hullEdge;
	DBuf<uint8_t> hullVR;
	DBuf<int4> hullInfo, hullInfo2;
	DBuf<float4> cvxWit;     // separation records of the convex pairs (3 per pair)
	DBuf<uint32_t> cvxList;  // convex pairs to search and touching ones: 2 counts, 2 lists
	int nConvexBodies = 0;   // convex bodies ever added (the convex kernels run when nonzero)
	pxh::HullSet hullSet() const
	{
		pxh::HullSet h;
		h.v = hullV.p; h.pl = hullP.p; h.poly = hullPoly.p; h.vr = hullVR.p; h.edge = hullEdge.p;
		h.info = hullInfo.p; h.info2 = hullInfo2.p; h.bc = hullBC.p; h.be = hullBE.p;
		return h;
	}

	// bodies
	int n = 0;
	DBuf<float4> pos, rot, lin, ang, geom, invI;
	DBuf<int> scene, local;
	std::vector<int> hScene;              // world of each body
	std::vector<int> hBodyStart{0};       // body range of each world (size nWorlds + 1)
	DBuf<int> bodyStart;
	// planes
	int m = 0;
	DBuf<float4> planes;
	DBuf<int> planeLocal;
	std::vector<int> hPlaneScene;
	std::vector<std::vector<int>> worldPlanes;
	// each world's tag: its identity in the pair keys that carry friction
	// patches and colours between steps; unchanged when other worlds are removed
	std::vector<int> hWorldTag;
	int nextTag = 0;
	DBuf<int> worldTag;
	// body-plane pair table
	int Q = 0;
	DBuf<int> qBody, qPlane;
	bool topologyDirty = true;

	// per-step buffers
	DBuf<float4> bmin, bmax, smin, smax, gridF;
	DBuf<int2> gridI;
	DBuf<uint32_t> rows, rowOff, rowStart, vals, vals2, hist, scanTmp, counts, flags, gridAcc;
	DBuf<uint64_t> keys, keys2;
	DBuf<uint2> pairs;
	DBuf<float4> ctN, ctP;
	DBuf<int> conA, conB, conSlot, conColor, order;
	DBuf<uint64_t> conKey, conPKey, best, mask;
	DBuf<int> colList, colList2;
	DBuf<uint32_t> colFlags;
	DBuf<uint32_t> colorCount, scratch;
	DBuf<unsigned long long> counters;    // contact points, friction rows, dynamic-dynamic constraints
	int64_t statFriction = 0, statDynCon = 0, statRounds = 0, statDynFriction = 0, statSolveLaunches = 0;
	pxh::SolverBodies sb{};
	DBuf<float4> sbVel, sbP, sbDQ, sbI0, sbI1, sbOrig, sbRec;
	DBuf<int4> rhI;
	DBuf<float4> rhF, rnA, rnB, rfT, rfA, rfB;
	DBuf<float> rnF, rfF, rnU;
	// friction patches of the previous and the current step (double-buffered by
	// patchCur), the previous step's row count and the pair-key table
	DBuf<float4> patch[2];
	DBuf<uint64_t> patchKey[2];
	DBuf<uint32_t> patchBroken[2];
	DBuf<int> patchColor[2];
	DBuf<int> prevRow;
	DBuf<int4> conRec;   // packed per-constraint record the preparation gathers
	DBuf<uint64_t> htKey;
	DBuf<int> htVal;
	DBuf<uint32_t> bodyNext;
	int patchCur = 0;
	int nPrevRows = 0;
	// capacities and tables sized on the host so a step needs few host round trips
	uint64_t* sortedKeys = nullptr;       // broadphase keys after the sort
	size_t pairCap = 0;                   // pair buffer capacity (grown when exceeded)
	uint32_t totalRows = 0;               // broadphase row table size (rowOff on the device)
	int maxPlanes = 0;                    // most planes in one world
	DBuf<uint32_t> colCounts;             // colouring compaction offsets
	// Direct GPU API forces and torques as accelerations for the next step
	// (allocated, zeroed, on the first one) and the velocities before the
	// last step (with enable_body_accelerations)
	DBuf<float4> extLin, extAng, prevLin, prevAng;
	bool extPending = false;
	float lastDt = 0.0f;
	// touch events of the last step (report_touch_events): flags, their
	// scans and the compacted world-relative pair keys of found / lost pairs
	DBuf<uint32_t> evFound, evFoundScan, evMatched, evLost, evLostScan;
	DBuf<uint2> evFoundKeys, evLostKeys;
	int evC = -1, evR = 0;   // constraints and previous rows of the last step (-1: no events yet)
	// per-world solve (the batched form): worlds by descending size, the
	// (world, colour) row starts, the work-queue head and the largest colour
	int maxWorldBodies = 0;
	DBuf<int> worldOrder;
	DBuf<uint32_t> segStart, solveQueue;
	bool lastWorldSolve = false;          // the last step used the per-world solve
	// launch shapes (pxh_scene_set_tuning): solve layout, per-world solve
	// workgroups (0: 4 per CU), bodies staged in LDS per world
	int32_t tuning[4] = {PXH_SOLVE_LAYOUT_AUTO, 0, 512, 0};
	// autotuning (tuning[3] steps per candidate): every candidate launch shape
	// of the solve runs for that many real steps, its solve time taken from
	// events on the stream; the fastest is kept. Results do not depend on the
	// shape, so tuning during real steps changes no bit of the simulation.
	struct TuneCand
	{
		int32_t t[3];
		double ms = 0.0;
		int steps = 0;
	};
	std::vector<TuneCand> tuneCands, tuneLog;
	size_t tuneIdx = 0;
	hipEvent_t tuneEv[2] = {nullptr, nullptr};
	bool tuneTimed = false;
	// aggregates (PxAggregate): each body's aggregate (-1: none) and whether
	// each aggregate's members collide with each other; the broadphase group
	// of every body (members of an aggregate without self-collisions share
	// one group, so they never pair) rebuilt when these change
	std::vector<int> hAggregate;
	std::vector<char> hAggSelf;
	bool aggDirty = false;
	DBuf<int> bpGroup;
	// joints (TGS): host copies in insertion order (kind, bodies, frames as
	// p xyz, q xyzw); on the device in colour order with the colour ranges and
	// the sorted body-pair keys that keep jointed bodies from colliding
	std::vector<int> hJType, hJA, hJB;
	std::vector<float> hJF0, hJF1;
	std::vector<float> hJP;   // 10 limit / drive parameters per joint
	std::vector<int> jColorStart{0};
	int nJKeys = 0;
	DBuf<int4> jInfo;
	DBuf<float4> jFrame, jParam, jHead, jRows;
	DBuf<uint64_t> jKeys;
	// static triangle meshes, at most one per world: host copies of each mesh's
	// world-space triangles (w of the vertices: vertex id), its broadphase
	// bounds; on the device the triangles and trees of every mesh, per world
	// the mesh's tree range (-1: none) and bounds, and the bodies of the
	// worlds with a mesh
	struct MeshInst
	{
		int world = 0;
		std::vector<float4> v0, v1, v2, n, lo, hi;
		float4 blo{}, bhi{};
	};
	std::vector<MeshInst> meshes;
	bool meshDirty = false;
	int nMeshBodies = 0;
	DBuf<float4> mV0, mV1, mV2, mN, mLo, mHi, mNodeLo, mNodeHi, meshLo, meshHi;
	DBuf<int4> mNodeInfo, worldMesh;
	DBuf<int> mLeafTri, meshBody;
	DBuf<uint32_t> meshOverflow;
	pxh::MeshSet meshSet() const
	{
		pxh::MeshSet m;
		m.v0 = mV0.p; m.v1 = mV1.p; m.v2 = mV2.p; m.n = mN.p; m.lo = mLo.p; m.hi = mHi.p;
		m.nodeLo = mNodeLo.p; m.nodeHi = mNodeHi.p; m.nodeInfo = mNodeInfo.p; m.leafTri = mLeafTri.p;
		return m;
	}
	// sleeping (desc.enable_sleeping): per body the sleep flags, wake counter,
	// accumulated motion velocities and island label; per island label the
	// pending wake-up (set between steps by wake-ups); per step the counted
	// interactions, joints per body, not-ready flags and union-find parents
	DBuf<uint32_t> slFlags, slWakeIsl, slNotReady;
	DBuf<float> slWc;
	DBuf<float4> slAccL, slAccA;
	DBuf<int> slIsl, slCounted, slJCount, slParent, slList;
	DBuf<float> slWcOut;
	DBuf<int> slAsleepOut;
	bool slJCountDirty = true;
	pxh::SleepState sleepState()
	{
		pxh::SleepState ss{};
		if (!desc.enable_sleeping)
			return ss;
		ss.flags = slFlags.p;
		ss.wc = slWc.p;
		ss.accL = slAccL.p;
		ss.accA = slAccA.p;
		ss.isl = slIsl.p;
		ss.counted = slCounted.p;
		ss.jcount = slJCount.p;
		ss.wakeIsl = slWakeIsl.p;
		ss.notReady = slNotReady.p;
		ss.parent = slParent.p;
		ss.threshold = desc.sleep_threshold;
		return ss;
	}
	DBuf<unsigned char> gatherTmp;        // world removal: gather scratch and indices
	DBuf<int> gatherIdx;
	uint32_t* hostCounts = nullptr;       // pinned

	// optional per-stage GPU timing with HIP events on the scene stream
	struct ProfPair { hipEvent_t a, b; int stage; };
	uint32_t profiling = 0;   // bit i: stage i is timed
	uint32_t profStride = 1;  // dispatch-timed launches: one in profStride carries events
	uint64_t profSeq = 0;     // dispatch-timed launches of the current step so far
	uint64_t profStep = 0;    // steps since profiling was set
	std::vector<ProfPair> prof;
	size_t profUsed = 0;
	double profMs[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	int64_t profLaunches[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	int profBegin(int stage)
	{
		if (!(profiling >> stage & 1u))
			return -1;
		if (profUsed == prof.size())
		{
			ProfPair p{};
			if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess)
				return -1;
			prof.push_back(p);
		}
		prof[profUsed].stage = stage;
		(void)hipEventRecord(prof[profUsed].a, stream);
		return int(profUsed++);
	}
	void profEnd(int k)
	{
		if (k >= 0)
			(void)hipEventRecord(prof[k].b, stream);
	}
	// a pair whose events the launch itself records at the kernel's start and
	// end (hipExtLaunchKernelGGL), so the time is the kernel's own
	int profKernel(int stage, hipEvent_t* a, hipEvent_t* b)
	{
		*a = *b = nullptr;
		if (!(profiling >> stage & 1u))
			return -1;
		// launch j of step k is timed when (j + k) % profStride == 0: over
		// profStride consecutive steps every launch index is timed once
		if (profStride > 1 && ((profSeq++ + profStep) % profStride) != 0)
			return -1;
		if (profUsed == prof.size())
		{
			ProfPair p{};
			if (hipEventCreate(&p.a) != hipSuccess || hipEventCreate(&p.b) != hipSuccess)
				return -1;
			prof.push_back(p);
		}
		prof[profUsed].stage = stage;
		*a = prof[profUsed].a;
		*b = prof[profUsed].b;
		return int(profUsed++);
	}
	void profCollect()
	{
		for (size_t k = 0; k < profUsed; k++)
		{
			float ms = 0.0f;
			if (hipEventElapsedTime(&ms, prof[k].a, prof[k].b) == hipSuccess)
			{
				profMs[prof[k].stage] += ms;
				profLaunches[prof[k].stage]++;
			}
		}
		profUsed = 0;
	}

	// statistics of the last step
	int64_t statPairs = 0, statPlanePairs = 0, statConstraints = 0, statContacts = 0, statColors = 0, statSteps = 0;

	// every device buffer of the scene
	template <typename F>
	void forEachBuf(F f)
	{
		f(hullV); f(hullP); f(hullBC); f(hullBE); f(hullPoly); f(hullEdge); f(hullVR); f(hullInfo); f(hullInfo2);
		f(cvxWit); f(cvxList);
		f(pos); f(rot); f(lin); f(ang); f(geom); f(invI); f(scene); f(local); f(bodyStart);
		f(planes); f(planeLocal); f(worldTag); f(qBody); f(qPlane);
		f(bmin); f(bmax); f(smin); f(smax); f(gridF); f(gridI);
		f(rows); f(rowOff); f(rowStart); f(vals); f(vals2); f(hist); f(scanTmp); f(counts); f(flags); f(gridAcc);
		f(keys); f(keys2); f(pairs); f(ctN); f(ctP);
		f(conA); f(conB); f(conSlot); f(conColor); f(order); f(conKey); f(conPKey); f(best); f(mask);
		f(colList); f(colList2); f(colFlags); f(colorCount); f(scratch); f(counters);
		f(sbVel); f(sbP); f(sbDQ); f(sbI0); f(sbI1); f(sbOrig); f(sbRec);
		f(rhI); f(rhF); f(rnA); f(rnB); f(rfT); f(rfA); f(rfB); f(rnF); f(rfF); f(rnU);
		for (int k = 0; k < 2; k++)
		{
			f(patch[k]); f(patchKey[k]); f(patchBroken[k]); f(patchColor[k]);
		}
		f(prevRow); f(conRec); f(htKey); f(htVal); f(bodyNext); f(colCounts); f(gatherTmp); f(gatherIdx);
		f(worldOrder); f(segStart); f(solveQueue); f(extLin); f(extAng); f(prevLin); f(prevAng);
		f(evFound); f(evFoundScan); f(evMatched); f(evLost); f(evLostScan); f(evFoundKeys); f(evLostKeys);
		f(jInfo); f(jFrame); f(jParam); f(jHead); f(jRows); f(jKeys); f(bpGroup);
	}
	// device bytes the scene holds (PxSimulationStatistics::gpuMemHeap)
	int64_t deviceBytes()
	{
		int64_t total = 0;
		forEachBuf([&](const auto& b) { total += int64_t(b.bytes()); });
		return total;
	}
};

struct pxh_broadphase_t
{
	hipStream_t stream = nullptr;
	int n = 0;
	DBuf<float4> bmin, bmax, smin, smax, gridF;
	DBuf<int2> gridI;
	DBuf<int> scene, group, bodyStart;
	DBuf<uint32_t> rows, rowOff, rowStart, vals, vals2, hist, scanTmp, counts, flags, flags2, gridAcc;
	DBuf<uint64_t> keys, keys2, pkPrev, pkCur, pkTmp;
	DBuf<uint2> pairs, created, deleted;
	uint64_t* sortedKeys = nullptr;
	size_t pairCap = 0;
	int nPrev = 0, nCur = 0, nCreated = 0, nDeleted = 0;
	uint32_t* hostCounts = nullptr;
};

namespace {

int worldCount(const pxh_scene_t* s)
{
	return std::max(int(s->hBodyStart.size()) - 1, int(s->worldPlanes.size()));
}

// whole host vector to a device buffer (grown as needed)
template <typename T>
int uploadVec(DBuf<T>& b, const std::vector<T>& v)
{
	PXH_CHECK(b.ensure(std::max<size_t>(v.size(), 1)));
	if (!v.empty())
		PXH_CHECK(hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
	return PXH_OK;
}

// Bounding-volume tree over triangles items[b, e) of one mesh (their bounds
// lo/hi): split at the median centroid along the widest axis of the
// centroids (ties by triangle index), leaves of at most 4; nodes depth-first,
// info = (first leaf entry, leaf count or 0, index after the subtree).
void buildMeshTree(const std::vector<float4>& lo, const std::vector<float4>& hi, std::vector<int>& items, int b, int e,
                   std::vector<float4>& nlo, std::vector<float4>& nhi, std::vector<int4>& info,
                   std::vector<int>& leafTri)
{
	const int me = int(info.size());
	float4 l = make_float4(3.4e38f, 3.4e38f, 3.4e38f, 0.0f), h = make_float4(-3.4e38f, -3.4e38f, -3.4e38f, 0.0f);
	float4 cl = l, ch = h;
	for (int k = b; k < e; k++)
	{
		const float4 a = lo[items[k]], c = hi[items[k]];
		l.x = std::min(l.x, a.x); l.y = std::min(l.y, a.y); l.z = std::min(l.z, a.z);
		h.x = std::max(h.x, c.x); h.y = std::max(h.y, c.y); h.z = std::max(h.z, c.z);
		const float mx = 0.5f * (a.x + c.x), my = 0.5f * (a.y + c.y), mz = 0.5f * (a.z + c.z);
		cl.x = std::min(cl.x, mx); cl.y = std::min(cl.y, my); cl.z = std::min(cl.z, mz);
		ch.x = std::max(ch.x, mx); ch.y = std::max(ch.y, my); ch.z = std::max(ch.z, mz);
	}
	nlo.push_back(l);
	nhi.push_back(h);
	info.push_back(make_int4(0, 0, 0, 0));
	if (e - b <= 4)
	{
		info[me] = make_int4(int(leafTri.size()), e - b, me + 1, 0);
		for (int k = b; k < e; k++)
			leafTri.push_back(items[k]);
		return;
	}
	const float ex = ch.x - cl.x, ey = ch.y - cl.y, ez = ch.z - cl.z;
	const int axis = (ex >= ey && ex >= ez) ? 0 : (ey >= ez ? 1 : 2);
	auto key = [&](int t) {
		const float4 a = lo[t], c = hi[t];
		return axis == 0 ? a.x + c.x : (axis == 1 ? a.y + c.y : a.z + c.z);
	};
	std::sort(items.begin() + b, items.begin() + e, [&](int x, int y) {
		const float kx = key(x), ky = key(y);
		return kx < ky || (kx == ky && x < y);
	});
	const int mid = (b + e) / 2;
	buildMeshTree(lo, hi, items, b, mid, nlo, nhi, info, leafTri);
	buildMeshTree(lo, hi, items, mid, e, nlo, nhi, info, leafTri);
	info[me] = make_int4(0, 0, int(info.size()), 0);
}

// every mesh's triangles and tree to the device, world ranges and bounds
int uploadMeshes(pxh_scene_t* s)
{
	std::vector<float4> v0, v1, v2, n, lo, hi, nlo, nhi;
	std::vector<int4> info;
	std::vector<int> leaf;
	const int nW = worldCount(s);
	std::vector<int4> wm(size_t(std::max(nW, 1)), make_int4(-1, 0, 0, 0));
	std::vector<float4> wlo(size_t(std::max(nW, 1))), whi(size_t(std::max(nW, 1)));
	for (const auto& m : s->meshes)
	{
		const int tb = int(v0.size()), nb = int(info.size());
		v0.insert(v0.end(), m.v0.begin(), m.v0.end());
		v1.insert(v1.end(), m.v1.begin(), m.v1.end());
		v2.insert(v2.end(), m.v2.begin(), m.v2.end());
		n.insert(n.end(), m.n.begin(), m.n.end());
		lo.insert(lo.end(), m.lo.begin(), m.lo.end());
		hi.insert(hi.end(), m.hi.begin(), m.hi.end());
		std::vector<int> items;
		for (int k = 0; k < int(m.v0.size()); k++)
			items.push_back(k);
		std::vector<float4> tl, th;
		std::vector<int4> ti;
		std::vector<int> tlf;
		if (!items.empty())
			buildMeshTree(m.lo, m.hi, items, 0, int(items.size()), tl, th, ti, tlf);
		const int lb = int(leaf.size());
		for (auto& x : ti)
		{
			if (x.y > 0)
				x.x += lb;
			x.z += nb;
			info.push_back(x);
		}
		nlo.insert(nlo.end(), tl.begin(), tl.end());
		nhi.insert(nhi.end(), th.begin(), th.end());
		for (int t : tlf)
			leaf.push_back(tb + t);
		if (m.world < nW)
		{
			wm[m.world] = make_int4(nb, int(ti.size()), 0, 0);
			wlo[m.world] = m.blo;
			whi[m.world] = m.bhi;
		}
	}
	if (int r = uploadVec(s->mV0, v0)) return r;
	if (int r = uploadVec(s->mV1, v1)) return r;
	if (int r = uploadVec(s->mV2, v2)) return r;
	if (int r = uploadVec(s->mN, n)) return r;
	if (int r = uploadVec(s->mLo, lo)) return r;
	if (int r = uploadVec(s->mHi, hi)) return r;
	if (int r = uploadVec(s->mNodeLo, nlo)) return r;
	if (int r = uploadVec(s->mNodeHi, nhi)) return r;
	if (int r = uploadVec(s->mNodeInfo, info)) return r;
	if (int r = uploadVec(s->mLeafTri, leaf)) return r;
	if (int r = uploadVec(s->worldMesh, wm)) return r;
	if (int r = uploadVec(s->meshLo, wlo)) return r;
	if (int r = uploadVec(s->meshHi, whi)) return r;
	s->meshDirty = false;
	return PXH_OK;
}

int rebuild_topology(pxh_scene_t* s)
{
	// tags of new worlds, in creation order
	while (int(s->hWorldTag.size()) < worldCount(s))
		s->hWorldTag.push_back(s->nextTag++);
	PXH_CHECK(s->worldTag.ensure(std::max<size_t>(s->hWorldTag.size(), 1)));
	if (!s->hWorldTag.empty())
		PXH_CHECK(hipMemcpy(s->worldTag.p, s->hWorldTag.data(), s->hWorldTag.size() * sizeof(int),
		                    hipMemcpyHostToDevice));
	// body-plane pairs: every body against the planes of its world
	std::vector<int> qb, qp;
	for (int i = 0; i < s->n; i++)
	{
		const int w = s->hScene[i];
		if (w < int(s->worldPlanes.size()))
			for (int p : s->worldPlanes[w])
			{
				qb.push_back(i);
				qp.push_back(p);
			}
	}
	// bodies of the worlds with a mesh
	{
		std::vector<char> has(size_t(std::max(worldCount(s), 1)), 0);
		for (const auto& m : s->meshes)
			if (m.world < int(has.size()))
				has[m.world] = 1;
		std::vector<int> mb;
		for (int i = 0; i < s->n; i++)
			if (has[s->hScene[i]])
				mb.push_back(i);
		s->nMeshBodies = int(mb.size());
		if (int r = uploadVec(s->meshBody, mb))
			return r;
		s->meshDirty = true;
	}
	s->Q = int(qb.size());
	PXH_CHECK(s->qBody.ensure(std::max<size_t>(qb.size(), 1)));
	PXH_CHECK(s->qPlane.ensure(std::max<size_t>(qp.size(), 1)));
	if (s->Q)
	{
		PXH_CHECK(hipMemcpy(s->qBody.p, qb.data(), qb.size() * sizeof(int), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->qPlane.p, qp.data(), qp.size() * sizeof(int), hipMemcpyHostToDevice));
	}
	// broadphase rows: world w gets at most 2 * (bodies in w) rows
	const int nW = int(s->hBodyStart.size()) - 1;
	std::vector<uint32_t> ro(size_t(nW) + 1, 0u);
	for (int w = 0; w < nW; w++)
	{
		const uint32_t nb = uint32_t(s->hBodyStart[w + 1] - s->hBodyStart[w]);
		ro[w + 1] = ro[w] + std::min<uint32_t>(std::max<uint32_t>(2u * nb, 1u), 1u << 20);
	}
	s->totalRows = ro[nW];
	PXH_CHECK(s->rowOff.ensure(ro.size()));
	PXH_CHECK(hipMemcpy(s->rowOff.p, ro.data(), ro.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
	s->maxPlanes = 0;
	for (const auto& wp : s->worldPlanes)
		s->maxPlanes = std::max(s->maxPlanes, int(wp.size()));
	// worlds by descending body count (stable): the per-world solve's queue order
	{
		std::vector<int> order(size_t(std::max(nW, 0)));
		s->maxWorldBodies = 0;
		for (int w = 0; w < nW; w++)
		{
			order[w] = w;
			s->maxWorldBodies = std::max(s->maxWorldBodies, s->hBodyStart[w + 1] - s->hBodyStart[w]);
		}
		std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
			return s->hBodyStart[a + 1] - s->hBodyStart[a] > s->hBodyStart[b + 1] - s->hBodyStart[b];
		});
		if (int r = uploadVec(s->worldOrder, order))
			return r;
	}
	PXH_CHECK(s->bodyStart.ensure(s->hBodyStart.size()));
	PXH_CHECK(hipMemcpy(s->bodyStart.p, s->hBodyStart.data(), s->hBodyStart.size() * sizeof(int),
	                    hipMemcpyHostToDevice));
	s->topologyDirty = false;
	return PXH_OK;
}

// pinned host staging of a scene: colour ranges in [0, kHostColorSlots), the
// step's device counters (u64) at kHostCounterSlot
constexpr size_t kHostCountsBytes = 65536;
constexpr int kHostColorSlots = 15000;
constexpr int kHostCounterSlot = 15000;

// b[i] = b[idx[i]] for i < n, through a scratch buffer that stays with the scene
template <typename T>
int gatherBuf(DBuf<T>& b, DBuf<unsigned char>& tmp, const int* idx, int n, hipStream_t st)
{
	PXH_CHECK(tmp.ensure(sizeof(T) * size_t(std::max(n, 1))));
	pxh::launch_gather(n, idx, b.p, tmp.p, int(sizeof(T)), st);
	PXH_CHECK(hipMemcpyAsync(b.p, tmp.p, sizeof(T) * size_t(n), hipMemcpyDeviceToDevice, st));
	return PXH_OK;
}


// appends one hull to the scene's host-side hull arrays
int appendHull(pxh_scene_t* s, const pxh::Hull& h)
{
	for (const pxh::HullPolygon& hp : h.polys)
		if (hp.count > 16)
			return fail(PXH_ERR_UNSUPPORTED, "convex mesh: polygon with more than 16 vertices");
	const int nv = int(h.verts.size() / 3), np = int(h.polys.size()), ne = int(h.edgeVerts.size() / 2);
	s->hHullInfo.push_back(make_int4(int(s->hHullV.size()), nv, int(s->hHullP.size()), np));
	s->hHullInfo2.push_back(make_int4(int(s->hHullVR.size()), int(s->hHullEdge.size()), ne, 0));
	for (int i = 0; i < nv; i++)
		s->hHullV.push_back(make_float4(h.verts[3 * i], h.verts[3 * i + 1], h.verts[3 * i + 2], 0.0f));
	for (const pxh::HullPolygon& hp : h.polys)
	{
		s->hHullP.push_back(make_float4(hp.plane[0], hp.plane[1], hp.plane[2], hp.plane[3]));
		s->hHullPoly.push_back(uint32_t(hp.first) | (uint32_t(hp.count) << 16));
	}
	s->hHullVR.insert(s->hHullVR.end(), h.vref.begin(), h.vref.end());
	for (int e = 0; e < ne; e++)
		s->hHullEdge.push_back(uint32_t(h.edgeVerts[2 * e]) | (uint32_t(h.edgeVerts[2 * e + 1]) << 8) |
		                       (uint32_t(h.edgeFaces[2 * e]) << 16) | (uint32_t(h.edgeFaces[2 * e + 1]) << 24));
	// GJK margin of ConvexHullV: 10% of the smallest internal extent
	// (geomutils/src/gjk/GuVecConvexHull.h:76-93)
	const float me = std::min(h.internalExtents[0], std::min(h.internalExtents[1], h.internalExtents[2]));
	s->hHullBC.push_back(make_float4(h.aabbCenter[0], h.aabbCenter[1], h.aabbCenter[2], me * 0.1f));
	float rad = 0.0f;
	for (int i = 0; i < nv; i++)
		rad = std::max(rad, std::sqrt(h.verts[3 * i] * h.verts[3 * i] + h.verts[3 * i + 1] * h.verts[3 * i + 1] +
		                              h.verts[3 * i + 2] * h.verts[3 * i + 2]));
	// the bounding radius is padded by a relative 1e-5 so float rounding
	// in the pair test never discards a touching pair
	s->hHullBE.push_back(make_float4(h.aabbExtents[0], h.aabbExtents[1], h.aabbExtents[2], rad * 1.00001f + 1e-6f));
	return PXH_OK;
}

// the host-side hull arrays to the device
int uploadHulls(pxh_scene_t* s)
{
	PXH_CHECK(hipStreamSynchronize(s->stream));
	if (int r = uploadVec(s->hullV, s->hHullV))
		return r;
	if (int r = uploadVec(s->hullP, s->hHullP))
		return r;
	if (int r = uploadVec(s->hullBC, s->hHullBC))
		return r;
	if (int r = uploadVec(s->hullBE, s->hHullBE))
		return r;
	if (int r = uploadVec(s->hullPoly, s->hHullPoly))
		return r;
	if (int r = uploadVec(s->hullEdge, s->hHullEdge))
		return r;
	if (int r = uploadVec(s->hullVR, s->hHullVR))
		return r;
	if (int r = uploadVec(s->hullInfo, s->hHullInfo))
		return r;
	return uploadVec(s->hullInfo2, s->hHullInfo2);
}

}  // namespace

extern "C" {

int32_t pxh_get_version(void) { return PXH_VERSION; }

const char* pxh_last_error(void) { return gLastError.c_str(); }

void pxh_default_scene_desc(pxh_scene_desc* d)
{
	if (!d)
		return;
	std::memset(d, 0, sizeof(*d));
	d->gravity[1] = -9.81f;
	d->solver_type = PXH_SOLVER_TGS;
	d->position_iterations = 4;
	d->velocity_iterations = 1;
	d->contact_offset = 0.02f;
	d->rest_offset = 0.0f;
	d->static_friction = 0.5f;
	d->dynamic_friction = 0.5f;
	d->restitution = 0.0f;
	d->linear_damping = 0.0f;
	d->angular_damping = 0.05f;
	d->max_linear_velocity_sq = 1e32f;
	d->max_angular_velocity_sq = 100.0f * 100.0f;
	d->bounce_threshold_velocity = 2.0f;
	d->friction_offset_threshold = 0.04f;
	d->friction_correlation_distance = 0.025f;
	d->tolerance_length = 1.0f;
	d->max_pairs_per_body = 4;
	d->enable_body_accelerations = 0;
	d->report_touch_events = 0;
	d->enable_sleeping = 0;
	d->sleep_threshold = 5e-5f * 10.0f * 10.0f;
}

int32_t pxh_scene_create(const pxh_scene_desc* desc, pxh_scene* out)
{
	if (!desc || !out)
		return fail(PXH_ERR_INVALID, "pxh_scene_create: null argument");
	if (desc->solver_type != PXH_SOLVER_TGS && desc->solver_type != PXH_SOLVER_PGS)
		return fail(PXH_ERR_INVALID, "pxh_scene_create: unknown solver type");
	if (desc->position_iterations < 1 || desc->velocity_iterations < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_create: iteration counts out of range");
	if (desc->max_pairs_per_body < 1)
		return fail(PXH_ERR_INVALID, "pxh_scene_create: max_pairs_per_body must be at least 1");
	if (desc->restitution != 0.0f)
		return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_create: restitution other than 0 is not implemented");
	if (desc->enable_sleeping && !(desc->sleep_threshold > 0.0f))
		return fail(PXH_ERR_INVALID, "pxh_scene_create: sleeping needs a positive sleep threshold");
	if (desc->enable_sleeping && desc->report_touch_events)
		return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_create: touch events do not combine with sleeping");
	int dev = 0;
	PXH_CHECK(hipGetDevice(&dev));
	pxh_scene_t* s = new pxh_scene_t();
	s->desc = *desc;
	if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
	    hipHostMalloc(reinterpret_cast<void**>(&s->hostCounts), kHostCountsBytes, hipHostMallocDefault) != hipSuccess)
	{
		delete s;
		return fail(PXH_ERR_DEVICE, "pxh_scene_create: stream or pinned buffer creation failed");
	}
	*out = s;
	return PXH_OK;
}

int32_t pxh_scene_release(pxh_scene s)
{
	if (!s)
		return PXH_OK;
	(void)hipStreamSynchronize(s->stream);
	s->forEachBuf([](auto& b) { b.release(); });
	for (hipEvent_t e : s->tuneEv)
		if (e)
			(void)hipEventDestroy(e);
	if (s->hostCounts)
		(void)hipHostFree(s->hostCounts);
	if (s->stream)
		(void)hipStreamDestroy(s->stream);
	delete s;
	return PXH_OK;
}

int32_t pxh_scene_add_bodies(pxh_scene s, int32_t n, const int32_t* type, const float* dims, const float* pose,
                             const float* lin_vel, const float* ang_vel, const float* mass, const float* inertia,
                             const int32_t* world)
{
	if (!s || n < 0 || (n > 0 && (!type || !dims || !pose || !mass || !inertia)))
		return fail(PXH_ERR_INVALID, "pxh_scene_add_bodies: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_bodies: not allowed while simulating");
	const int base = s->n;
	const int lastWorld = s->hScene.empty() ? 0 : s->hScene.back();
	std::vector<float4> hp(n), hq(n), hl(n), ha(n), hg(n), hi(n);
	std::vector<int> hw(n), hloc(n);
	for (int i = 0; i < n; i++)
	{
		const int w = world ? world[i] : 0;
		if (w < lastWorld || (i > 0 && w < hw[i - 1]) || w < 0)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_bodies: world ids must be non-decreasing");
		if (type[i] != PXH_GEOM_SPHERE && type[i] != PXH_GEOM_CAPSULE && type[i] != PXH_GEOM_BOX &&
		    type[i] != PXH_GEOM_CONVEXMESH)
			return fail(PXH_ERR_UNSUPPORTED,
			            "pxh_scene_add_bodies: only sphere, capsule, box and convex mesh shapes are implemented");
		if (type[i] == PXH_GEOM_CAPSULE && !(dims[4 * i] > 0.0f && dims[4 * i + 1] >= 0.0f))
			return fail(PXH_ERR_INVALID, "pxh_scene_add_bodies: capsule needs radius > 0 and half height >= 0");
		if (type[i] == PXH_GEOM_CONVEXMESH &&
		    !(dims[4 * i] >= 0.0f && int(dims[4 * i]) < int(s->hHullInfo.size()) && float(int(dims[4 * i])) == dims[4 * i]))
			return fail(PXH_ERR_INVALID, "pxh_scene_add_bodies: convex mesh index out of range");
		if (!(mass[i] > 0.0f))
			return fail(PXH_ERR_INVALID, "pxh_scene_add_bodies: mass must be positive");
		hw[i] = w;
		const float* p = pose + 7 * i;
		hp[i] = make_float4(p[0], p[1], p[2], 0.0f);
		hq[i] = make_float4(p[3], p[4], p[5], p[6]);
		const float im = 1.0f / mass[i];
		hl[i] = make_float4(lin_vel ? lin_vel[3 * i] : 0.0f, lin_vel ? lin_vel[3 * i + 1] : 0.0f,
		                    lin_vel ? lin_vel[3 * i + 2] : 0.0f, im);
		ha[i] = make_float4(ang_vel ? ang_vel[3 * i] : 0.0f, ang_vel ? ang_vel[3 * i + 1] : 0.0f,
		                    ang_vel ? ang_vel[3 * i + 2] : 0.0f, 0.0f);
		hg[i] = make_float4(dims[4 * i], dims[4 * i + 1], dims[4 * i + 2], float(type[i]));
		const float* I = inertia + 3 * i;
		hi[i] = make_float4(I[0] > 0.0f ? 1.0f / I[0] : 0.0f, I[1] > 0.0f ? 1.0f / I[1] : 0.0f,
		                    I[2] > 0.0f ? 1.0f / I[2] : 0.0f, 0.0f);
	}
	// world body ranges and scene-local indices
	for (int i = 0; i < n; i++)
	{
		const int w = hw[i];
		while (int(s->hBodyStart.size()) - 1 <= w)
			s->hBodyStart.push_back(s->hBodyStart.back());
		hloc[i] = s->hBodyStart[w + 1] - s->hBodyStart[w];
		s->hBodyStart[w + 1]++;
		s->hScene.push_back(w);
	}
	const size_t nn = size_t(base) + n;
	hipStream_t st = s->stream;
	PXH_CHECK(s->pos.grow_keep(nn, base, st));
	PXH_CHECK(s->rot.grow_keep(nn, base, st));
	PXH_CHECK(s->lin.grow_keep(nn, base, st));
	PXH_CHECK(s->ang.grow_keep(nn, base, st));
	PXH_CHECK(s->geom.grow_keep(nn, base, st));
	PXH_CHECK(s->invI.grow_keep(nn, base, st));
	PXH_CHECK(s->scene.grow_keep(nn, base, st));
	PXH_CHECK(s->local.grow_keep(nn, base, st));
	if (s->extLin.p)
	{
		// the new bodies start without forces
		PXH_CHECK(s->extLin.grow_keep(nn, base, st));
		PXH_CHECK(s->extAng.grow_keep(nn, base, st));
		if (n > 0)
		{
			PXH_CHECK(hipMemsetAsync(s->extLin.p + base, 0, sizeof(float4) * size_t(n), st));
			PXH_CHECK(hipMemsetAsync(s->extAng.p + base, 0, sizeof(float4) * size_t(n), st));
		}
	}
	s->lastDt = 0.0f;   // no acceleration to read until the next step
	if (s->desc.enable_sleeping)
	{
		// new bodies are awake with the wake counter reset value (20 * 0.02 s,
		// PxSceneDesc::wakeCounterResetValue), each its own island
		PXH_CHECK(s->slFlags.grow_keep(nn, base, st));
		PXH_CHECK(s->slWakeIsl.grow_keep(nn, base, st));
		PXH_CHECK(s->slWc.grow_keep(nn, base, st));
		PXH_CHECK(s->slAccL.grow_keep(nn, base, st));
		PXH_CHECK(s->slAccA.grow_keep(nn, base, st));
		PXH_CHECK(s->slIsl.grow_keep(nn, base, st));
		if (n > 0)
		{
			std::vector<uint32_t> zu(size_t(n), 0u);
			std::vector<float> wc(size_t(n), 20.0f * 0.02f);
			std::vector<float4> z4(size_t(n), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
			std::vector<int> isl;
			for (int i = 0; i < n; i++)
				isl.push_back(base + i);
			PXH_CHECK(hipStreamSynchronize(st));
			PXH_CHECK(hipMemcpy(s->slFlags.p + base, zu.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
			PXH_CHECK(hipMemcpy(s->slWakeIsl.p + base, zu.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
			PXH_CHECK(hipMemcpy(s->slWc.p + base, wc.data(), n * sizeof(float), hipMemcpyHostToDevice));
			PXH_CHECK(hipMemcpy(s->slAccL.p + base, z4.data(), n * sizeof(float4), hipMemcpyHostToDevice));
			PXH_CHECK(hipMemcpy(s->slAccA.p + base, z4.data(), n * sizeof(float4), hipMemcpyHostToDevice));
			PXH_CHECK(hipMemcpy(s->slIsl.p + base, isl.data(), n * sizeof(int), hipMemcpyHostToDevice));
		}
		s->slJCountDirty = true;
	}
	if (n > 0)
	{
		PXH_CHECK(hipMemcpy(s->pos.p + base, hp.data(), n * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->rot.p + base, hq.data(), n * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->lin.p + base, hl.data(), n * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->ang.p + base, ha.data(), n * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->geom.p + base, hg.data(), n * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->invI.p + base, hi.data(), n * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->scene.p + base, hw.data(), n * sizeof(int), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->local.p + base, hloc.data(), n * sizeof(int), hipMemcpyHostToDevice));
	}
	for (int i = 0; i < n; i++)
		s->nConvexBodies += type[i] == PXH_GEOM_CONVEXMESH;
	s->n = int(nn);
	s->topologyDirty = true;
	if (!s->hAggSelf.empty())
	{
		s->hAggregate.resize(size_t(s->n), -1);
		s->aggDirty = true;
	}
	return base;
}

int32_t pxh_scene_add_planes(pxh_scene s, int32_t m, const float* plane, const int32_t* world)
{
	if (!s || m < 0 || (m > 0 && !plane))
		return fail(PXH_ERR_INVALID, "pxh_scene_add_planes: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_planes: not allowed while simulating");
	std::vector<float4> hp(m);
	std::vector<int> hloc(m);
	for (int i = 0; i < m; i++)
	{
		const int w = world ? world[i] : 0;
		if (w < 0)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_planes: negative world id");
		const float* p = plane + 4 * i;
		const float l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
		if (!(l > 0.0f))
			return fail(PXH_ERR_INVALID, "pxh_scene_add_planes: zero normal");
		hp[i] = make_float4(p[0] / l, p[1] / l, p[2] / l, p[3] / l);
		if (int(s->worldPlanes.size()) <= w)
			s->worldPlanes.resize(w + 1);
		hloc[i] = int(s->worldPlanes[w].size());
		s->worldPlanes[w].push_back(s->m + i);
		s->hPlaneScene.push_back(w);
	}
	const size_t nm = size_t(s->m) + m;
	PXH_CHECK(s->planes.grow_keep(nm, s->m, s->stream));
	PXH_CHECK(s->planeLocal.grow_keep(nm, s->m, s->stream));
	if (m > 0)
	{
		PXH_CHECK(hipMemcpy(s->planes.p + s->m, hp.data(), m * sizeof(float4), hipMemcpyHostToDevice));
		PXH_CHECK(hipMemcpy(s->planeLocal.p + s->m, hloc.data(), m * sizeof(int), hipMemcpyHostToDevice));
	}
	s->m = int(nm);
	s->topologyDirty = true;
	return PXH_OK;
}

int32_t pxh_scene_add_aggregate(pxh_scene s, int32_t k, const int32_t* bodies, int32_t self_collisions)
{
	if (!s || k < 0 || (k > 0 && !bodies))
		return fail(PXH_ERR_INVALID, "pxh_scene_add_aggregate: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_aggregate: not allowed while simulating");
	s->hAggregate.resize(size_t(s->n), -1);
	for (int i = 0; i < k; i++)
	{
		if (bodies[i] < 0 || bodies[i] >= s->n)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_aggregate: body index out of range");
		if (s->hAggregate[bodies[i]] >= 0)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_aggregate: a body is already in an aggregate");
		for (int j = 0; j < i; j++)
			if (bodies[j] == bodies[i])
				return fail(PXH_ERR_INVALID, "pxh_scene_add_aggregate: a body is listed twice");
	}
	const int id = int(s->hAggSelf.size());
	s->hAggSelf.push_back(self_collisions ? 1 : 0);
	for (int i = 0; i < k; i++)
		s->hAggregate[bodies[i]] = id;
	s->aggDirty = true;
	return id;
}

namespace {

// Joint colours and device arrays from the host copies. Joints between two
// bodies take colours first fit, newest joint first, over the colours each
// body's joints hold; each joint to the world then takes the colour after all
// of its body's other joint colours (one per such joint, in joint order): the
// order in which the reference partitions a jointed chain, constraints on
// static bodies after their body's dynamic ones
// (lowleveldynamics/src/DyConstraintPartition.cpp:168-179, 604-616).
int uploadJoints(pxh_scene_t* s)
{
	const int J = int(s->hJType.size());
	std::vector<int> color(size_t(J), 0);
	std::vector<uint64_t> used(size_t(std::max(s->n, 1)), 0ull);
	std::vector<int> next(size_t(std::max(s->n, 1)), 0);
	int nColors = 0;
	for (int i = J - 1; i >= 0; i--)
	{
		const int a = s->hJA[i], b = s->hJB[i];
		if (a < 0 || b < 0)
			continue;
		const uint64_t u = used[a] | used[b];
		if (u == ~0ull)
			return fail(PXH_ERR_CAPACITY, "pxh_scene_add_joints: a body is in more than 64 joints");
		int c = 0;
		while ((u >> c) & 1ull)
			c++;
		color[i] = c;
		used[a] |= 1ull << c;
		used[b] |= 1ull << c;
		next[a] = std::max(next[a], c + 1);
		next[b] = std::max(next[b], c + 1);
		nColors = std::max(nColors, c + 1);
	}
	for (int i = 0; i < J; i++)
	{
		const int a = s->hJA[i], b = s->hJB[i];
		if (a >= 0 && b >= 0)
			continue;
		const int body = a >= 0 ? a : b;
		color[i] = next[body]++;
		nColors = std::max(nColors, color[i] + 1);
	}
	// colour order, joint order within a colour
	std::vector<int> order;
	for (int i = 0; i < J; i++)
		order.push_back(i);
	std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return color[x] < color[y]; });
	s->jColorStart.assign(size_t(nColors) + 1, 0);
	for (int i = 0; i < J; i++)
		s->jColorStart[color[i] + 1]++;
	for (int c = 0; c < nColors; c++)
		s->jColorStart[c + 1] += s->jColorStart[c];
	std::vector<int4> info(size_t(std::max(J, 1)));
	std::vector<float4> frame(size_t(4) * std::max(J, 1));
	std::vector<float4> param(size_t(pxh::JOINT_PARAM) * std::max(J, 1), make_float4(0.0f, 0.0f, 0.0f, 0.0f));
	std::vector<uint64_t> keys;
	for (int k = 0; k < J; k++)
	{
		const int i = order[k];
		info[k] = make_int4(s->hJType[i], s->hJA[i], s->hJB[i], color[i]);
		const float* f0 = &s->hJF0[7 * size_t(i)];
		const float* f1 = &s->hJF1[7 * size_t(i)];
		frame[4 * k] = make_float4(f0[0], f0[1], f0[2], 0.0f);
		frame[4 * k + 1] = make_float4(f0[3], f0[4], f0[5], f0[6]);
		frame[4 * k + 2] = make_float4(f1[0], f1[1], f1[2], 0.0f);
		frame[4 * k + 3] = make_float4(f1[3], f1[4], f1[5], f1[6]);
		const float* jp = &s->hJP[10 * size_t(i)];
		param[3 * k] = make_float4(jp[0], jp[1], jp[2], jp[3]);
		param[3 * k + 1] = make_float4(jp[4], jp[5], jp[6], jp[7]);
		param[3 * k + 2] = make_float4(jp[8], jp[9], 0.0f, 0.0f);
		// jointed bodies do not collide unless the joint enables it
		// (PxConstraintFlag::eCOLLISION_ENABLED: parameter flag 8)
		if (s->hJA[i] >= 0 && s->hJB[i] >= 0 && !(int(jp[0]) & 8))
		{
			const uint32_t lo = uint32_t(std::min(s->hJA[i], s->hJB[i])), hi = uint32_t(std::max(s->hJA[i], s->hJB[i]));
			keys.push_back((uint64_t(lo) << 32) | hi);
		}
	}
	std::sort(keys.begin(), keys.end());
	keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
	s->nJKeys = int(keys.size());
	s->slJCountDirty = true;
	PXH_CHECK(hipStreamSynchronize(s->stream));
	PXH_CHECK(s->jInfo.ensure(info.size()));
	PXH_CHECK(s->jFrame.ensure(frame.size()));
	PXH_CHECK(s->jParam.ensure(param.size()));
	PXH_CHECK(s->jHead.ensure(size_t(pxh::JOINT_HEAD) * std::max(J, 1)));
	PXH_CHECK(s->jRows.ensure(size_t(pxh::JOINT_ROWS) * std::max(J, 1)));
	PXH_CHECK(s->jKeys.ensure(std::max<size_t>(keys.size(), 1)));
	PXH_CHECK(hipMemcpy(s->jInfo.p, info.data(), info.size() * sizeof(int4), hipMemcpyHostToDevice));
	PXH_CHECK(hipMemcpy(s->jFrame.p, frame.data(), frame.size() * sizeof(float4), hipMemcpyHostToDevice));
	PXH_CHECK(hipMemcpy(s->jParam.p, param.data(), param.size() * sizeof(float4), hipMemcpyHostToDevice));
	if (!keys.empty())
		PXH_CHECK(hipMemcpy(s->jKeys.p, keys.data(), keys.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
	return PXH_OK;
}

}  // namespace

int32_t pxh_scene_add_joints(pxh_scene s, int32_t k, const int32_t* type, const int32_t* body0, const int32_t* body1,
                             const float* frame0, const float* frame1)
{
	if (!s || k < 0 || (k > 0 && (!type || !body0 || !body1 || !frame0 || !frame1)))
		return fail(PXH_ERR_INVALID, "pxh_scene_add_joints: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_joints: not allowed while simulating");
	for (int i = 0; i < k; i++)
	{
		if (type[i] < PXH_JOINT_FIXED || type[i] > PXH_JOINT_DISTANCE)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_joints: unknown joint type");
		if (body0[i] < -1 || body1[i] < -1 || body0[i] >= s->n || body1[i] >= s->n)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_joints: body index out of range");
		if (body0[i] == body1[i])
			return fail(PXH_ERR_INVALID, "pxh_scene_add_joints: a joint needs two different actors, one dynamic");
		if (body0[i] >= 0 && body1[i] >= 0 && s->hScene[body0[i]] != s->hScene[body1[i]])
			return fail(PXH_ERR_INVALID, "pxh_scene_add_joints: the bodies are in different worlds");
		for (int c = 0; c < 7; c++)
			if (!std::isfinite(frame0[7 * i + c]) || !std::isfinite(frame1[7 * i + c]))
				return fail(PXH_ERR_INVALID, "pxh_scene_add_joints: frame is not finite");
	}
	const int first = int(s->hJType.size());
	for (int i = 0; i < k; i++)
	{
		s->hJType.push_back(type[i]);
		s->hJA.push_back(body0[i]);
		s->hJB.push_back(body1[i]);
		s->hJF0.insert(s->hJF0.end(), frame0 + 7 * i, frame0 + 7 * i + 7);
		s->hJF1.insert(s->hJF1.end(), frame1 + 7 * i, frame1 + 7 * i + 7);
		// no limit or drive; gear ratio 1. A distance joint starts with
		// PxDistanceJointCreate's defaults: max distance 0 enforced,
		// tolerance 0.025 (ExtDistanceJoint.cpp:37-47)
		const float none[10] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
		const float dist[10] = {1.0f, 0.0f, 0.0f, 0.025f, 0.0f, 0.0f, 0.0f, 1.0f, 0.0f, 0.0f};
		const float* d0 = type[i] == PXH_JOINT_DISTANCE ? dist : none;
		s->hJP.insert(s->hJP.end(), d0, d0 + 10);
	}
	if (int r = uploadJoints(s))
	{
		s->hJType.resize(first); s->hJA.resize(first); s->hJB.resize(first);
		s->hJF0.resize(7 * size_t(first)); s->hJF1.resize(7 * size_t(first)); s->hJP.resize(10 * size_t(first));
		(void)uploadJoints(s);
		return r;
	}
	return first;
}

int32_t pxh_scene_set_joint_params(pxh_scene s, int32_t k, const int32_t* joints, const float* params)
{
	if (!s || k < 0 || (k > 0 && (!joints || !params)))
		return fail(PXH_ERR_INVALID, "pxh_scene_set_joint_params: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_set_joint_params: not allowed while simulating");
	const int J = int(s->hJType.size());
	for (int i = 0; i < k; i++)
	{
		if (joints[i] < 0 || joints[i] >= J)
			return fail(PXH_ERR_INVALID, "pxh_scene_set_joint_params: joint index out of range");
		const float* q = params + 10 * size_t(i);
		for (int c = 0; c < 10; c++)
			if (!std::isfinite(q[c]) && !(c == 6 && q[c] > 0.0f))
				return fail(PXH_ERR_INVALID, "pxh_scene_set_joint_params: parameter is not finite");
		const int f = int(q[0]);
		if (float(f) != q[0] || f < 0 || f > 31)
			return fail(PXH_ERR_INVALID, "pxh_scene_set_joint_params: unknown flags");
		const int t = s->hJType[joints[i]];
		if ((f & 7) && t != PXH_JOINT_REVOLUTE && t != PXH_JOINT_PRISMATIC && t != PXH_JOINT_SPHERICAL &&
		    t != PXH_JOINT_DISTANCE)
			return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_set_joint_params: fixed joints take no limits");
		if ((f & 22) && t != PXH_JOINT_REVOLUTE && t != PXH_JOINT_DISTANCE)
			return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_set_joint_params: only revolute joints take a drive");
		if (q[3] < 0.0f || (q[3] > 1.0f && t != PXH_JOINT_DISTANCE) || q[4] < 0.0f || q[6] < 0.0f || q[8] < 0.0f ||
		    q[9] < 0.0f)
			return fail(PXH_ERR_INVALID, "pxh_scene_set_joint_params: restitution outside [0, 1] or a negative "
			                             "bounce threshold, force limit, stiffness or damping");
	}
	for (int i = 0; i < k; i++)
		std::copy(params + 10 * size_t(i), params + 10 * size_t(i) + 10, s->hJP.begin() + 10 * size_t(joints[i]));
	return uploadJoints(s);
}

int32_t pxh_scene_add_convex_meshes(pxh_scene s, int32_t k, const int32_t* counts, const float* points,
                                    int32_t* first)
{
	if (!s || k < 0 || (k > 0 && (!counts || !points)))
		return fail(PXH_ERR_INVALID, "pxh_scene_add_convex_meshes: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_convex_meshes: not allowed while simulating");
	if (first)
		*first = int32_t(s->hHullInfo.size());
	const float* pts = points;
	for (int m = 0; m < k; m++)
	{
		pxh::Hull h;
		if (counts[m] < 4 || pxh::build_hull(pts, counts[m], h))
			return fail(PXH_ERR_INVALID, "pxh_scene_add_convex_meshes: degenerate or oversized point cloud");
		pts += 3 * size_t(counts[m]);
		if (int r = appendHull(s, h))
			return r;
	}
	return uploadHulls(s);
}

int32_t pxh_scene_add_cooked_convex(pxh_scene s, int32_t nv, const float* verts, int32_t np, const float* planes,
                                    const int32_t* polys, const uint8_t* indices, const float* com, int32_t* index)
{
	if (!s || nv < 4 || np < 4 || !verts || !planes || !polys || !indices || !com)
		return fail(PXH_ERR_INVALID, "pxh_scene_add_cooked_convex: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_cooked_convex: not allowed while simulating");
	if (nv > 255 || np > 255)
		return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_add_cooked_convex: more than 255 vertices or polygons");
	pxh::Hull h;
	h.verts.assign(verts, verts + 3 * size_t(nv));
	int refs = 0;
	for (int p = 0; p < np; p++)
	{
		const int base = polys[2 * p], cnt = polys[2 * p + 1];
		if (cnt < 3 || base < 0)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_cooked_convex: polygon with fewer than 3 vertices");
		pxh::HullPolygon hp;
		for (int c = 0; c < 4; c++)
			hp.plane[c] = planes[4 * p + c];
		hp.first = int(h.vref.size());
		hp.count = cnt;
		for (int v = 0; v < cnt; v++)
		{
			const int vi = indices[base + v];
			if (vi >= nv)
				return fail(PXH_ERR_INVALID, "pxh_scene_add_cooked_convex: vertex index out of range");
			h.vref.push_back(uint8_t(vi));
		}
		refs += cnt;
		h.polys.push_back(hp);
	}
	pxh::hull_edges_bounds(h);
	// every edge must join two polygons (a closed hull)
	for (size_t e = 0; e < h.edgeFaces.size() / 2; e++)
		if (h.edgeFaces[2 * e] == h.edgeFaces[2 * e + 1])
			return fail(PXH_ERR_INVALID, "pxh_scene_add_cooked_convex: polygons do not close the hull");
	pxh::hull_internal_objects(h, com);
	if (index)
		*index = int32_t(s->hHullInfo.size());
	if (int r = appendHull(s, h))
		return r;
	return uploadHulls(s);
}

int32_t pxh_scene_remove_worlds(pxh_scene s, const int32_t* worlds, int32_t k)
{
	if (!s || k < 0 || (k > 0 && !worlds))
		return fail(PXH_ERR_INVALID, "pxh_scene_remove_worlds: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_remove_worlds: not allowed while simulating");
	const int nW = worldCount(s);
	while (int(s->hWorldTag.size()) < nW)
		s->hWorldTag.push_back(s->nextTag++);
	std::vector<char> drop(size_t(nW), 0);
	for (int i = 0; i < k; i++)
	{
		if (worlds[i] < 0 || worlds[i] >= nW)
			return fail(PXH_ERR_INVALID, "pxh_scene_remove_worlds: world id out of range");
		drop[worlds[i]] = 1;
	}
	std::vector<int> newIdx(size_t(nW), -1), tags;
	int nw = 0;
	for (int w = 0; w < nW; w++)
		if (!drop[w])
		{
			newIdx[w] = nw++;
			tags.push_back(s->hWorldTag[w]);
		}
	// bodies of the kept worlds, in order
	std::vector<int> keep, hScene, hStart(1, 0), newWorldOfBody;
	const int nBW = int(s->hBodyStart.size()) - 1;
	for (int w = 0; w < nW; w++)
	{
		if (drop[w])
			continue;
		if (w < nBW)
			for (int i = s->hBodyStart[w]; i < s->hBodyStart[w + 1]; i++)
			{
				keep.push_back(i);
				hScene.push_back(newIdx[w]);
			}
		hStart.push_back(int(keep.size()));
	}
	// planes of the kept worlds
	std::vector<int> keepPl, hPlaneScene;
	std::vector<std::vector<int>> worldPlanes(static_cast<size_t>(nw));
	for (int p = 0; p < s->m; p++)
	{
		const int w = s->hPlaneScene[p];
		if (drop[w])
			continue;
		worldPlanes[newIdx[w]].push_back(int(keepPl.size()));
		hPlaneScene.push_back(newIdx[w]);
		keepPl.push_back(p);
	}
	hipStream_t st = s->stream;
	PXH_CHECK(hipStreamSynchronize(st));
	const int nn = int(keep.size()), mm = int(keepPl.size());
	DBuf<int>& idx = s->gatherIdx;
	PXH_CHECK(idx.ensure(std::max<size_t>(keep.size() + keepPl.size(), 1)));
	if (nn > 0)
	{
		PXH_CHECK(hipMemcpy(idx.p, keep.data(), keep.size() * sizeof(int), hipMemcpyHostToDevice));
		DBuf<float4>* f4s[] = {&s->pos, &s->rot, &s->lin, &s->ang, &s->geom, &s->invI, &s->extLin, &s->extAng};
		for (auto* b : f4s)
			if (b->p)
				if (int r = gatherBuf(*b, s->gatherTmp, idx.p, nn, st))
					return r;
		s->lastDt = 0.0f;
		if (int r = gatherBuf(s->local, s->gatherTmp, idx.p, nn, st))
			return r;
		if (s->desc.enable_sleeping)
		{
			// sleep state of the kept bodies; island labels (body indices of
			// the same world) renumbered like the bodies
			if (int r = gatherBuf(s->slFlags, s->gatherTmp, idx.p, nn, st))
				return r;
			if (int r = gatherBuf(s->slWakeIsl, s->gatherTmp, idx.p, nn, st))
				return r;
			if (int r = gatherBuf(s->slWc, s->gatherTmp, idx.p, nn, st))
				return r;
			if (int r = gatherBuf(s->slAccL, s->gatherTmp, idx.p, nn, st))
				return r;
			if (int r = gatherBuf(s->slAccA, s->gatherTmp, idx.p, nn, st))
				return r;
			if (int r = gatherBuf(s->slIsl, s->gatherTmp, idx.p, nn, st))
				return r;
			std::vector<int> isl(static_cast<size_t>(nn), 0), newBody(static_cast<size_t>(keep.back() + 1), -1);
			for (int i = 0; i < nn; i++)
				newBody[keep[i]] = i;
			PXH_CHECK(hipMemcpyAsync(isl.data(), s->slIsl.p, size_t(nn) * sizeof(int), hipMemcpyDeviceToHost, st));
			PXH_CHECK(hipStreamSynchronize(st));
			for (int i = 0; i < nn; i++)
				isl[i] = isl[i] >= 0 && isl[i] < int(newBody.size()) && newBody[isl[i]] >= 0 ? newBody[isl[i]] : i;
			PXH_CHECK(hipMemcpy(s->slIsl.p, isl.data(), size_t(nn) * sizeof(int), hipMemcpyHostToDevice));
			s->slJCountDirty = true;
		}
		PXH_CHECK(s->scene.ensure(size_t(nn)));
		PXH_CHECK(hipMemcpyAsync(s->scene.p, hScene.data(), hScene.size() * sizeof(int), hipMemcpyHostToDevice, st));
	}
	if (mm > 0)
	{
		int* pidx = idx.p + keep.size();
		PXH_CHECK(hipMemcpy(pidx, keepPl.data(), keepPl.size() * sizeof(int), hipMemcpyHostToDevice));
		if (int r = gatherBuf(s->planes, s->gatherTmp, pidx, mm, st))
			return r;
		if (int r = gatherBuf(s->planeLocal, s->gatherTmp, pidx, mm, st))
			return r;
	}
	PXH_CHECK(hipStreamSynchronize(st));
	s->n = nn;
	s->m = mm;
	s->hScene.swap(hScene);
	s->hBodyStart.swap(hStart);
	s->hPlaneScene.swap(hPlaneScene);
	s->worldPlanes.swap(worldPlanes);
	s->hWorldTag.swap(tags);
	s->topologyDirty = true;
	// meshes of the kept worlds, renumbered
	{
		int dropped = 0;
		for (int w = 0; w < nW; w++)
			dropped += drop[w];
		std::vector<pxh_scene_t::MeshInst> kept;
		for (auto& m : s->meshes)
		{
			if (m.world < nW && drop[m.world])
				continue;
			m.world = m.world < nW ? newIdx[m.world] : m.world - dropped;
			kept.push_back(std::move(m));
		}
		s->meshes.swap(kept);
		s->meshDirty = true;
	}
	// aggregate membership of the kept bodies
	if (!s->hAggSelf.empty())
	{
		std::vector<int> agg(size_t(nn), -1);
		for (int i = 0; i < nn; i++)
			agg[i] = keep[i] < int(s->hAggregate.size()) ? s->hAggregate[keep[i]] : -1;
		s->hAggregate.swap(agg);
		s->aggDirty = true;
	}
	// joints of the kept worlds, bodies renumbered (a joint to the world
	// belongs to its body's world)
	if (!s->hJType.empty())
	{
		std::vector<int> newBody(size_t(keep.empty() ? 0 : keep.back() + 1), -1);
		for (int i = 0; i < nn; i++)
			newBody[keep[i]] = i;
		auto remap = [&](int b) { return b < 0 ? -1 : (b < int(newBody.size()) ? newBody[b] : -1); };
		std::vector<int> t, a, b;
		std::vector<float> f0, f1, jp;
		for (size_t i = 0; i < s->hJType.size(); i++)
		{
			const int na = remap(s->hJA[i]), nb = remap(s->hJB[i]);
			if ((s->hJA[i] >= 0 && na < 0) || (s->hJB[i] >= 0 && nb < 0))
				continue;
			t.push_back(s->hJType[i]);
			a.push_back(na);
			b.push_back(nb);
			f0.insert(f0.end(), s->hJF0.begin() + 7 * i, s->hJF0.begin() + 7 * i + 7);
			f1.insert(f1.end(), s->hJF1.begin() + 7 * i, s->hJF1.begin() + 7 * i + 7);
			jp.insert(jp.end(), s->hJP.begin() + 10 * i, s->hJP.begin() + 10 * i + 10);
		}
		s->hJType.swap(t); s->hJA.swap(a); s->hJB.swap(b); s->hJF0.swap(f0); s->hJF1.swap(f1); s->hJP.swap(jp);
		if (int r = uploadJoints(s))
			return r;
	}
	return PXH_OK;
}

int32_t pxh_scene_add_triangle_mesh(pxh_scene s, int32_t nv, const float* verts, int32_t nt, const int32_t* tris,
                                    const float* pose, int32_t world)
{
	if (!s || nv <= 0 || nt <= 0 || !verts || !tris || !pose || world < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_add_triangle_mesh: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_add_triangle_mesh: not allowed while simulating");
	for (const auto& m : s->meshes)
		if (m.world == world)
			return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_add_triangle_mesh: one mesh per world");
	for (int k = 0; k < 3 * nt; k++)
		if (tris[k] < 0 || tris[k] >= nv)
			return fail(PXH_ERR_INVALID, "pxh_scene_add_triangle_mesh: vertex index out of range");
	// PxQuat::rotate and the quaternion's matrix, in the oracle's arithmetic
	const float qx = pose[3], qy = pose[4], qz = pose[5], qw = pose[6];
	auto rot = [&](float x, float y, float z, float* o) {
		const float vx = 2.0f * x, vy = 2.0f * y, vz = 2.0f * z;
		const float w2 = qw * qw - 0.5f;
		const float dot2 = qx * vx + qy * vy + qz * vz;
		o[0] = vx * w2 + (qy * vz - qz * vy) * qw + qx * dot2;
		o[1] = vy * w2 + (qz * vx - qx * vz) * qw + qy * dot2;
		o[2] = vz * w2 + (qx * vy - qy * vx) * qw + qz * dot2;
	};
	std::vector<float> P(size_t(3) * nv);
	float lmin[3] = {3.4e38f, 3.4e38f, 3.4e38f}, lmax[3] = {-3.4e38f, -3.4e38f, -3.4e38f};
	for (int k = 0; k < nv; k++)
	{
		float r[3];
		rot(verts[3 * k], verts[3 * k + 1], verts[3 * k + 2], r);
		for (int c = 0; c < 3; c++)
		{
			P[3 * k + c] = r[c] + pose[c];
			lmin[c] = std::min(lmin[c], verts[3 * k + c]);
			lmax[c] = std::max(lmax[c], verts[3 * k + c]);
		}
	}
	pxh_scene_t::MeshInst m;
	m.world = world;
	for (int k = 0; k < nt; k++)
	{
		const int a = tris[3 * k], b = tris[3 * k + 1], c = tris[3 * k + 2];
		const float* A = &P[3 * a];
		const float* B = &P[3 * b];
		const float* C = &P[3 * c];
		const float e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]}, e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
		const float n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
		const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
		if (!(nn > 0.0f))
			continue;   // degenerate triangle: not stored
		const float inv = 1.0f / std::sqrt(nn);
		auto vtx = [&](const float* v, int id) {
			float id_f;
			std::memcpy(&id_f, &id, sizeof(float));
			return make_float4(v[0], v[1], v[2], id_f);
		};
		m.v0.push_back(vtx(A, a));
		m.v1.push_back(vtx(B, b));
		m.v2.push_back(vtx(C, c));
		m.n.push_back(make_float4(n[0] * inv, n[1] * inv, n[2] * inv, 0.0f));
		m.lo.push_back(make_float4(std::fmin(std::fmin(A[0], B[0]), C[0]), std::fmin(std::fmin(A[1], B[1]), C[1]),
		                           std::fmin(std::fmin(A[2], B[2]), C[2]), 0.0f));
		m.hi.push_back(make_float4(std::fmax(std::fmax(A[0], B[0]), C[0]), std::fmax(std::fmax(A[1], B[1]), C[1]),
		                           std::fmax(std::fmax(A[2], B[2]), C[2]), 0.0f));
	}
	if (m.v0.empty())
		return fail(PXH_ERR_INVALID, "pxh_scene_add_triangle_mesh: every triangle is degenerate");
	// broadphase bounds: the local box about the pose (Gu::computeBounds of a
	// triangle mesh shape, geomutils/src/GuBounds.cpp:354-395)
	{
		const float cx = 0.5f * (lmin[0] + lmax[0]), cy = 0.5f * (lmin[1] + lmax[1]), cz = 0.5f * (lmin[2] + lmax[2]);
		const float ex = 0.5f * (lmax[0] - lmin[0]), ey = 0.5f * (lmax[1] - lmin[1]), ez = 0.5f * (lmax[2] - lmin[2]);
		float c[3];
		rot(cx, cy, cz, c);
		const float x2 = qx + qx, y2 = qy + qy, z2 = qz + qz;
		const float xx = x2 * qx, yy = y2 * qy, zz = z2 * qz, xy = x2 * qy, xz = x2 * qz, xw = x2 * qw;
		const float yz = y2 * qz, yw = y2 * qw, zw = z2 * qw;
		const float R[3][3] = {{1.0f - yy - zz, xy - zw, xz + yw}, {xy + zw, 1.0f - xx - zz, yz - xw},
		                       {xz - yw, yz + xw, 1.0f - xx - yy}};   // rows
		float w[3];
		for (int r = 0; r < 3; r++)
			w[r] = std::fabs(R[r][0]) * ex + std::fabs(R[r][1]) * ey + std::fabs(R[r][2]) * ez;
		m.blo = make_float4(c[0] + pose[0] - w[0], c[1] + pose[1] - w[1], c[2] + pose[2] - w[2], 0.0f);
		m.bhi = make_float4(c[0] + pose[0] + w[0], c[1] + pose[1] + w[1], c[2] + pose[2] + w[2], 0.0f);
	}
	s->meshes.push_back(std::move(m));
	s->meshDirty = true;
	s->topologyDirty = true;
	return int32_t(s->meshes.size()) - 1;
}

int32_t pxh_scene_add_heightfield(pxh_scene s, int32_t rows, int32_t cols, const int16_t* heights,
                                  const uint8_t* tess, float height_scale, float row_scale, float column_scale,
                                  const float* pose, int32_t world)
{
	if (!s || rows < 2 || cols < 2 || !heights || !pose || !(row_scale > 0.0f) || !(column_scale > 0.0f) ||
	    !(height_scale > 0.0f))
		return fail(PXH_ERR_INVALID, "pxh_scene_add_heightfield: bad arguments");
	// the heightfield's vertices and triangles (Gu::HeightField::
	// getTriangleVertexIndices, geomutils/src/hf/GuHeightField.h:708-760):
	// vertex row * cols + col at (row * row_scale, height * height_scale,
	// col * column_scale); cell (row, col) splits along the diagonal from its
	// zeroth vertex when that sample's tessellation flag is set
	std::vector<float> v(size_t(3) * rows * cols);
	for (int r = 0; r < rows; r++)
		for (int c = 0; c < cols; c++)
		{
			float* o = &v[3 * (size_t(r) * cols + c)];
			o[0] = float(r) * row_scale;
			o[1] = float(heights[size_t(r) * cols + c]) * height_scale;
			o[2] = float(c) * column_scale;
		}
	std::vector<int32_t> t;
	for (int r = 0; r + 1 < rows; r++)
		for (int c = 0; c + 1 < cols; c++)
		{
			const int cell = r * cols + c;
			if (tess && tess[cell])
			{
				const int a[6] = {cell + cols, cell, cell + cols + 1, cell + 1, cell + cols + 1, cell};
				t.insert(t.end(), a, a + 6);
			}
			else
			{
				const int a[6] = {cell, cell + 1, cell + cols, cell + cols + 1, cell + cols, cell + 1};
				t.insert(t.end(), a, a + 6);
			}
		}
	return pxh_scene_add_triangle_mesh(s, rows * cols, v.data(), int32_t(t.size() / 3), t.data(), pose, world);
}

int32_t pxh_scene_get_sleep_state(pxh_scene s, float* wake_counter, int32_t* asleep, int32_t n)
{
	if (!s || n < 0 || n > s->n)
		return fail(PXH_ERR_INVALID, "pxh_scene_get_sleep_state: bad arguments");
	if (!s->desc.enable_sleeping)
		return fail(PXH_ERR_STATE, "pxh_scene_get_sleep_state: sleeping is off for this scene");
	if (n == 0)
		return PXH_OK;
	hipStream_t st = s->stream;
	PXH_CHECK(s->slWcOut.ensure(size_t(n)));
	PXH_CHECK(s->slAsleepOut.ensure(size_t(n)));
	pxh::launch_sleep_read(n, s->sleepState(), s->slWcOut.p, s->slAsleepOut.p, st);
	if (wake_counter)
		PXH_CHECK(hipMemcpyAsync(wake_counter, s->slWcOut.p, size_t(n) * sizeof(float), hipMemcpyDeviceToHost, st));
	if (asleep)
		PXH_CHECK(hipMemcpyAsync(asleep, s->slAsleepOut.p, size_t(n) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
	PXH_CHECK(hipStreamSynchronize(st));
	return PXH_OK;
}

int32_t pxh_scene_wake_bodies(pxh_scene s, const int32_t* bodies, int32_t k)
{
	if (!s || k < 0 || (k > 0 && !bodies))
		return fail(PXH_ERR_INVALID, "pxh_scene_wake_bodies: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_wake_bodies: not allowed while simulating");
	for (int i = 0; i < k; i++)
		if (bodies[i] < 0 || bodies[i] >= s->n)
			return fail(PXH_ERR_INVALID, "pxh_scene_wake_bodies: body index out of range");
	if (!s->desc.enable_sleeping || k == 0)
		return PXH_OK;   // bodies that never sleep are awake
	hipStream_t st = s->stream;
	PXH_CHECK(s->slList.ensure(size_t(k)));
	PXH_CHECK(hipMemcpyAsync(s->slList.p, bodies, size_t(k) * sizeof(int32_t), hipMemcpyHostToDevice, st));
	pxh::launch_sleep_wake_bodies(k, s->slList.p, s->sleepState(), st);
	PXH_CHECK(hipStreamSynchronize(st));
	return PXH_OK;
}

int32_t pxh_scene_world_range(pxh_scene s, int32_t world, int32_t* first, int32_t* count)
{
	if (!s || world < 0 || world >= worldCount(s))
		return fail(PXH_ERR_INVALID, "pxh_scene_world_range: bad arguments");
	const int nBW = int(s->hBodyStart.size()) - 1;
	const int f = world < nBW ? s->hBodyStart[world] : s->n;
	const int c = world < nBW ? s->hBodyStart[world + 1] - f : 0;
	if (first) *first = f;
	if (count) *count = c;
	return PXH_OK;
}

int32_t pxh_scene_world_count(pxh_scene s)
{
	if (!s)
		return fail(PXH_ERR_INVALID, "pxh_scene_world_count: null scene");
	return worldCount(s);
}

int32_t pxh_scene_simulate(pxh_scene s, float dt)
{
	if (!s || !(dt > 0.0f))
		return fail(PXH_ERR_INVALID, "pxh_scene_simulate: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_simulate: previous step not fetched");
	if (s->topologyDirty)
	{
		int r = rebuild_topology(s);
		if (r)
			return r;
	}
	if (s->meshDirty)
	{
		int r = uploadMeshes(s);
		if (r)
			return r;
	}
	const pxh_scene_desc& d = s->desc;
	if (!s->hJType.empty() && d.solver_type != PXH_SOLVER_TGS)
		return fail(PXH_ERR_UNSUPPORTED, "pxh_scene_simulate: joints run with the TGS solver only");
	hipStream_t st = s->stream;
	const int n = s->n;
	const int nWorlds = int(s->hBodyStart.size()) - 1;
	s->tuneTimed = false;
	if (s->tuning[3] > 0 && n > 0)
	{
		if (s->tuneCands.empty())
		{
			// the colour launches, and where every world fits one workgroup the
			// per-world solve at 8, 4, 2 and 1 workgroups per compute unit with
			// velocity records in LDS or in HBM (the shapes that are usually
			// fastest first, so that slow ones are dropped after one step)
			int dev = 0, cus = 256;
			(void)hipGetDevice(&dev);
			(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
			s->tuneCands.push_back(pxh_scene_t::TuneCand{{PXH_SOLVE_LAYOUT_COLORS, 0, s->tuning[2]}});
			if (d.solver_type == PXH_SOLVER_TGS && s->maxWorldBodies <= PXH_WORLD_SOLVE_MAX_BODIES)
				for (int per : {8, 4, 2, 1})
					for (int lds : {std::min(1000, std::max(s->maxWorldBodies, 1)), 0})
						s->tuneCands.push_back(pxh_scene_t::TuneCand{{PXH_SOLVE_LAYOUT_WORLDS, per * cus, lds}});
			s->tuneIdx = 0;
			for (hipEvent_t& e : s->tuneEv)
				if (!e)
					PXH_CHECK(hipEventCreate(&e));
		}
		for (int i = 0; i < 3; i++)
			s->tuning[i] = s->tuneCands[s->tuneIdx].t[i];
		s->tuneTimed = true;
	}
	s->statSteps++;
	s->profStep += s->profSeq > 0 ? 1 : 0;
	s->profSeq = 0;
	s->statPairs = s->statPlanePairs = s->statConstraints = s->statContacts = s->statColors = 0;
	s->statFriction = s->statDynCon = s->statDynFriction = 0;
	if (n == 0)
		return PXH_OK;
	const float co = d.contact_offset, cd = 2.0f * co;

	// 1. bounds + broadphase; the pair count stays on the device (dP)
	PXH_CHECK(s->bmin.ensure(n));
	PXH_CHECK(s->bmax.ensure(n));
	int pk = s->profBegin(0);
	pxh::launch_bounds(n, s->pos.p, s->rot.p, s->geom.p, s->hullSet(), co, s->bmin.p, s->bmax.p, st);
	s->profEnd(pk);
	s->pairCap = std::max(s->pairCap, size_t(s->desc.max_pairs_per_body) * n);
	if (s->aggDirty)
	{
		// broadphase groups: odd = one body's own, even = an aggregate's
		std::vector<int> grp(size_t(std::max(n, 1)));
		for (int i = 0; i < n; i++)
		{
			const int a = s->hAggregate[i];
			grp[i] = (a >= 0 && !s->hAggSelf[a]) ? 2 * a : 2 * i + 1;
		}
		PXH_CHECK(s->bpGroup.ensure(grp.size()));
		PXH_CHECK(hipMemcpyAsync(s->bpGroup.p, grp.data(), grp.size() * sizeof(int), hipMemcpyHostToDevice, st));
		PXH_CHECK(hipStreamSynchronize(st));
		s->aggDirty = false;
	}
	const int* bpGroup = s->hAggSelf.empty() ? nullptr : s->bpGroup.p;
	int r = broadphase(s, n, nWorlds, s->scene.p, bpGroup, s->rowOff.p, s->totalRows, st, s->pairCap);
	if (r)
		return r;
	const uint32_t* dP = s->counts.p + pxh::PAIR_CLASSES * int64_t(n);
	const int Q = s->Q;
	s->statPlanePairs = Q;

	// sleeping: joints per body when bodies or joints changed; per-step buffers
	const bool sleeping = d.enable_sleeping != 0;
	if (sleeping)
	{
		if (s->slJCountDirty)
		{
			std::vector<int> jc(size_t(n), 0);
			for (size_t j = 0; j < s->hJType.size(); j++)
			{
				if (s->hJA[j] >= 0) jc[s->hJA[j]]++;
				if (s->hJB[j] >= 0) jc[s->hJB[j]]++;
			}
			PXH_CHECK(s->slJCount.ensure(size_t(n)));
			PXH_CHECK(hipStreamSynchronize(st));
			PXH_CHECK(hipMemcpy(s->slJCount.p, jc.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
			s->slJCountDirty = false;
		}
		PXH_CHECK(s->slCounted.ensure(size_t(n)));
		PXH_CHECK(s->slNotReady.ensure(size_t(n)));
		PXH_CHECK(s->slParent.ensure(size_t(n)));
	}

	// 2. narrowphase into per-slot contact sets (slots: body-plane pairs, then
	//    body pairs up to the count on the device) and 3. the constraint list
	//    (touching slots in slot order). One host round trip reads the pair
	//    and constraint counts; when the pairs outgrew their buffer it is
	//    enlarged and the stage repeated.
	int P = 0, C = 0;
	for (int attempt = 0;; attempt++)
	{
		const int cap = int(s->pairCap);
		const int pairEnd = Q + cap, nMB = s->nMeshBodies;
		const int S = pairEnd + nMB * pxh::MESH_MAXP;   // slots: planes, body pairs, mesh patches
		if (attempt > 0)
		{
			r = bp_write_pairs(s, n, s->scene.p, bpGroup, s->rowOff.p, st, s->pairCap);
			if (r)
				return r;
		}
		PXH_CHECK(s->ctN.ensure(size_t(S)));
		PXH_CHECK(s->ctP.ensure(size_t(4) * S));
		const uint32_t* dConvex = s->counts.p + pxh::PAIR_CLASS_FIRST_CONVEX * int64_t(n);
		pk = s->profBegin(3);
		pxh::launch_np_planes(Q, 0, s->qBody.p, s->qPlane.p, s->pos.p, s->rot.p, s->geom.p, s->planes.p, cd, s->ctN.p,
		                      s->ctP.p, st);
		if (s->nConvexBodies > 0)
			pxh::launch_np_planes_convex(Q, 0, s->qBody.p, s->qPlane.p, s->pos.p, s->rot.p, s->geom.p, s->planes.p,
			                             s->hullSet(), cd, d.tolerance_length, s->ctN.p, s->ctP.p, st);
		s->profEnd(pk);
		pk = s->profBegin(2);
		pxh::launch_np_pairs(cap, dP, dConvex, Q, s->pairs.p, s->pos.p, s->rot.p, s->geom.p, cd, d.tolerance_length,
		                     s->ctN.p, s->ctP.p, st);
		if (s->nConvexBodies > 0)
		{
			PXH_CHECK(s->cvxWit.ensure(size_t(3) * std::max(cap, 1)));
			PXH_CHECK(s->cvxList.ensure(2 + 2 * size_t(cap)));
			PXH_CHECK(hipMemsetAsync(s->cvxList.p, 0, 2 * sizeof(uint32_t), st));
			pxh::launch_np_convex(cap, dP, dConvex, Q, s->pairs.p, s->pos.p, s->rot.p, s->geom.p, s->hullSet(), cd,
			                      d.tolerance_length, s->cvxWit.p, s->cvxList.p, s->ctN.p, s->ctP.p, st);
		}
		s->profEnd(pk);
		if (nMB > 0)
		{
			PXH_CHECK(s->meshOverflow.ensure(1));
			pxh::launch_fill_u32(s->meshOverflow.p, 1, 0u, st);
			pxh::launch_np_mesh(nMB, s->meshBody.p, s->scene.p, s->worldMesh.p, s->meshSet(), s->pos.p, s->rot.p,
			                    s->geom.p, s->hullSet(), cd, pairEnd, s->ctN.p, s->ctP.p, s->meshOverflow.p, st);
		}
		pxh::launch_joint_filter(cap, Q, dP, s->pairs.p, s->jKeys.p, s->nJKeys, s->ctN.p, st);
		if (sleeping)
		{
			// the sleep stages act once on the whole pair set: make sure it fit
			const uint32_t* srcP[1] = {dP};
			r = readCounts(st, s->hostCounts, srcP, 1);
			if (r)
				return r;
			if (int(s->hostCounts[0]) > cap)
			{
				if (attempt > 0)
					return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: pair buffer overflow after growth");
				s->pairCap = size_t(s->hostCounts[0]) + size_t(s->hostCounts[0]) / 4;
				continue;
			}
			const pxh::SleepState ss = s->sleepState();
			pxh::launch_sleep_count(n, Q, cap, dP, s->pairs.p, s->qBody.p, s->qPlane.p, s->planes.p, s->bmin.p,
			                        s->bmax.p, co, s->jKeys.p, s->nJKeys, nMB, s->meshBody.p, s->scene.p, s->meshLo.p,
			                        s->meshHi.p, ss, st);
			pxh::launch_sleep_wake(n, Q, cap, dP, s->pairs.p, s->qBody.p, nMB, s->meshBody.p, s->ctN.p, ss, st);
			pxh::launch_sleep_islands(n, Q, cap, dP, s->pairs.p, s->qBody.p, nMB, s->meshBody.p, s->ctN.p,
			                          int(s->hJType.size()), s->jInfo.p, s->lin.p, s->ang.p, ss, st);
		}
		PXH_CHECK(s->flags.ensure(size_t(S) + 1));
		PXH_CHECK(s->scanTmp.ensure(pxh::scan_tmp_size(int64_t(S) + 1) + 1024));
		pxh::launch_touch_flags(S, Q, pairEnd, dP, s->ctN.p, s->flags.p, st);
		PXH_CHECK(hipMemsetAsync(s->flags.p + S, 0, sizeof(uint32_t), st));
		pxh::launch_scan(s->flags.p, int64_t(S) + 1, s->scanTmp.p, st);
		const size_t Sc = std::max(S, 1);
		PXH_CHECK(s->conA.ensure(Sc));
		PXH_CHECK(s->conB.ensure(Sc));
		PXH_CHECK(s->conSlot.ensure(Sc));
		PXH_CHECK(s->conColor.ensure(Sc));
		PXH_CHECK(s->conKey.ensure(Sc));
		PXH_CHECK(s->conPKey.ensure(Sc));
		pxh::launch_con_build(S, Q, pairEnd, dP, s->ctN.p, s->flags.p, s->pairs.p, s->qBody.p, s->qPlane.p,
		                      s->meshBody.p, s->m, s->local.p, s->planeLocal.p, s->scene.p, s->worldTag.p, s->conA.p,
		                      s->conB.p, s->conSlot.p, s->conKey.p, s->conPKey.p, s->conColor.p, st);
		const uint32_t* src[3] = {dP, s->flags.p + S, nMB > 0 ? s->meshOverflow.p : dP};
		r = readCounts(st, s->hostCounts, src, 3);
		if (r)
			return r;
		P = int(s->hostCounts[0]);
		C = int(s->hostCounts[1]);
		if (nMB > 0 && s->hostCounts[2])
			return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: more than 128 mesh triangles meet one body's bounds");
		if (P <= cap)
			break;
		if (attempt > 0)
			return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: pair buffer overflow after growth");
		s->pairCap = size_t(P) + size_t(P) / 4;
	}
	// grow ahead of need so the next steps keep to the single round trip
	if (size_t(P) + size_t(P) / 8 > s->pairCap)
		s->pairCap = size_t(P) + size_t(P) / 4;
	s->statPairs = P;
	s->statConstraints = C;
	PXH_CHECK(s->order.ensure(std::max(C, 1)));

	// 4. previous step's rows: a pair-key table over them gives each
	//    constraint its previous row (friction patch) and, for body-body
	//    constraints, its previous colour
	const size_t Cc = std::max(C, 1);
	const int cur = s->patchCur, prv = cur ^ 1;
	PXH_CHECK(s->best.ensure(n));
	PXH_CHECK(s->mask.ensure(n));
	PXH_CHECK(s->prevRow.ensure(Cc));
	PXH_CHECK(s->conRec.ensure(2 * Cc));
	pxh::launch_fill_u64(s->mask.p, n, 0ull, st);
	pxh::PatchIO pio{};
	pio.corrDist = d.friction_correlation_distance;
	pio.conPKey = s->conPKey.p;
	pio.conRec = s->conRec.p;
	if (s->nPrevRows > 0)
	{
		uint32_t size = 64;
		while (size < 2u * uint32_t(s->nPrevRows))
			size <<= 1;
		PXH_CHECK(s->htKey.ensure(size));
		PXH_CHECK(s->htVal.ensure(size));
		pxh::launch_patch_table(s->nPrevRows, s->patchKey[prv].p, s->htKey.p, s->htVal.p, size - 1, st);
		pio.htKey = s->htKey.p;
		pio.htVal = s->htVal.p;
		pio.htMask = size - 1;
		pio.prev = s->patch[prv].p;
		pio.prevBroken = s->patchBroken[prv].p;
		pio.prevColor = s->patchColor[prv].p;
		pio.prevRow = s->prevRow.p;
	}
	// every step: the packed constraint records (and, after a previous step,
	// each constraint's previous row and colour)
	pxh::launch_con_prev(C, s->conA.p, s->conB.p, s->conSlot.p, pio, s->conColor.p, s->mask.p, st);

	// touch found / lost events: constraints without a previous row, and
	// previous rows that no constraint found, as pair keys
	if (d.report_touch_events)
	{
		const int R = s->nPrevRows;
		PXH_CHECK(s->evFound.ensure(Cc + 1));
		PXH_CHECK(s->evFoundScan.ensure(Cc + 1));
		PXH_CHECK(s->evFoundKeys.ensure(Cc));
		PXH_CHECK(s->evMatched.ensure(size_t(R) + 1));
		PXH_CHECK(s->evLost.ensure(size_t(R) + 1));
		PXH_CHECK(s->evLostScan.ensure(size_t(R) + 1));
		PXH_CHECK(s->evLostKeys.ensure(std::max<size_t>(R, 1)));
		PXH_CHECK(s->scanTmp.ensure(pxh::scan_tmp_size(int64_t(std::max(C, R)) + 1) + 1024));
		pxh::launch_touch_events(C, R > 0 ? s->prevRow.p : nullptr, R, s->evFound.p, s->evMatched.p, s->evLost.p, st);
		pxh::launch_fill_u32(s->evFound.p + C, 1, 0u, st);
		pxh::launch_fill_u32(s->evLost.p + R, 1, 0u, st);
		PXH_CHECK(hipMemcpyAsync(s->evFoundScan.p, s->evFound.p, sizeof(uint32_t) * (size_t(C) + 1),
		                         hipMemcpyDeviceToDevice, st));
		pxh::launch_scan(s->evFoundScan.p, int64_t(C) + 1, s->scanTmp.p, st);
		pxh::launch_compact_keys(C, s->conPKey.p, s->evFound.p, s->evFoundScan.p, s->evFoundKeys.p, st);
		PXH_CHECK(hipMemcpyAsync(s->evLostScan.p, s->evLost.p, sizeof(uint32_t) * (size_t(R) + 1),
		                         hipMemcpyDeviceToDevice, st));
		pxh::launch_scan(s->evLostScan.p, int64_t(R) + 1, s->scanTmp.p, st);
		if (R > 0)
			pxh::launch_compact_keys(R, s->patchKey[prv].p, s->evLost.p, s->evLostScan.p, s->evLostKeys.p, st);
		s->evC = C;
		s->evR = R;
	}

	// 5. colouring of the remaining body-body constraints: rounds of local key
	//    minima, each taking its first free colour, over a compacted list of the
	//    uncoloured constraints whose count stays on the device; then the
	//    body-plane colours and the colour order. One host round trip reads the
	//    colour ranges and what is left uncoloured (a first step, with nothing
	//    carried over, may need further rounds).
	PXH_CHECK(s->colList.ensure(Cc));
	PXH_CHECK(s->colList2.ensure(Cc));
	PXH_CHECK(s->colFlags.ensure(Cc + 1));
	PXH_CHECK(s->colCounts.ensure(Cc + 1));
	PXH_CHECK(s->bodyNext.ensure(n));
	PXH_CHECK(s->keys.ensure(Cc));
	PXH_CHECK(s->keys2.ensure(Cc));
	PXH_CHECK(s->vals.ensure(Cc));
	PXH_CHECK(s->vals2.ensure(Cc));
	{
		const int nb = (C + pxh::radix_tile() - 1) / pxh::radix_tile();
		PXH_CHECK(s->hist.ensure(size_t(256) * std::max(nb, 1)));
		PXH_CHECK(s->scanTmp.ensure(pxh::scan_tmp_size(std::max<int64_t>(int64_t(256) * std::max(nb, 1), C + 1)) +
		                            1024));
	}
	int rounds = 0;
	const int kMaxRounds = 960;
	const int* list = nullptr;         // null: every constraint
	const uint32_t* mDev = nullptr;    // null: count C
	int bound = C;                     // host bound on the list length
	if (64 + kMaxRounds + s->maxPlanes + 2 > kHostColorSlots)
		return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: too many planes in one world");
	auto compact = [&]() {
		int* out = list == s->colList.p ? s->colList2.p : s->colList.p;
		pxh::launch_color_uncolored(bound, mDev, list, s->conColor.p, s->colFlags.p, st);
		(void)hipMemcpyAsync(s->colCounts.p, s->colFlags.p, sizeof(uint32_t) * (size_t(bound) + 1),
		                     hipMemcpyDeviceToDevice, st);
		pxh::launch_scan(s->colCounts.p, int64_t(bound) + 1, s->scanTmp.p, st);
		pxh::launch_color_compact(bound, list, s->colFlags.p, s->colCounts.p, out, st);
		list = out;
		mDev = s->colCounts.p + bound;
	};
	int nColors = 0;
	std::vector<uint32_t> colorStart(1, 0);
	int remaining = C;
	// the per-world solve: TGS, and every world small enough for one workgroup
	// (layout auto), or as the tuning asks
	const int layout = s->tuning[0];
	const bool worldSolve = d.solver_type == PXH_SOLVER_TGS && nWorlds > 0 && s->hJType.empty() &&
	                        (layout == PXH_SOLVE_LAYOUT_WORLDS ||
	                         (layout == PXH_SOLVE_LAYOUT_AUTO && s->maxWorldBodies <= PXH_WORLD_SOLVE_MAX_BODIES));
	s->lastWorldSolve = worldSolve;
	int CB = 0;   // per-world solve: colour slots per world
	if (worldSolve)
		PXH_CHECK(s->solveQueue.ensure(2));
	while (C > 0 && worldSolve)
	{
		compact();
		const int batch = s->nPrevRows > 0 && rounds == 0 ? 8 : 48;
		for (int k = 0; k < batch && rounds < kMaxRounds; k++, rounds++)
			pxh::launch_color_round(bound, mDev, list, rounds, s->conA.p, s->conB.p, s->conKey.p, s->conColor.p,
			                        s->best.p, s->mask.p, st);
		compact();
		pxh::launch_static_colors(n, C, s->conA.p, s->conB.p, s->conColor.p, s->bodyNext.p, st);
		// row order: stable radix sort of (world, colour, index); the row
		// starts of every (world, colour) come from the sorted keys
		const int colorBound = 64 + rounds + s->maxPlanes + 1;
		CB = colorBound + 1;
		const uint64_t slots = uint64_t(nWorlds) * uint64_t(CB);
		if (slots >= (1ull << 31))
			return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: too many (world, colour) slots for the per-world solve");
		PXH_CHECK(s->segStart.ensure(size_t(slots) + 2));
		pxh::launch_color_keys_world(C, s->conColor.p, s->conA.p, s->scene.p, uint32_t(colorBound), s->keys.p,
		                             s->vals.p, st);
		bool alt = false;
		pxh::launch_radix_sort(s->keys.p, s->vals.p, s->keys2.p, s->vals2.p, C, 32,
		                       32 + ((bitsFor(slots) + 7) / 8) * 8, s->hist.p, s->scanTmp.p, st, alt);
		const uint64_t* ck = alt ? s->keys2.p : s->keys.p;
		const uint32_t* cv = alt ? s->vals2.p : s->vals.p;
		pxh::launch_row_start(C, ck, s->segStart.p, uint32_t(slots), st);
		PXH_CHECK(hipMemcpyAsync(s->order.p, cv, sizeof(int) * C, hipMemcpyDeviceToDevice, st));
		pxh::launch_fill_u32(s->solveQueue.p + 1, 1, 0u, st);
		pxh::launch_max_color(C, s->conColor.p, s->solveQueue.p + 1, st);
		PXH_CHECK(hipMemcpyAsync(s->hostCounts, mDev, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		PXH_CHECK(hipMemcpyAsync(s->hostCounts + 1, s->solveQueue.p + 1, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		PXH_CHECK(hipStreamSynchronize(st));
		remaining = int(s->hostCounts[0]);
		if (remaining == 0)
		{
			nColors = int(s->hostCounts[1]) + 1;
			break;
		}
		if (rounds >= kMaxRounds)
			return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: colouring did not finish in 960 rounds");
		bound = remaining;
	}
	if (worldSolve && C == 0)
	{
		// nothing to solve; the per-world solve still integrates every world
		CB = 1;
		PXH_CHECK(s->segStart.ensure(size_t(nWorlds) + 2));
		pxh::launch_fill_u32(s->segStart.p, int64_t(nWorlds) + 1, 0u, st);
	}
	while (C > 0 && !worldSolve)
	{
		compact();
		const int batch = s->nPrevRows > 0 && rounds == 0 ? 8 : 48;
		for (int k = 0; k < batch && rounds < kMaxRounds; k++, rounds++)
			pxh::launch_color_round(bound, mDev, list, rounds, s->conA.p, s->conB.p, s->conKey.p, s->conColor.p,
			                        s->best.p, s->mask.p, st);
		compact();
		pxh::launch_static_colors(n, C, s->conA.p, s->conB.p, s->conColor.p, s->bodyNext.p, st);
		// colour order: stable radix sort of (colour, index); colours stay below
		// 64 + rounds + planes per world
		const int colorBound = 64 + rounds + s->maxPlanes + 1;
		PXH_CHECK(s->colorCount.ensure(size_t(colorBound) + 2));
		pxh::launch_color_keys(C, s->conColor.p, uint32_t(colorBound), s->keys.p, s->vals.p, st);
		bool alt = false;
		pxh::launch_radix_sort(s->keys.p, s->vals.p, s->keys2.p, s->vals2.p, C, 32,
		                       32 + ((bitsFor(uint64_t(colorBound)) + 7) / 8) * 8, s->hist.p, s->scanTmp.p, st, alt);
		const uint64_t* ck = alt ? s->keys2.p : s->keys.p;
		const uint32_t* cv = alt ? s->vals2.p : s->vals.p;
		pxh::launch_row_start(C, ck, s->colorCount.p, uint32_t(colorBound), st);
		PXH_CHECK(hipMemcpyAsync(s->order.p, cv, sizeof(int) * C, hipMemcpyDeviceToDevice, st));
		PXH_CHECK(hipMemcpyAsync(s->hostCounts, s->colorCount.p, sizeof(uint32_t) * (size_t(colorBound) + 1),
		                         hipMemcpyDeviceToHost, st));
		PXH_CHECK(hipMemcpyAsync(s->hostCounts + colorBound + 1, mDev, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
		PXH_CHECK(hipStreamSynchronize(st));
		remaining = int(s->hostCounts[colorBound + 1]);
		if (remaining == 0)
		{
			colorStart.assign(s->hostCounts, s->hostCounts + colorBound + 1);
			nColors = colorBound;
			while (nColors > 0 && colorStart[nColors] == colorStart[nColors - 1])
				nColors--;
			colorStart.resize(size_t(nColors) + 1);
			break;
		}
		if (rounds >= kMaxRounds)
			return fail(PXH_ERR_CAPACITY, "pxh_scene_simulate: colouring did not finish in 960 rounds");
		bound = remaining;
	}
	s->statColors = nColors;
	s->statRounds = rounds;

	// 6. solver bodies + contact rows
	PXH_CHECK(s->sbVel.ensure(size_t(4) * n)); PXH_CHECK(s->sbP.ensure(n)); PXH_CHECK(s->sbDQ.ensure(n));
	PXH_CHECK(s->sbI0.ensure(n)); PXH_CHECK(s->sbI1.ensure(n)); PXH_CHECK(s->sbOrig.ensure(n));
	PXH_CHECK(s->sbRec.ensure(size_t(pxh::PREP_REC) * n));
	pxh::SolverBodies sb{s->sbVel.p, s->sbP.p, s->sbDQ.p, s->sbI0.p, s->sbI1.p, s->sbOrig.p, s->sbRec.p};
	const float3 gdt = make_float3(d.gravity[0] * dt, d.gravity[1] * dt, d.gravity[2] * dt);
	const bool pgs = d.solver_type == PXH_SOLVER_PGS;
	if (d.enable_body_accelerations)
	{
		PXH_CHECK(s->prevLin.ensure(n));
		PXH_CHECK(s->prevAng.ensure(n));
	}
	pxh::launch_preintegrate(n, s->pos.p, s->rot.p, s->lin.p, s->ang.p, s->invI.p, gdt, d.linear_damping * dt,
	                         d.angular_damping * dt, d.max_linear_velocity_sq, d.max_angular_velocity_sq, pgs, sb, dt,
	                         s->extPending ? s->extLin.p : nullptr, s->extPending ? s->extAng.p : nullptr,
	                         d.enable_body_accelerations ? s->prevLin.p : nullptr,
	                         d.enable_body_accelerations ? s->prevAng.p : nullptr, sleeping ? s->slFlags.p : nullptr, st);
	s->extPending = false;
	s->lastDt = dt;

	PXH_CHECK(s->rhI.ensure(Cc)); PXH_CHECK(s->rhF.ensure(Cc));
	PXH_CHECK(s->rnA.ensure(4 * Cc)); PXH_CHECK(s->rnB.ensure(4 * Cc)); PXH_CHECK(s->rnF.ensure(4 * Cc));
	PXH_CHECK(s->rfT.ensure(4 * Cc)); PXH_CHECK(s->rfA.ensure(4 * Cc)); PXH_CHECK(s->rfB.ensure(4 * Cc));
	PXH_CHECK(s->rfF.ensure(4 * Cc));
	pxh::Rows rows{s->rhI.p, s->rhF.p, s->rnA.p, s->rnB.p, s->rnF.p, s->rfT.p, s->rfA.p, s->rfB.p, s->rfF.p};
	const int P_ = d.position_iterations, V_ = d.velocity_iterations;
	const float stepDt = dt / float(P_);
	const float invStepDt = 1.0f / stepDt;
	const float biasCoefficient = 2.0f * std::sqrt(1.0f / float(P_));
	const float scale = std::min(0.8f, biasCoefficient);
	// joint rows from the poses at the step's start (DyTGSDynamics.cpp:
	// 1180-1256; erp = half the bias coefficient, DyTGSContactPrep.cpp:2018)
	const int nJ = int(s->hJType.size());
	if (nJ > 0)
		pxh::launch_joint_prep(nJ, s->jInfo.p, s->jFrame.p, s->jParam.p, s->pos.p, s->rot.p, sb, stepDt, invStepDt,
		                       1.0f / dt,
		                       0.5f * biasCoefficient, d.tolerance_length, s->jHead.p, s->jRows.p,
		                       sleeping ? s->slFlags.p : nullptr, st);
	// four step counters, then four per-block partial sums of the preparation
	PXH_CHECK(s->counters.ensure(4 + 4 * ((Cc + 255) / 256)));
	// friction patches and colours of this step's rows go to the [cur] buffers
	PXH_CHECK(s->patch[cur].ensure(6 * Cc));
	PXH_CHECK(s->patchKey[cur].ensure(Cc));
	PXH_CHECK(s->patchBroken[cur].ensure(Cc));
	PXH_CHECK(s->patchColor[cur].ensure(Cc));
	pio.cur = s->patch[cur].p;
	pio.curKey = s->patchKey[cur].p;
	pio.curColor = s->patchColor[cur].p;
	pio.conColor = s->conColor.p;
	pk = s->profBegin(4);
	if (pgs)
	{
		PXH_CHECK(s->rnU.ensure(4 * Cc));
		pxh::launch_prep_pgs(C, s->order.p, s->conA.p, s->conB.p, s->conSlot.p, s->ctN.p, s->ctP.p, s->pos.p,
		                     s->rot.p, sb, 1.0f / dt, d.static_friction, d.dynamic_friction, 2.0f * d.rest_offset,
		                     d.friction_offset_threshold, rows, s->rnU.p, pio, s->counters.p, st);
	}
	else
		pxh::launch_prep(C, s->order.p, s->conA.p, s->conB.p, s->conSlot.p, s->ctN.p, s->ctP.p, s->pos.p, s->rot.p,
		                 sb, invStepDt, scale, 2.0f * d.rest_offset, d.friction_offset_threshold, rows, pio,
		                 s->counters.p, st);
	PXH_CHECK(hipMemcpyAsync(s->hostCounts + kHostCounterSlot, s->counters.p, 4 * sizeof(unsigned long long),
	                         hipMemcpyDeviceToHost, st));
	s->profEnd(pk);

	// 7. solve. PGS (solveV_Blocks, lowleveldynamics/src/DySolverControl.cpp:163-380):
	//    position passes by colour, friction in the last three only (no
	//    friction-every-iteration flag), then the velocity changes are kept
	//    for the pose update and the velocity passes run on the concluded rows.
	//    TGS: position iterations with substep integration, then velocity
	//    iterations (DyTGSDynamics.cpp:2515-2790).
	//    PGS always runs at least one velocity pass (DySolverControl.cpp:343).
	if (s->tuneTimed)
		PXH_CHECK(hipEventRecord(s->tuneEv[0], st));
	s->statSolveLaunches = 0;
	const int pgsPasses = P_ + std::max(V_, 1);
	for (int it = 0; pgs && it < pgsPasses; it++)
	{
		const bool posIter = it < P_;
		const bool doFriction = !posIter || P_ - it <= 3;
		for (int c = 0; c < nColors; c++)
		{
			if (colorStart[c + 1] == colorStart[c])
				continue;
			hipEvent_t e0, e1;
			s->profKernel(5, &e0, &e1);
			s->statSolveLaunches++;
			pxh::launch_solve_pgs(int(colorStart[c]), int(colorStart[c + 1]), C, rows, s->rnU.p, sb, doFriction,
			                      !posIter, it == pgsPasses - 1 ? s->patchBroken[cur].p : nullptr, st, e0, e1);
		}
		if (it == P_ - 1)
			pxh::launch_pgs_save_motion(n, sb, st);
	}
	if (!pgs && worldSolve)
	{
		int grid = s->tuning[1];
		if (grid <= 0)
		{
			int dev = 0, cus = 256;
			(void)hipGetDevice(&dev);
			(void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
			grid = 4 * cus;
		}
		grid = std::min(grid, nWorlds);
		hipEvent_t e0, e1;
		s->profKernel(5, &e0, &e1);
		s->statSolveLaunches++;
		pxh::launch_solve_worlds(grid, nWorlds, s->worldOrder.p, s->solveQueue.p, s->bodyStart.p, s->segStart.p, CB,
		                         C, rows, sb, P_, V_, stepDt, d.static_friction, d.dynamic_friction, invStepDt,
		                         invStepDt * scale, scale, s->patchBroken[cur].p, s->tuning[2], st, e0, e1);
	}
	for (int it = 0; !pgs && !worldSolve && it < P_ + V_; it++)
	{
		const bool posIter = it < P_;
		// joints by colour, then contacts by colour; the last position
		// iteration concludes the joint rows (conclude1DStep)
		const float elapsed = stepDt * float(std::min(it, P_));
		for (int c = 0; c + 1 < int(s->jColorStart.size()); c++)
			pxh::launch_joint_solve(s->jColorStart[c], s->jColorStart[c + 1], s->jInfo.p, s->jHead.p, s->jRows.p, sb,
			                        elapsed, it == P_ - 1, sleeping ? s->slFlags.p : nullptr, st);
		for (int c = 0; c < nColors; c++)
		{
			if (colorStart[c + 1] == colorStart[c])
				continue;
			hipEvent_t e0, e1;
			s->profKernel(5, &e0, &e1);
			s->statSolveLaunches++;
			pxh::launch_solve(int(colorStart[c]), int(colorStart[c + 1]), C, rows, sb,
			                  posIter ? -3.402823466e+38f : 0.0f, d.static_friction, d.dynamic_friction, invStepDt,
			                  invStepDt * scale, scale, it == P_ + V_ - 1 ? s->patchBroken[cur].p : nullptr, st, e0,
			                  e1);
		}
		if (posIter)
		{
			pk = s->profBegin(6);
			pxh::launch_integrate(n, stepDt, sb, st);
			s->profEnd(pk);
		}
	}

	if (s->tuneTimed)
		PXH_CHECK(hipEventRecord(s->tuneEv[1], st));
	s->nPrevRows = C;
	s->patchCur = cur ^ 1;

	// 8. write back poses and velocities
	//    (with the sleep check of every awake body)
	if (pgs)
		pxh::launch_pgs_finalize(n, dt, s->pos.p, s->rot.p, s->lin.p, s->ang.p, s->invI.p, sb, s->sleepState(), st);
	else
		pxh::launch_finalize(n, dt, s->pos.p, s->rot.p, s->lin.p, s->ang.p, s->invI.p, sb, s->sleepState(), st);
	PXH_CHECK(hipGetLastError());
	s->simulating = true;
	return PXH_OK;
}

int32_t pxh_scene_fetch_results(pxh_scene s, int32_t block)
{
	if (!s)
		return fail(PXH_ERR_INVALID, "pxh_scene_fetch_results: null scene");
	if (!s->simulating)
		return PXH_OK;
	if (!block)
	{
		hipError_t q = hipStreamQuery(s->stream);
		if (q == hipErrorNotReady)
			return 1;
		if (q != hipSuccess)
			return fail(PXH_ERR_DEVICE, hipGetErrorString(q));
	}
	else
		PXH_CHECK(hipStreamSynchronize(s->stream));
	s->simulating = false;
	if (s->tuneTimed)
	{
		s->tuneTimed = false;
		float ms = 0.0f;
		PXH_CHECK(hipEventElapsedTime(&ms, s->tuneEv[0], s->tuneEv[1]));
		pxh_scene_t::TuneCand& c = s->tuneCands[s->tuneIdx];
		c.ms += ms;
		++c.steps;
		// a candidate whose first step is more than 1.25 times the best mean so
		// far is not tried again (each tried step is a real, timed step)
		double bestMean = 0.0;
		for (size_t k = 0; k < s->tuneIdx; k++)
			if (bestMean == 0.0 || s->tuneCands[k].ms / s->tuneCands[k].steps < bestMean)
				bestMean = s->tuneCands[k].ms / s->tuneCands[k].steps;
		const bool slow = c.steps == 1 && bestMean > 0.0 && ms > 1.25 * bestMean;
		if ((c.steps >= s->tuning[3] || slow) && ++s->tuneIdx == s->tuneCands.size())
		{
			size_t best = 0;
			for (size_t k = 1; k < s->tuneCands.size(); k++)
				if (s->tuneCands[k].ms / s->tuneCands[k].steps < s->tuneCands[best].ms / s->tuneCands[best].steps)
					best = k;
			for (int i = 0; i < 3; i++)
				s->tuning[i] = s->tuneCands[best].t[i];
			s->tuning[3] = 0;
			s->tuneLog.swap(s->tuneCands);
			s->tuneCands.clear();
		}
	}
	// event pairs are read when the profile is asked for (reading each pair
	// costs host time between steps); the pool is drained past 64k pairs
	if (s->profUsed > 65536)
		s->profCollect();
	if (s->statConstraints > 0)
	{
		unsigned long long c3[4];
		std::memcpy(c3, s->hostCounts + kHostCounterSlot, sizeof(c3));
		s->statContacts = int64_t(c3[0]);
		s->statFriction = int64_t(c3[1]);
		s->statDynCon = int64_t(c3[2]);
		s->statDynFriction = int64_t(c3[3]);
	}
	return PXH_OK;
}

int32_t pxh_scene_set_tuning(pxh_scene s, const int32_t* t, int32_t n)
{
	if (!s || n < 0 || (n > 0 && !t))
		return fail(PXH_ERR_INVALID, "pxh_scene_set_tuning: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_set_tuning: not allowed while simulating");
	if (n > 0 && (t[0] < PXH_SOLVE_LAYOUT_AUTO || t[0] > PXH_SOLVE_LAYOUT_WORLDS))
		return fail(PXH_ERR_INVALID, "pxh_scene_set_tuning: unknown solve layout");
	if (n > 1 && t[1] < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_set_tuning: negative workgroup count");
	if (n > 2 && (t[2] < 0 || t[2] > 1000))
		return fail(PXH_ERR_INVALID, "pxh_scene_set_tuning: LDS bodies out of [0, 1000]");
	if (n > 3 && t[3] < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_set_tuning: negative autotuning steps");
	for (int i = 0; i < std::min(n, 4); i++)
		s->tuning[i] = t[i];
	s->tuneCands.clear();
	return PXH_OK;
}

int32_t pxh_scene_get_tuning_log(pxh_scene s, int32_t* shapes, double* ms, int32_t max)
{
	if (!s || max < 0 || (max > 0 && (!shapes || !ms)))
		return fail(PXH_ERR_INVALID, "pxh_scene_get_tuning_log: bad arguments");
	const int k = std::min<int>(max, int(s->tuneLog.size()));
	for (int i = 0; i < k; i++)
	{
		const pxh_scene_t::TuneCand& c = s->tuneLog[i];
		for (int j = 0; j < 3; j++)
			shapes[3 * i + j] = c.t[j];
		ms[i] = c.steps ? c.ms / c.steps : 0.0;
	}
	return int32_t(s->tuneLog.size());
}

int32_t pxh_scene_get_tuning(pxh_scene s, int32_t* t, int32_t n)
{
	if (!s || n < 0 || (n > 0 && !t))
		return fail(PXH_ERR_INVALID, "pxh_scene_get_tuning: bad arguments");
	const int k = std::min(n, 4);
	for (int i = 0; i < k; i++)
		t[i] = s->tuning[i];
	return k;
}

int32_t pxh_scene_set_profiling(pxh_scene s, int32_t enable)
{
	if (!s)
		return fail(PXH_ERR_INVALID, "pxh_scene_set_profiling: null scene");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_set_profiling: not allowed while simulating");
	s->profCollect();
	s->profiling = uint32_t(enable) & 0xffu;
	s->profSeq = 0;
	s->profStep = 0;
	for (int i = 0; i < 8; i++)
	{
		s->profMs[i] = 0.0;
		s->profLaunches[i] = 0;
	}
	return PXH_OK;
}

int32_t pxh_scene_set_profiling_stride(pxh_scene s, int32_t stride)
{
	if (!s || stride < 1)
		return fail(PXH_ERR_INVALID, "pxh_scene_set_profiling_stride: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_set_profiling_stride: not allowed while simulating");
	s->profStride = uint32_t(stride);
	s->profSeq = 0;
	s->profStep = 0;
	return PXH_OK;
}

int32_t pxh_scene_get_profile(pxh_scene s, double* ms, int64_t* launches, int32_t n)
{
	if (!s || n < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_get_profile: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_get_profile: not allowed while simulating");
	s->profCollect();
	const int k = std::min(n, 8);
	for (int i = 0; i < k; i++)
	{
		if (ms) ms[i] = s->profMs[i];
		if (launches) launches[i] = s->profLaunches[i];
	}
	return k;
}

int32_t pxh_scene_get_rigid_dynamic_data(pxh_scene s, void* data, const uint32_t* index, int32_t kind, int32_t n,
                                         void* stream)
{
	if (!s || !data || n < 0 || kind < 0 || kind > 4)
		return fail(PXH_ERR_INVALID, "pxh_scene_get_rigid_dynamic_data: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_get_rigid_dynamic_data: not allowed while simulating");
	if (kind >= PXH_READ_LINEAR_ACCELERATION && !s->desc.enable_body_accelerations)
		return fail(PXH_ERR_STATE, "pxh_scene_get_rigid_dynamic_data: accelerations need enable_body_accelerations");
	if (kind >= PXH_READ_LINEAR_ACCELERATION && !(s->lastDt > 0.0f && s->prevLin.p))
		return fail(PXH_ERR_STATE, "pxh_scene_get_rigid_dynamic_data: no step since the bodies changed");
	if (!index && n > s->n)
		return fail(PXH_ERR_INVALID, "pxh_scene_get_rigid_dynamic_data: n exceeds body count");
	hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
	if (stream)
	{
		// order after the scene's stream
		hipEvent_t ev;
		PXH_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		PXH_CHECK(hipEventRecord(ev, s->stream));
		PXH_CHECK(hipStreamWaitEvent(st, ev, 0));
		PXH_CHECK(hipEventDestroy(ev));
	}
	pxh::launch_get_data(n, index, kind, s->pos.p, s->rot.p, s->lin.p, s->ang.p, s->prevLin.p, s->prevAng.p,
	                     s->lastDt > 0.0f ? 1.0f / s->lastDt : 0.0f, static_cast<float*>(data), st);
	PXH_CHECK(hipGetLastError());
	if (!stream)
		PXH_CHECK(hipStreamSynchronize(st));
	return PXH_OK;
}

int32_t pxh_scene_set_rigid_dynamic_data(pxh_scene s, const void* data, const uint32_t* index, int32_t kind, int32_t n,
                                         void* stream)
{
	if (!s || !data || n < 0 || kind < 0 || kind > 4)
		return fail(PXH_ERR_INVALID, "pxh_scene_set_rigid_dynamic_data: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_set_rigid_dynamic_data: not allowed while simulating");
	if (!index && n > s->n)
		return fail(PXH_ERR_INVALID, "pxh_scene_set_rigid_dynamic_data: n exceeds body count");
	if (stream)
	{
		// the scene's stream waits for the caller's data
		hipEvent_t ev;
		PXH_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		PXH_CHECK(hipEventRecord(ev, static_cast<hipStream_t>(stream)));
		PXH_CHECK(hipStreamWaitEvent(s->stream, ev, 0));
		PXH_CHECK(hipEventDestroy(ev));
	}
	if (kind >= PXH_WRITE_FORCE && !s->extLin.p && s->n > 0)
	{
		PXH_CHECK(s->extLin.ensure(size_t(s->n)));
		PXH_CHECK(s->extAng.ensure(size_t(s->n)));
		PXH_CHECK(hipMemsetAsync(s->extLin.p, 0, sizeof(float4) * s->extLin.cap, s->stream));
		PXH_CHECK(hipMemsetAsync(s->extAng.p, 0, sizeof(float4) * s->extAng.cap, s->stream));
	}
	pxh::launch_set_data(n, index, kind, s->pos.p, s->rot.p, s->lin.p, s->ang.p, s->invI.p, s->extLin.p, s->extAng.p,
	                     static_cast<const float*>(data), s->sleepState(), s->stream);
	if (kind >= PXH_WRITE_FORCE)
		s->extPending = true;
	PXH_CHECK(hipGetLastError());
	PXH_CHECK(hipStreamSynchronize(s->stream));
	return PXH_OK;
}

int32_t pxh_scene_get_touch_events(pxh_scene s, int32_t* found, int64_t max_found, int32_t* lost, int64_t max_lost,
                                   int64_t* counts)
{
	if (!s || !counts || max_found < 0 || max_lost < 0 || (max_found > 0 && !found) || (max_lost > 0 && !lost))
		return fail(PXH_ERR_INVALID, "pxh_scene_get_touch_events: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_get_touch_events: not allowed while simulating");
	if (!s->desc.report_touch_events)
		return fail(PXH_ERR_STATE, "pxh_scene_get_touch_events: the scene does not report touch events");
	counts[0] = counts[1] = 0;
	if (s->evC < 0)
		return PXH_OK;
	uint32_t nf = 0, nl = 0;
	PXH_CHECK(hipMemcpy(&nf, s->evFoundScan.p + s->evC, sizeof(uint32_t), hipMemcpyDeviceToHost));
	PXH_CHECK(hipMemcpy(&nl, s->evLostScan.p + s->evR, sizeof(uint32_t), hipMemcpyDeviceToHost));
	std::vector<uint2> fk(nf), lk(nl);
	if (nf)
		PXH_CHECK(hipMemcpy(fk.data(), s->evFoundKeys.p, sizeof(uint2) * nf, hipMemcpyDeviceToHost));
	if (nl)
		PXH_CHECK(hipMemcpy(lk.data(), s->evLostKeys.p, sizeof(uint2) * nl, hipMemcpyDeviceToHost));
	// world-relative keys (world tag 16 | local a 24 | local b or plane 24)
	// to body indices of the current worlds; pairs of removed worlds are dropped
	std::vector<int> worldOfTag(65536, -1);
	for (size_t w = 0; w < s->hWorldTag.size(); w++)
		worldOfTag[size_t(s->hWorldTag[w]) & 0xFFFF] = int(w);
	const int nBW = int(s->hBodyStart.size()) - 1;
	auto decode = [&](const uint2& k, int32_t* ab) {
		const uint64_t key = (uint64_t(k.x) << 32) | k.y;
		const int w = worldOfTag[size_t(key >> 48)];
		if (w < 0 || w >= nBW)
			return false;
		const uint32_t la = uint32_t(key >> 24) & 0xFFFFFFu, lb = uint32_t(key) & 0xFFFFFFu;
		ab[0] = s->hBodyStart[w] + int(la);
		if (lb & 0x800000u)
		{
			const uint32_t pl = lb & 0x7FFFFFu;
			if (size_t(w) >= s->worldPlanes.size() || pl >= s->worldPlanes[w].size())
				return false;
			ab[1] = -1 - s->worldPlanes[w][pl];
		}
		else
			ab[1] = s->hBodyStart[w] + int(lb);
		return true;
	};
	int64_t kf = 0, kl = 0;
	for (const uint2& k : fk)
	{
		int32_t ab[2];
		if (!decode(k, ab))
			continue;
		if (kf < max_found)
		{
			found[2 * kf] = ab[0];
			found[2 * kf + 1] = ab[1];
		}
		kf++;
	}
	for (const uint2& k : lk)
	{
		int32_t ab[2];
		if (!decode(k, ab))
			continue;
		if (kl < max_lost)
		{
			lost[2 * kl] = ab[0];
			lost[2 * kl + 1] = ab[1];
		}
		kl++;
	}
	counts[0] = kf;
	counts[1] = kl;
	return PXH_OK;
}

int32_t pxh_scene_get_stats(pxh_scene s, int64_t* out, int32_t n)
{
	if (!s || !out || n < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_get_stats: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_get_stats: not allowed while simulating");
	const int64_t v[13] = {s->n,          s->statPairs, s->statPlanePairs, s->statConstraints, s->statContacts,
	                       s->statColors, s->statSteps, s->statFriction,   s->statDynCon,      s->statRounds,
	                       s->statDynFriction, s->deviceBytes(), s->statSolveLaunches};
	const int k = std::min(n, 13);
	for (int i = 0; i < k; i++)
		out[i] = v[i];
	return k;
}

int64_t pxh_scene_get_constraints(pxh_scene s, int32_t* out, int64_t maxc)
{
	if (!s || !out)
		return fail(PXH_ERR_INVALID, "pxh_scene_get_constraints: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_get_constraints: not allowed while simulating");
	const int64_t C = s->statConstraints;
	const int64_t k = std::min(C, maxc);
	if (k <= 0)
		return C;
	std::vector<int> a(k), b(k), col(k), slot(k);
	PXH_CHECK(hipMemcpy(a.data(), s->conA.p, k * sizeof(int), hipMemcpyDeviceToHost));
	PXH_CHECK(hipMemcpy(b.data(), s->conB.p, k * sizeof(int), hipMemcpyDeviceToHost));
	PXH_CHECK(hipMemcpy(col.data(), s->conColor.p, k * sizeof(int), hipMemcpyDeviceToHost));
	PXH_CHECK(hipMemcpy(slot.data(), s->conSlot.p, k * sizeof(int), hipMemcpyDeviceToHost));
	std::vector<float4> hn(size_t(s->statPairs + s->statPlanePairs) + 1);
	PXH_CHECK(hipMemcpy(hn.data(), s->ctN.p, (hn.size() - 1) * sizeof(float4), hipMemcpyDeviceToHost));
	for (int64_t i = 0; i < k; i++)
	{
		int cnt;
		std::memcpy(&cnt, &hn[slot[i]].w, 4);
		out[4 * i] = a[i];
		out[4 * i + 1] = b[i];
		out[4 * i + 2] = col[i];
		out[4 * i + 3] = cnt;
	}
	return C;
}

int32_t pxh_scene_copy_contact_data(pxh_scene s, void* data, uint32_t* nb_pairs, int32_t max_pairs, void* stream)
{
	if (!s || !data || max_pairs < 0)
		return fail(PXH_ERR_INVALID, "pxh_scene_copy_contact_data: bad arguments");
	if (s->simulating)
		return fail(PXH_ERR_STATE, "pxh_scene_copy_contact_data: not allowed while simulating");
	hipStream_t st = stream ? static_cast<hipStream_t>(stream) : s->stream;
	if (stream)
	{
		hipEvent_t ev;
		PXH_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		PXH_CHECK(hipEventRecord(ev, s->stream));
		PXH_CHECK(hipStreamWaitEvent(st, ev, 0));
		PXH_CHECK(hipEventDestroy(ev));
	}
	const int C = int(s->statConstraints);
	if (C > 0 && s->n > 0)
		pxh::launch_copy_contacts(C, max_pairs, s->order.p, s->conA.p, s->conB.p, s->conSlot.p, s->ctN.p, s->ctP.p,
		                          s->rnF.p, static_cast<float4*>(data), nb_pairs, st);
	else if (nb_pairs)
		PXH_CHECK(hipMemsetAsync(nb_pairs, 0, sizeof(uint32_t), st));
	if (!stream)
		PXH_CHECK(hipStreamSynchronize(st));
	return PXH_OK;
}

void* pxh_scene_get_stream(pxh_scene s) { return s ? static_cast<void*>(s->stream) : nullptr; }

// ---------------------------------------------------------------------------
// standalone broadphase
// ---------------------------------------------------------------------------
int32_t pxh_bp_create(int32_t capacity, pxh_broadphase* out)
{
	if (!out || capacity < 0)
		return fail(PXH_ERR_INVALID, "pxh_bp_create: bad arguments");
	pxh_broadphase_t* bp = new pxh_broadphase_t();
	if (hipStreamCreateWithFlags(&bp->stream, hipStreamNonBlocking) != hipSuccess ||
	    hipHostMalloc(reinterpret_cast<void**>(&bp->hostCounts), 4096, hipHostMallocDefault) != hipSuccess)
	{
		delete bp;
		return fail(PXH_ERR_DEVICE, "pxh_bp_create: stream or pinned buffer creation failed");
	}
	if (capacity > 0)
	{
		(void)bp->pkPrev.ensure(capacity);
		(void)bp->pkCur.ensure(capacity);
	}
	*out = bp;
	return PXH_OK;
}

int32_t pxh_bp_update(pxh_broadphase bp, const float* bounds, const uint32_t* groups, const float* distances,
                      int32_t n, void* stream)
{
	if (!bp || n < 0 || (n > 0 && !bounds))
		return fail(PXH_ERR_INVALID, "pxh_bp_update: bad arguments");
	hipStream_t st = bp->stream;
	if (stream)
	{
		hipEvent_t ev;
		PXH_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		PXH_CHECK(hipEventRecord(ev, static_cast<hipStream_t>(stream)));
		PXH_CHECK(hipStreamWaitEvent(st, ev, 0));
		PXH_CHECK(hipEventDestroy(ev));
	}
	return pxh::bp_update(bp, bounds, groups, distances, n, st, bp->hostCounts, true, nullptr);
}

int32_t pxh_bp_update_handles(pxh_broadphase bp, const float* bounds, const uint32_t* groups, const float* distances,
                              const uint32_t* ids, uint32_t id_bound, uint64_t filter_lut, int32_t n, void* stream)
{
	if (!bp || n < 0 || (n > 0 && (!bounds || !ids)) || (n > 0 && id_bound == 0))
		return fail(PXH_ERR_INVALID, "pxh_bp_update_handles: bad arguments");
	hipStream_t st = bp->stream;
	if (stream)
	{
		hipEvent_t ev;
		PXH_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
		PXH_CHECK(hipEventRecord(ev, static_cast<hipStream_t>(stream)));
		PXH_CHECK(hipStreamWaitEvent(st, ev, 0));
		PXH_CHECK(hipEventDestroy(ev));
	}
	return pxh::bp_update(bp, bounds, groups, distances, n, st, bp->hostCounts, true, nullptr, ids, id_bound,
	                      filter_lut);
}

int32_t pxh_bp_fetch_results(pxh_broadphase bp, const uint32_t** created, int32_t* n_created, const uint32_t** deleted,
                             int32_t* n_deleted)
{
	if (!bp)
		return fail(PXH_ERR_INVALID, "pxh_bp_fetch_results: null broadphase");
	if (created) *created = reinterpret_cast<const uint32_t*>(bp->created.p);
	if (n_created) *n_created = bp->nCreated;
	if (deleted) *deleted = reinterpret_cast<const uint32_t*>(bp->deleted.p);
	if (n_deleted) *n_deleted = bp->nDeleted;
	return PXH_OK;
}

int32_t pxh_bp_release(pxh_broadphase bp)
{
	if (!bp)
		return PXH_OK;
	(void)hipStreamSynchronize(bp->stream);
	bp->bmin.release(); bp->bmax.release(); bp->smin.release(); bp->smax.release(); bp->gridF.release();
	bp->gridI.release(); bp->scene.release(); bp->group.release(); bp->bodyStart.release();
	bp->rows.release(); bp->rowOff.release(); bp->rowStart.release(); bp->vals.release(); bp->vals2.release();
	bp->hist.release(); bp->scanTmp.release(); bp->gridAcc.release(); bp->counts.release(); bp->flags.release(); bp->flags2.release();
	bp->keys.release(); bp->keys2.release(); bp->pkPrev.release(); bp->pkCur.release(); bp->pkTmp.release();
	bp->pairs.release(); bp->created.release(); bp->deleted.release();
	if (bp->hostCounts)
		(void)hipHostFree(bp->hostCounts);
	if (bp->stream)
		(void)hipStreamDestroy(bp->stream);
	delete bp;
	return PXH_OK;
}

int32_t pxh_generate_contacts(int32_t n, const int32_t* type0, const float* geom0, const float* pose0,
                              const int32_t* type1, const float* geom1, const float* pose1, float contact_distance,
                              float tolerance_length, float* out, int32_t* counts, void* stream)
{
	if (n < 0 || (n > 0 && (!type0 || !geom0 || !pose0 || !type1 || !geom1 || !pose1 || !out || !counts)))
		return fail(PXH_ERR_INVALID, "pxh_generate_contacts: bad arguments");
	hipStream_t st = static_cast<hipStream_t>(stream);
	pxh::launch_np_immediate(n, type0, geom0, pose0, type1, geom1, pose1, contact_distance, tolerance_length, out,
	                         counts, st);
	PXH_CHECK(hipGetLastError());
	if (!stream)
		PXH_CHECK(hipStreamSynchronize(st));
	return PXH_OK;
}

}  // extern "C"
