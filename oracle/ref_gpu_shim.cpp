// Test harness (test infrastructure only) over the reference PhysX SDK built
// with its GPU support on and the HIP plug-in linked in static plug-in mode
// (oracle/ref.mk, target gpu). It drives the SDK's own standalone broadphase,
// PxCreateBroadPhase (immediatemode/src/NpImmediateMode.cpp:2083-2128,
// 2131-2150), either as eGPU (ImmGPUBP: PxvGetPhysXGpu -> PxCreatePhysXGpu ->
// createGpuBroadPhase, i.e. the plug-in over libpxhip) or as one of the CPU
// types (eABP, ePABP, eSAP), so the two can be compared on the same updates.
//
// C ABI for ctypes:
//   pxrefgpu_bp_create(type, discardKineKine, discardStaticKine) -> handle
//     (type: 0 eSAP, 1 eMBP, 2 eABP, 3 ePABP, 4 eGPU; NULL on failure)
//   pxrefgpu_bp_update(h, created, nc, updated, nu, removed, nr, bounds, groups,
//                      distances, capacity) -> 0
//   pxrefgpu_bp_results(h, created, maxc, deleted, maxd, counts[2]) -> 0
//   pxrefgpu_bp_release(h)
//   pxrefgpu_static_group / dynamic_group(id) / kinematic_group(id)
//   pxrefgpu_plugin_info(name, max) -> 1 when the GPU plug-in created a
//     context manager (name: its device name)
#include "PxPhysicsAPI.h"
#include "PxBroadPhase.h"
#include "PxImmediateMode.h"
#include "gpu/PxGpu.h"

#include <cstdint>
#include <cstring>
#include <vector>

using namespace physx;

namespace {

PxDefaultAllocator gAllocator;
PxDefaultErrorCallback gErrorCallback;
PxFoundation* gFoundation = NULL;
PxCudaContextManager* gCuda = NULL;

bool ensureFoundation()
{
	if (!gFoundation)
		gFoundation = PxCreateFoundation(PX_PHYSICS_VERSION, gAllocator, gErrorCallback);
	return gFoundation != NULL;
}

bool ensureCuda()
{
	if (!ensureFoundation())
		return false;
	if (!gCuda)
	{
		PxCudaContextManagerDesc desc;
		gCuda = PxCreateCudaContextManager(*gFoundation, desc);
		if (gCuda && !gCuda->contextIsValid())
		{
			gCuda->release();
			gCuda = NULL;
		}
	}
	return gCuda != NULL;
}

struct Harness
{
	PxBroadPhase* bp = NULL;
	std::vector<PxBroadPhasePair> created, deleted;
};

}  // namespace

extern "C" {

void* pxrefgpu_bp_create(int type, int discardKineKine, int discardStaticKine)
{
	if (!ensureFoundation())
		return NULL;
	PxBroadPhaseDesc desc(static_cast<PxBroadPhaseType::Enum>(type));
	desc.mDiscardKinematicVsKinematic = discardKineKine != 0;
	desc.mDiscardStaticVsKinematic = discardStaticKine != 0;
	if (desc.mType == PxBroadPhaseType::eGPU)
	{
		if (!ensureCuda())
			return NULL;
		desc.mContextManager = gCuda;
	}
	Harness* h = new Harness();
	h->bp = PxCreateBroadPhase(desc);
	if (!h->bp)
	{
		delete h;
		return NULL;
	}
	return h;
}

int pxrefgpu_bp_update(void* hv, const uint32_t* created, int nc, const uint32_t* updated, int nu,
                       const uint32_t* removed, int nr, const float* bounds, const uint32_t* groups,
                       const float* distances, int capacity)
{
	Harness* h = static_cast<Harness*>(hv);
	if (!h)
		return -1;
	const PxBroadPhaseUpdateData ud(created, PxU32(nc), updated, PxU32(nu), removed, PxU32(nr),
	                                reinterpret_cast<const PxBounds3*>(bounds), groups, distances, PxU32(capacity));
	h->bp->update(ud);
	PxBroadPhaseResults res;
	h->bp->fetchResults(res);
	h->created.assign(res.mCreatedPairs, res.mCreatedPairs + res.mNbCreatedPairs);
	h->deleted.assign(res.mDeletedPairs, res.mDeletedPairs + res.mNbDeletedPairs);
	return 0;
}

// created / deleted pairs (id0, id1) of the last update, up to max each;
// counts[0..1] receive the full counts
int pxrefgpu_bp_results(void* hv, uint32_t* created, int maxc, uint32_t* deleted, int maxd, int* counts)
{
	Harness* h = static_cast<Harness*>(hv);
	if (!h)
		return -1;
	counts[0] = int(h->created.size());
	counts[1] = int(h->deleted.size());
	for (int i = 0; i < maxc && i < counts[0]; i++)
	{
		created[2 * i] = h->created[i].mID0;
		created[2 * i + 1] = h->created[i].mID1;
	}
	for (int i = 0; i < maxd && i < counts[1]; i++)
	{
		deleted[2 * i] = h->deleted[i].mID0;
		deleted[2 * i + 1] = h->deleted[i].mID1;
	}
	return 0;
}

void pxrefgpu_bp_release(void* hv)
{
	Harness* h = static_cast<Harness*>(hv);
	if (!h)
		return;
	h->bp->release();
	delete h;
}

uint32_t pxrefgpu_static_group() { return PxGetBroadPhaseStaticFilterGroup(); }
uint32_t pxrefgpu_dynamic_group(uint32_t id) { return PxGetBroadPhaseDynamicFilterGroup(id); }
uint32_t pxrefgpu_kinematic_group(uint32_t id) { return PxGetBroadPhaseKinematicFilterGroup(id); }

int pxrefgpu_plugin_info(char* name, int max)
{
	if (!ensureCuda())
		return 0;
	if (name && max > 0)
	{
		std::strncpy(name, gCuda->getDeviceName(), size_t(max) - 1);
		name[max - 1] = 0;
	}
	return 1;
}

}  // extern "C"
