// PhysX GPU plug-in on HIP: the objects the reference SDK obtains from its
// GPU plug-in factory, implemented over libpxhip for an MI355X.
//
// Built against the PhysX 5.6.1 headers with PX_PHYSX_GPU_STATIC, the mode in
// which the SDK calls PxCreatePhysXGpu directly (lowlevel/api/src/
// px_globals.cpp:74-75) and expects the plug-in library to provide the public
// GPU entry points of include/gpu/PxGpu.h:70-144 as well as the plug-in's own
// (physxgpu/include/PxPhysXGpu.h:203-241).
//
//   PxCreatePhysXGpu            -> HipPhysXGpu (PxPhysXGpu.h:101-200)
//   PxCreateCudaContextManager  -> HipContextManager (PxCudaContextManager.h)
//                                  with HipContext (PxCudaContext.h:75-183)
//   createGpuMemoryManager      -> HipMemoryManager (PxsMemoryManager.h)
//   createGpuHeapMemoryAllocatorManager -> HipHeapManager (PxsHeapMemoryAllocator.h)
//   getGpuKernelWranglerManager -> HipWranglerManager (PxsKernelWrangler.h)
//   createGpuBroadPhase         -> HipBroadPhase (lowlevelaabb/include/BpBroadPhase.h:98-206)
//                                  over pxh_bp_update_handles
//
// The scene pipeline objects (AABB manager, narrowphase context, simulation
// controller, dynamics context) are not adapted: the whole rigid-body step
// is reached through pxh_scene_* instead. Their factory methods report
// eINTERNAL_ERROR and return NULL, as the reference does when the GPU pipeline
// cannot be created.
#include "PxPhysXGpu.h"
#include "BpBroadPhase.h"
#include "BpBroadPhaseUpdate.h"
#include "BpFiltering.h"
#include "PxsHeapMemoryAllocator.h"
#include "PxsKernelWrangler.h"
#include "PxsMemoryManager.h"
#include "cudamanager/PxCudaContext.h"
#include "cudamanager/PxCudaContextManager.h"
#include "foundation/PxAllocator.h"
#include "foundation/PxFoundation.h"
#include "gpu/PxPhysicsGpu.h"

#include "pxh_hip_rt.h"
#include "pxhip.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace physx;

namespace {

// CUresult codes (the CUDA driver API's, which pxh_hip_rt.h returns too):
// success, invalid value, out of memory, not ready, not supported, unknown
enum : PxU32 { kOk = 0, kInvalid = 1, kOutOfMemory = 2, kNotReady = 600, kNotSupported = 801, kUnknown = 999 };

PxCUresult res(int rtCode) { return PxCUresult(PxU32(rtCode)); }
void* hs(CUstream s) { return reinterpret_cast<void*>(s); }
void* he(CUevent e) { return reinterpret_cast<void*>(e); }
void* dp(CUdeviceptr p) { return reinterpret_cast<void*>(p); }

PxFoundation* gFoundation = NULL;
PxProfilerCallback* gProfiler = NULL;

void reportError(PxErrorCode::Enum code, const char* msg, int line)
{
	if (gFoundation)
		gFoundation->error(code, __FILE__, line, "%s", msg);
}

// PxCudaContext on the HIP runtime (PxCudaContext.h:75-183). Errors latch in
// lastError; abort mode drops further work, as the reference's does
// (cudamanager/src/CudaContextManager.cpp:1158-1163, 1480-1487).
class HipContext : public PxCudaContext
{
public:
	explicit HipContext(PxDeviceAllocatorCallback* alloc) : mAlloc(alloc) {}
	void release() override { delete this; }

	PxCUresult memAlloc(CUdeviceptr* dptr, size_t bytesize) override
	{
		if (mAbort)
			return PxCUresult(kUnknown);
		void* p = NULL;
		if (mAlloc)
		{
			if (!mAlloc->memAlloc(&p, bytesize))
				return latch(PxCUresult(kOutOfMemory));
		}
		else if (pxhrt::malloc(&p, bytesize) != pxhrt::kOk)
			return latch(PxCUresult(kOutOfMemory));
		*dptr = CUdeviceptr(p);
		return PxCUresult(kOk);
	}
	PxCUresult memFree(CUdeviceptr dptr) override
	{
		if (mAlloc)
			return mAlloc->memFree(dp(dptr)) ? PxCUresult(kOk) : latch(PxCUresult(kInvalid));
		return latch(res(pxhrt::free(dp(dptr))));
	}
	PxCUresult memHostAlloc(void** pp, size_t bytesize, unsigned int flags) override
	{
		(void)flags;   // portable | device-mapped: HIP pinned memory is both
		return latch(res(pxhrt::hostMalloc(pp, bytesize)));
	}
	PxCUresult memFreeHost(void* p) override { return latch(res(pxhrt::hostFree(p))); }
	PxCUresult memHostGetDevicePointer(CUdeviceptr* pdptr, void* p, unsigned int flags) override
	{
		void* d = NULL;
		const PxCUresult r = latch(res(pxhrt::hostDevicePointer(&d, p, flags)));
		*pdptr = CUdeviceptr(d);
		return r;
	}
	// kernels of this plug-in are linked HIP code: there are no CUDA modules
	PxCUresult moduleLoadDataEx(CUmodule*, const void*, unsigned int, PxCUjit_option*, void**) override
	{
		return PxCUresult(kNotSupported);
	}
	PxCUresult moduleGetFunction(CUfunction*, CUmodule, const char*) override { return PxCUresult(kNotSupported); }
	PxCUresult moduleUnload(CUmodule) override { return PxCUresult(kNotSupported); }
	PxCUresult streamCreate(CUstream* s, unsigned int flags) override
	{
		void* h = NULL;
		const PxCUresult r = latch(res(pxhrt::streamCreate(&h, flags != 0)));
		*s = reinterpret_cast<CUstream>(h);
		return r;
	}
	// every stream keeps the default priority (streams of different
	// priorities are not used on this hardware)
	PxCUresult streamCreateWithPriority(CUstream* s, unsigned int flags, int) override { return streamCreate(s, flags); }
	PxCUresult streamFlush(CUstream) override { return PxCUresult(kOk); }
	PxCUresult streamWaitEvent(CUstream s, CUevent e, unsigned int flags) override
	{
		return latch(res(pxhrt::streamWaitEvent(hs(s), he(e), flags)));
	}
	PxCUresult streamWaitEvent(CUstream s, CUevent e) override { return streamWaitEvent(s, e, 0); }
	PxCUresult streamDestroy(CUstream s) override { return latch(res(pxhrt::streamDestroy(hs(s)))); }
	PxCUresult streamSynchronize(CUstream s) override { return latch(res(pxhrt::streamSynchronize(hs(s)))); }
	PxCUresult eventCreate(CUevent* e, unsigned int flags) override
	{
		void* h = NULL;
		const PxCUresult r = latch(res(pxhrt::eventCreate(&h, flags)));
		*e = reinterpret_cast<CUevent>(h);
		return r;
	}
	PxCUresult eventRecord(CUevent e, CUstream s) override { return latch(res(pxhrt::eventRecord(he(e), hs(s)))); }
	PxCUresult eventQuery(CUevent e) override { return res(pxhrt::eventQuery(he(e))); }
	PxCUresult eventSynchronize(CUevent e) override { return latch(res(pxhrt::eventSynchronize(he(e)))); }
	PxCUresult eventDestroy(CUevent e) override { return latch(res(pxhrt::eventDestroy(he(e)))); }
	PxCUresult launchKernel(CUfunction, unsigned int, unsigned int, unsigned int, unsigned int, unsigned int,
	                        unsigned int, unsigned int, CUstream, PxCudaKernelParam*, size_t, void**, const char*,
	                        int) override
	{
		return PxCUresult(kNotSupported);
	}
	PxCUresult launchKernel(CUfunction, PxU32, PxU32, PxU32, PxU32, PxU32, PxU32, PxU32, CUstream, void**, void**,
	                        const char*, int) override
	{
		return PxCUresult(kNotSupported);
	}
	PxCUresult memcpyDtoH(void* dst, CUdeviceptr src, size_t n) override
	{
		return latch(res(pxhrt::memcpy(dst, dp(src), n, pxhrt::kDeviceToHost)));
	}
	PxCUresult memcpyDtoHAsync(void* dst, CUdeviceptr src, size_t n, CUstream s) override
	{
		return latch(res(pxhrt::memcpyAsync(dst, dp(src), n, pxhrt::kDeviceToHost, hs(s))));
	}
	PxCUresult memcpyHtoD(CUdeviceptr dst, const void* src, size_t n) override
	{
		return latch(res(pxhrt::memcpy(dp(dst), src, n, pxhrt::kHostToDevice)));
	}
	PxCUresult memcpyHtoDAsync(CUdeviceptr dst, const void* src, size_t n, CUstream s) override
	{
		return latch(res(pxhrt::memcpyAsync(dp(dst), src, n, pxhrt::kHostToDevice, hs(s))));
	}
	PxCUresult memcpyDtoDAsync(CUdeviceptr dst, CUdeviceptr src, size_t n, CUstream s) override
	{
		return latch(res(pxhrt::memcpyAsync(dp(dst), dp(src), n, pxhrt::kDeviceToDevice, hs(s))));
	}
	PxCUresult memcpyDtoD(CUdeviceptr dst, CUdeviceptr src, size_t n) override
	{
		return latch(res(pxhrt::memcpy(dp(dst), dp(src), n, pxhrt::kDeviceToDevice)));
	}
	PxCUresult memcpyPeerAsync(CUdeviceptr dst, CUcontext, CUdeviceptr src, CUcontext, size_t n, CUstream s) override
	{
		return latch(res(pxhrt::memcpyAsync(dp(dst), dp(src), n, pxhrt::kDefault, hs(s))));
	}
	PxCUresult memsetD32Async(CUdeviceptr dst, unsigned int v, size_t n, CUstream s) override
	{
		return latch(res(pxhrt::memsetD32(dp(dst), v, n, hs(s))));
	}
	PxCUresult memsetD8Async(CUdeviceptr dst, unsigned char v, size_t n, CUstream s) override
	{
		return latch(res(pxhrt::memsetD8(dp(dst), v, n, hs(s))));
	}
	PxCUresult memsetD32(CUdeviceptr dst, unsigned int v, size_t n) override
	{
		return latch(res(pxhrt::memsetD32(dp(dst), v, n, NULL)));
	}
	PxCUresult memsetD16(CUdeviceptr dst, unsigned short v, size_t n) override
	{
		return latch(res(pxhrt::memsetD16(dp(dst), v, n)));
	}
	PxCUresult memsetD8(CUdeviceptr dst, unsigned char v, size_t n) override
	{
		return latch(res(pxhrt::memsetD8(dp(dst), v, n, NULL)));
	}
	PxCUresult getLastError() override
	{
		const PxU32 e = mLast;
		mLast = kOk;
		return PxCUresult(e != kOk ? e : PxU32(pxhrt::lastError()));
	}
	void setAbortMode(bool abort) override { mAbort = abort; }
	bool isInAbortMode() override { return mAbort; }

private:
	PxCUresult latch(PxCUresult r)
	{
		if (PxU32(r) != kOk && mLast == kOk)
			mLast = PxU32(r);
		return r;
	}
	PxDeviceAllocatorCallback* mAlloc;
	PxU32 mLast = kOk;
	bool mAbort = false;
};

// PxCudaContextManager on the HIP runtime for one device. The context is the
// runtime's primary context of that device: acquiring it makes the device
// current on the calling thread. supportsArchSM30 answers "the GPU pipeline
// can run", which is what the SDK asks it (simulationcontroller/src/
// ScScene.cpp:778-790).
class HipContextManager : public PxCudaContextManager
{
public:
	HipContextManager(int device, PxDeviceAllocatorCallback* alloc) : mDevice(device), mContext(new HipContext(alloc))
	{
		mValid = pxhrt::setDevice(device) == pxhrt::kOk && pxhrt::deviceInfo(device, &mInfo) == pxhrt::kOk;
		(void)pxhrt::driverVersion(&mDriver);
	}
	~HipContextManager() override { mContext->release(); }

	CUdeviceptr getMappedDevicePtr(void* pinned) override
	{
		void* d = NULL;
		return pxhrt::hostDevicePointer(&d, pinned, 0) == pxhrt::kOk ? CUdeviceptr(d) : CUdeviceptr(0);
	}
	void acquireContext() override { (void)pxhrt::setDevice(mDevice); }
	bool tryAcquireContext() override { return pxhrt::setDevice(mDevice) == pxhrt::kOk; }
	void releaseContext() override {}
	CUcontext getContext() override { return reinterpret_cast<CUcontext>(this); }
	PxCudaContext* getCudaContext() override { return mContext; }
	bool contextIsValid() const override { return mValid; }
	bool supportsArchSM10() const override { return mValid; }
	bool supportsArchSM11() const override { return mValid; }
	bool supportsArchSM12() const override { return mValid; }
	bool supportsArchSM13() const override { return mValid; }
	bool supportsArchSM20() const override { return mValid; }
	bool supportsArchSM30() const override { return mValid; }
	bool supportsArchSM35() const override { return mValid; }
	bool supportsArchSM50() const override { return mValid; }
	bool supportsArchSM52() const override { return mValid; }
	bool supportsArchSM60() const override { return mValid; }
	bool isIntegrated() const override { return mInfo.integrated != 0; }
	bool canMapHostMemory() const override { return mInfo.canMapHostMemory != 0; }
	int getDriverVersion() const override { return mDriver; }
	size_t getDeviceTotalMemBytes() const override { return mInfo.totalMem; }
	int getMultiprocessorCount() const override { return mInfo.multiProcessors; }
	unsigned int getClockRate() const override { return unsigned(mInfo.clockKHz); }
	int getSharedMemPerBlock() const override { return int(mInfo.sharedPerBlock); }
	int getSharedMemPerMultiprocessor() const override { return int(mInfo.sharedPerMP); }
	unsigned int getMaxThreadsPerBlock() const override { return unsigned(mInfo.maxThreadsPerBlock); }
	const char* getDeviceName() const override { return mInfo.name; }
	CUdevice getDevice() const override { return mDevice; }
	void setUsingConcurrentStreams(bool v) override { mConcurrent = v; }
	bool getUsingConcurrentStreams() const override { return mConcurrent; }
	void getDeviceMemoryInfo(size_t& free, size_t& total) const override
	{
		free = total = 0;
		(void)pxhrt::memInfo(&free, &total);
	}
	CUmodule* getCuModules() override { return NULL; }
	void release() override { delete this; }

	void* allocDeviceBufferInternal(PxU64 n, const char*, PxI32) override
	{
		CUdeviceptr p = 0;
		return PxU32(mContext->memAlloc(&p, size_t(n))) == kOk ? dp(p) : NULL;
	}
	void* allocPinnedHostBufferInternal(PxU64 n, const char*, PxI32) override
	{
		void* p = NULL;
		return PxU32(mContext->memHostAlloc(&p, size_t(n), 0)) == kOk ? p : NULL;
	}
	void freeDeviceBufferInternal(void* p) override { (void)mContext->memFree(CUdeviceptr(p)); }
	void freePinnedHostBufferInternal(void* p) override { (void)mContext->memFreeHost(p); }
	void clearDeviceBufferAsyncInternal(void* p, PxU32 n, CUstream s, PxI32 v) override
	{
		(void)mContext->memsetD32Async(CUdeviceptr(p), unsigned(v), n / 4, s);
	}
	void copyDToHAsyncInternal(void* h, const void* d, PxU32 n, CUstream s) override
	{
		(void)mContext->memcpyDtoHAsync(h, CUdeviceptr(d), n, s);
	}
	void copyHToDAsyncInternal(void* d, const void* h, PxU32 n, CUstream s) override
	{
		(void)mContext->memcpyHtoDAsync(CUdeviceptr(d), h, n, s);
	}
	void copyDToDAsyncInternal(void* d, const void* src, PxU32 n, CUstream s) override
	{
		(void)mContext->memcpyDtoDAsync(CUdeviceptr(d), CUdeviceptr(src), n, s);
	}
	void copyDToHInternal(void* h, const void* d, PxU32 n) override { (void)mContext->memcpyDtoH(h, CUdeviceptr(d), n); }
	void copyHToDInternal(void* d, const void* h, PxU32 n) override { (void)mContext->memcpyHtoD(CUdeviceptr(d), h, n); }
	void memsetD8AsyncInternal(void* d, const PxU8& v, PxU32 n, CUstream s) override
	{
		(void)mContext->memsetD8Async(CUdeviceptr(d), v, n, s);
	}
	void memsetD32AsyncInternal(void* d, const PxU32& v, PxU32 n, CUstream s) override
	{
		(void)mContext->memsetD32Async(CUdeviceptr(d), v, n, s);
	}

private:
	int mDevice;
	HipContext* mContext;
	pxhrt::DeviceInfo mInfo{};
	int mDriver = 0;
	bool mValid = false;
	bool mConcurrent = true;
};

// Host (pinned) and device allocators of a context manager
// (PxsMemoryManager.h; the reference's is createPxgMemoryManager).
class PinnedAllocator : public PxVirtualAllocatorCallback
{
public:
	void* allocate(size_t size, int, const char*, int) override
	{
		void* p = NULL;
		return pxhrt::hostMalloc(&p, size) == pxhrt::kOk ? p : NULL;
	}
	void deallocate(void* p) override
	{
		if (p)
			(void)pxhrt::hostFree(p);
	}
};

class DeviceAllocator : public PxVirtualAllocatorCallback
{
public:
	void* allocate(size_t size, int, const char*, int) override
	{
		void* p = NULL;
		return pxhrt::malloc(&p, size) == pxhrt::kOk ? p : NULL;
	}
	void deallocate(void* p) override
	{
		if (p)
			(void)pxhrt::free(p);
	}
};

class HipMemoryManager : public PxsMemoryManager
{
public:
	PxVirtualAllocatorCallback* getHostMemoryAllocator() override { return &mHost; }
	PxVirtualAllocatorCallback* getDeviceMemoryAllocator() override { return &mDevice; }

private:
	PinnedAllocator mHost;
	DeviceAllocator mDevice;
};

// Device heap bookkeeping: libpxhip's broadphase holds its own buffers, so the
// heap reports no memory of its own.
class HipHeapManager : public PxsHeapMemoryAllocatorManager
{
public:
	HipHeapManager() { mMappedMemoryAllocators = NULL; }
	PxU64 getDeviceMemorySize() const override { return 0; }
	PxsHeapStats getDeviceHeapStats() const override { return PxsHeapStats(); }
	void flushDeferredDeallocs() override {}
};

// No CUDA kernels to wrangle: libpxhip's kernels are linked code.
class HipWranglerManager : public PxsKernelWranglerManager
{
public:
	explicit HipWranglerManager(PxCudaContextManager* cm)
	{
		mKernelWrangler = NULL;
		mCudaContextManager = cm;
	}
};

// BpFilter's 7 x 7 table as libpxhip's 8 x 8 bit mask (bit t0 * 8 + t1)
PxU64 lutBits(const Bp::BpFilter& f)
{
	const bool* lut = f.getLUT();
	PxU64 bits = 0;
	for (int a = 0; a < Bp::FilterType::COUNT; a++)
		for (int b = 0; b < Bp::FilterType::COUNT; b++)
			if (lut[a * Bp::FilterType::COUNT + b])
				bits |= 1ull << (a * 8 + b);
	return bits;
}

// Bp::BroadPhase over libpxhip's standalone broadphase. The SDK keeps every
// object's bounds, group and contact distance in arrays indexed by handle
// and passes the created / updated / removed handles each update
// (BpBroadPhaseUpdate.h). The live handles' data is uploaded and the
// overlaps are computed on the GPU in handle space, created and deleted being
// the differences from the previous update; pairs that disappear because an
// object was removed are not reported as deleted (BpBroadPhase.h:171-185).
class HipBroadPhase : public Bp::BroadPhase
{
public:
	explicit HipBroadPhase(PxU64 contextID) : mContextID(contextID) {}
	bool init()
	{
		if (pxh_bp_create(0, &mBp) != PXH_OK)
			return false;
		return pxhrt::streamCreate(&mStream, true) == pxhrt::kOk;
	}
	~HipBroadPhase() override
	{
		if (mBp)
			pxh_bp_release(mBp);
		for (void* p : {static_cast<void*>(dBounds), static_cast<void*>(dGroups), static_cast<void*>(dDist),
		                static_cast<void*>(dIds)})
			if (p)
				(void)pxhrt::free(p);
		if (mStream)
			(void)pxhrt::streamDestroy(mStream);
	}

	PxBroadPhaseType::Enum getType() const override { return PxBroadPhaseType::eGPU; }
	void release() override { PX_DELETE_THIS; }
	void preBroadPhase(const Bp::BroadPhaseUpdateData&) override {}
	void fetchBroadPhaseResults() override {}
	void freeBuffers() override {}
	void shiftOrigin(const PxVec3&, const PxBounds3*, const PxReal*) override {}
#if PX_CHECKED
	bool isValid(const Bp::BroadPhaseUpdateData&) const override { return true; }
#endif

	void update(PxcScratchAllocator*, const Bp::BroadPhaseUpdateData& ud, PxBaseTask*) override
	{
		mCreated.clear();
		mDeleted.clear();
		const PxU32 cap = ud.getCapacity();
		if (mLive.size() < cap)
			mLive.resize(cap, 0);
		std::vector<char> removedNow(mLive.size(), 0);
		for (PxU32 i = 0; i < ud.getNumRemovedHandles(); i++)
		{
			const PxU32 h = ud.getRemovedHandles()[i];
			if (h < mLive.size())
			{
				mLive[h] = 0;
				removedNow[h] = 1;
			}
		}
		for (PxU32 i = 0; i < ud.getNumCreatedHandles(); i++)
		{
			const PxU32 h = ud.getCreatedHandles()[i];
			if (h < mLive.size())
				mLive[h] = 1;
		}
		// the live handles' data, in handle order
		std::vector<PxU32> ids;
		for (PxU32 h = 0; h < PxU32(mLive.size()); h++)
			if (mLive[h])
				ids.push_back(h);
		const int n = int(ids.size());
		std::vector<float> b(static_cast<size_t>(6 * n)), d(static_cast<size_t>(n));
		std::vector<PxU32> g(static_cast<size_t>(n));
		const PxBounds3* aabb = ud.getAABBs();
		for (int k = 0; k < n; k++)
		{
			const PxBounds3& bb = aabb[ids[k]];
			const float v[6] = {bb.minimum.x, bb.minimum.y, bb.minimum.z, bb.maximum.x, bb.maximum.y, bb.maximum.z};
			std::memcpy(&b[size_t(6) * k], v, sizeof(v));
			d[k] = ud.getContactDistance()[ids[k]];
			g[k] = PxU32(ud.getGroups()[ids[k]]);
		}
		if (!reserve(n) ||
		    (n > 0 && (pxhrt::memcpyAsync(dBounds, b.data(), b.size() * 4, pxhrt::kHostToDevice, mStream) != pxhrt::kOk ||
		               pxhrt::memcpyAsync(dDist, d.data(), d.size() * 4, pxhrt::kHostToDevice, mStream) != pxhrt::kOk ||
		               pxhrt::memcpyAsync(dGroups, g.data(), g.size() * 4, pxhrt::kHostToDevice, mStream) != pxhrt::kOk ||
		               pxhrt::memcpyAsync(dIds, ids.data(), ids.size() * 4, pxhrt::kHostToDevice, mStream) != pxhrt::kOk)))
		{
			reportError(PxErrorCode::eOUT_OF_MEMORY, "GPU broadphase: staging the bounds failed", __LINE__);
			return;
		}
		const PxU32 idBound = std::max<PxU32>(PxU32(mLive.size()), 1u);
		if (pxh_bp_update_handles(mBp, dBounds, dGroups, dDist, dIds, idBound, lutBits(ud.getFilter()), n, mStream) !=
		    PXH_OK)
		{
			reportError(PxErrorCode::eINTERNAL_ERROR, pxh_last_error(), __LINE__);
			return;
		}
		const uint32_t *c = NULL, *x = NULL;
		int32_t nc = 0, nx = 0;
		pxh_bp_fetch_results(mBp, &c, &nc, &x, &nx);
		std::vector<PxU32> hc(size_t(2) * nc), hx(size_t(2) * nx);
		if ((nc > 0 && pxhrt::memcpy(hc.data(), c, hc.size() * 4, pxhrt::kDeviceToHost) != pxhrt::kOk) ||
		    (nx > 0 && pxhrt::memcpy(hx.data(), x, hx.size() * 4, pxhrt::kDeviceToHost) != pxhrt::kOk))
		{
			reportError(PxErrorCode::eINTERNAL_ERROR, "GPU broadphase: reading the pairs failed", __LINE__);
			return;
		}
		for (int k = 0; k < nc; k++)
			mCreated.push_back(Bp::BroadPhasePair(hc[2 * k], hc[2 * k + 1]));
		for (int k = 0; k < nx; k++)
		{
			const PxU32 a = hx[2 * k], bb = hx[2 * k + 1];
			if (!(a < removedNow.size() && removedNow[a]) && !(bb < removedNow.size() && removedNow[bb]))
				mDeleted.push_back(Bp::BroadPhasePair(a, bb));
		}
	}

	const Bp::BroadPhasePair* getCreatedPairs(PxU32& n) const override
	{
		n = PxU32(mCreated.size());
		return mCreated.empty() ? NULL : mCreated.data();
	}
	const Bp::BroadPhasePair* getDeletedPairs(PxU32& n) const override
	{
		n = PxU32(mDeleted.size());
		return mDeleted.empty() ? NULL : mDeleted.data();
	}

private:
	bool reserve(int n)
	{
		if (n <= mCap)
			return true;
		const int cap = std::max(n, 2 * mCap);
		for (void** p : {reinterpret_cast<void**>(&dBounds), reinterpret_cast<void**>(&dGroups),
		                 reinterpret_cast<void**>(&dDist), reinterpret_cast<void**>(&dIds)})
		{
			if (*p)
				(void)pxhrt::free(*p);
			*p = NULL;
		}
		mCap = 0;
		if (pxhrt::malloc(reinterpret_cast<void**>(&dBounds), sizeof(float) * 6 * size_t(cap)) != pxhrt::kOk ||
		    pxhrt::malloc(reinterpret_cast<void**>(&dGroups), sizeof(PxU32) * size_t(cap)) != pxhrt::kOk ||
		    pxhrt::malloc(reinterpret_cast<void**>(&dDist), sizeof(float) * size_t(cap)) != pxhrt::kOk ||
		    pxhrt::malloc(reinterpret_cast<void**>(&dIds), sizeof(PxU32) * size_t(cap)) != pxhrt::kOk)
			return false;
		mCap = cap;
		return true;
	}

	PxU64 mContextID;
	pxh_broadphase mBp = NULL;
	void* mStream = NULL;
	std::vector<char> mLive;
	std::vector<Bp::BroadPhasePair> mCreated, mDeleted;
	float* dBounds = NULL;
	PxU32* dGroups = NULL;
	float* dDist = NULL;
	PxU32* dIds = NULL;
	int mCap = 0;
};

// The plug-in factory (PxPhysXGpu.h:101-200; the reference's is PxgPhysXGpu,
// physxgpu/src/PxgPhysXGpu.cpp:44-250), a process singleton.
class HipPhysXGpu : public PxPhysXGpu, public PxUserAllocated
{
public:
	void release() override;

	PxsParticleBuffer* createParticleBuffer(PxU32, PxU32, PxCudaContextManager&) override { return NULL; }
	PxsParticleAndDiffuseBuffer* createParticleAndDiffuseBuffer(PxU32, PxU32, PxU32, PxCudaContextManager&) override
	{
		return NULL;
	}
	PxsParticleClothBuffer* createParticleClothBuffer(PxU32, PxU32, PxU32, PxU32, PxU32, PxCudaContextManager&) override
	{
		return NULL;
	}
	PxsParticleRigidBuffer* createParticleRigidBuffer(PxU32, PxU32, PxU32, PxCudaContextManager&) override
	{
		return NULL;
	}

	PxsMemoryManager* createGpuMemoryManager(PxCudaContextManager*) override { return PX_NEW(HipMemoryManager)(); }

	PxsHeapMemoryAllocatorManager* createGpuHeapMemoryAllocatorManager(const PxU32, PxsMemoryManager*,
	                                                                   const PxU32) override
	{
		return PX_NEW(HipHeapManager)();
	}

	PxsKernelWranglerManager* getGpuKernelWranglerManager(PxCudaContextManager* cm) override
	{
		if (!mWrangler)
			mWrangler = PX_NEW(HipWranglerManager)(cm);
		return mWrangler;
	}

	Bp::BroadPhase* createGpuBroadPhase(const PxGpuBroadPhaseDesc&, PxsKernelWranglerManager*, PxCudaContextManager* cm,
	                                    PxU32 gpuComputeVersion, const PxGpuDynamicsMemoryConfig&,
	                                    PxsHeapMemoryAllocatorManager*, PxU64 contextID) override
	{
		// the reference requires compute version 0 here (PxgPhysXGpu.cpp:157)
		if (gpuComputeVersion != 0 || !cm || !cm->contextIsValid())
			return NULL;
		cm->acquireContext();
		HipBroadPhase* bp = PX_NEW(HipBroadPhase)(contextID);
		if (!bp->init())
		{
			PX_DELETE(bp);
			reportError(PxErrorCode::eINTERNAL_ERROR, "GPU broadphase: creation failed", __LINE__);
			return NULL;
		}
		return bp;
	}

	Bp::AABBManagerBase* createGpuAABBManager(PxsKernelWranglerManager*, PxCudaContextManager*, const PxU32,
	                                          const PxGpuDynamicsMemoryConfig&, PxsHeapMemoryAllocatorManager*,
	                                          Bp::BroadPhase&, Bp::BoundsArray&, PxFloatArrayPinnedSafe&, PxU32, PxU32,
	                                          PxVirtualAllocator&, PxU64, PxPairFilteringMode::Enum,
	                                          PxPairFilteringMode::Enum) override
	{
		return notAdapted("createGpuAABBManager");
	}
	Bp::BoundsArray* createGpuBounds(PxVirtualAllocator&) override { return notAdapted("createGpuBounds"); }
	PxvNphaseImplementationContext* createGpuNphaseImplementationContext(
	    PxsContext&, PxsKernelWranglerManager*, PxvNphaseImplementationFallback*, const PxGpuDynamicsMemoryConfig&, void*,
	    void*, void*, PxBoundsArrayPinned&, IG::IslandSim*, physx::Dy::Context*, const PxU32,
	    PxsHeapMemoryAllocatorManager*, bool) override
	{
		return notAdapted("createGpuNphaseImplementationContext");
	}
	PxsSimulationController* createGpuSimulationController(PxsKernelWranglerManager*, PxCudaContextManager*,
	                                                       Dy::Context*, PxvNphaseImplementationContext*,
	                                                       Bp::BroadPhase*, bool, PxsSimulationControllerCallback*,
	                                                       PxU32, PxsHeapMemoryAllocatorManager*, PxU32, PxU32, PxU32,
	                                                       PxU32, bool) override
	{
		return notAdapted("createGpuSimulationController");
	}
	Dy::Context* createGpuDynamicsContext(Cm::FlushPool&, PxsKernelWranglerManager*, PxCudaContextManager*,
	                                      const PxGpuDynamicsMemoryConfig&, IG::SimpleIslandManager&, PxU32, PxU32, bool,
	                                      bool, bool, PxReal, PxU32, PxvSimStats&, PxsHeapMemoryAllocatorManager*, bool,
	                                      bool, PxSolverType::Enum, PxReal, bool, PxU64, bool) override
	{
		return notAdapted("createGpuDynamicsContext");
	}

private:
	struct Null
	{
		template <typename T>
		operator T*() const { return NULL; }
	};
	Null notAdapted(const char* what)
	{
		(void)what;
		reportError(PxErrorCode::eINTERNAL_ERROR,
		            "HIP GPU plug-in: the scene pipeline objects are not provided (step scenes with pxh_scene_*)",
		            __LINE__);
		return Null();
	}
	~HipPhysXGpu() override { PX_DELETE(mWrangler); }
	HipWranglerManager* mWrangler = NULL;
};

HipPhysXGpu* gInstance = NULL;

void HipPhysXGpu::release()
{
	if (gInstance == this)
		gInstance = NULL;
	PX_DELETE_THIS;
}

}  // namespace

// ---------------------------------------------------------------------------
// Exports: the plug-in's own (PxPhysXGpu.h:203-241) and, for the static
// plug-in mode, the SDK's public GPU entry points (include/gpu/PxGpu.h:70-144),
// which PxGpu.cpp forwards to the plug-in only in the dynamic mode.
// ---------------------------------------------------------------------------

PX_C_EXPORT PxPhysXGpu* PX_CALL_CONV PxCreatePhysXGpu()
{
	if (!gInstance)
		gInstance = PX_NEW(HipPhysXGpu)();
	return gInstance;
}

PX_C_EXPORT PxCudaContextManager* PX_CALL_CONV PxCreateCudaContextManager(PxFoundation& foundation,
                                                                         const PxCudaContextManagerDesc& desc,
                                                                         PxProfilerCallback* profilerCallback,
                                                                         bool launchSynchronous)
{
	(void)launchSynchronous;
	gFoundation = &foundation;
	if (profilerCallback)
		gProfiler = profilerCallback;
	int count = 0;
	if (pxhrt::deviceCount(&count) != pxhrt::kOk || count == 0)
	{
		foundation.error(PxErrorCode::eDEBUG_INFO, __FILE__, __LINE__, "no HIP device");
		return NULL;
	}
	const int device = desc.deviceOrdinal >= 0 && desc.deviceOrdinal < count ? desc.deviceOrdinal : 0;
	return new HipContextManager(device, desc.deviceAllocator);
}

PX_C_EXPORT void PX_CALL_CONV PxSetPhysXGpuProfilerCallback(PxProfilerCallback* cb) { gProfiler = cb; }

PX_C_EXPORT void PX_CALL_CONV PxSetPhysXGpuFoundationInstance(PxFoundation& foundation) { gFoundation = &foundation; }

PX_C_EXPORT int PX_CALL_CONV PxGetSuggestedCudaDeviceOrdinal(PxErrorCallback&)
{
	int count = 0;
	return pxhrt::deviceCount(&count) == pxhrt::kOk && count > 0 ? 0 : -1;
}

// no CUDA modules or kernels are registered: the HIP kernels are linked code
PX_C_EXPORT void PX_CALL_CONV PxGpuCudaRegisterFunction(int, const char*) {}
PX_C_EXPORT void** PX_CALL_CONV PxGpuCudaRegisterFatBinary(void*) { return NULL; }
PX_C_EXPORT PxKernelIndex* PX_CALL_CONV PxGpuGetCudaFunctionTable() { return NULL; }
PX_C_EXPORT PxU32 PX_CALL_CONV PxGpuGetCudaFunctionTableSize() { return 0; }
PX_C_EXPORT void** PX_CALL_CONV PxGpuGetCudaModuleTable() { return NULL; }
PX_C_EXPORT PxU32 PX_CALL_CONV PxGpuGetCudaModuleTableSize() { return 0; }
// the SDF / particle GPU utilities are outside this plug-in
PX_C_EXPORT PxPhysicsGpu* PX_CALL_CONV PxGpuCreatePhysicsGpu() { return NULL; }

PX_C_EXPORT void PX_CALL_CONV PxCudaRegisterFunction(int moduleIndex, const char* name)
{
	PxGpuCudaRegisterFunction(moduleIndex, name);
}
PX_C_EXPORT void** PX_CALL_CONV PxCudaRegisterFatBinary(void* fatBin) { return PxGpuCudaRegisterFatBinary(fatBin); }
PX_C_EXPORT void** PX_CALL_CONV PxGetCudaModuleTable() { return PxGpuGetCudaModuleTable(); }
PX_C_EXPORT PxU32 PX_CALL_CONV PxGetCudaModuleTableSize() { return PxGpuGetCudaModuleTableSize(); }
PX_C_EXPORT PxKernelIndex* PX_CALL_CONV PxGetCudaFunctionTable() { return PxGpuGetCudaFunctionTable(); }
PX_C_EXPORT PxU32 PX_CALL_CONV PxGetCudaFunctionTableSize() { return PxGpuGetCudaFunctionTableSize(); }
PX_C_EXPORT PxPhysicsGpu* PX_CALL_CONV PxGetPhysicsGpu() { return PxGpuCreatePhysicsGpu(); }
