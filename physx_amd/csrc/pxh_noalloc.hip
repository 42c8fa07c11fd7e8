// Allocation-free standalone broadphase update (pxh_bp_update_noalloc).
//
// The same launch sequence as pxh_bp_update (pxh::bp_update, the kernels and
// their order are shared, so the pair sets are the same bit for bit), with
// every intermediate buffer a fixed view into one file-scope device array,
// `scratch`, instead of growable hipMalloc buffers: for callers that may not
// allocate device memory. Nothing here calls hipMalloc, hipFree, hipHostMalloc
// or an allocating library routine; the caller passes inputs and outputs.
//
// Each GPU holds its own copy of `scratch` (PXH_BP_NOALLOC_SCRATCH_BYTES)
// until the process ends. Its device address is looked up once per device
// with hipGetSymbolAddress on the first call made with that device current;
// host code only ever passes that address to kernels and copies.
//
// The array keeps the sorted pair keys of the last successful update on its
// GPU, which is what created/deleted are diffed against: one process shares
// one broadphase state per GPU. Calls on one GPU are serialised by a host lock
// and complete before they return, so they cannot overlap within a process;
// two callers interleaving their updates on one GPU see each other's pairs as
// the previous update.
#include "pxh_bp_core.h"

#include <mutex>

namespace {

constexpr size_t kScratchBytes = PXH_BP_NOALLOC_SCRATCH_BYTES;
constexpr int kMaxBoxes = PXH_BP_NOALLOC_MAX_BOXES;
constexpr int kMaxPairs = PXH_BP_NOALLOC_MAX_PAIRS;
constexpr int kMaxDevices = 64;

// device memory beyond inputs and outputs: keys, sorted bounds, rows, pair
// buffers, scan and histogram space, and the previous update's pair keys
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// a fixed view into scratch: ensure() never grows
template <typename T>
struct FixedBuf
{
	T* p = nullptr;
	size_t cap = 0;
	hipError_t ensure(size_t n) const { return n <= cap ? hipSuccess : hipErrorOutOfMemory; }
};

// the context pxh::bp_update works on (same members as the growable one)
struct NoAllocBP
{
	int n = 0;
	FixedBuf<float4> bmin, bmax, smin, smax, gridF;
	FixedBuf<int2> gridI;
	FixedBuf<int> scene, group, bodyStart;
	FixedBuf<uint32_t> rows, rowOff, rowStart, vals, vals2, hist, scanTmp, counts, flags, flags2, gridAcc;
	FixedBuf<uint64_t> keys, keys2, pkPrev, pkCur, pkTmp;
	FixedBuf<uint2> pairs, created, deleted;
	uint64_t* sortedKeys = nullptr;
	size_t pairCap = 0;
	int nPrev = 0, nCur = 0, nCreated = 0, nDeleted = 0;
};

struct Carver
{
	unsigned char* base;
	size_t off = 0;
	template <typename T>
	void take(FixedBuf<T>& b, size_t count)
	{
		b.p = base ? reinterpret_cast<T*>(base + off) : nullptr;
		b.cap = count;
		off += (count * sizeof(T) + 255) & ~size_t(255);
	}
};

// Lays the context's buffers out from `base` (null: only measures); returns
// the bytes used. Sizes cover every ensure() of pxh::bp_update for up to
// kMaxBoxes boxes and kMaxPairs pairs.
size_t carve(unsigned char* base, NoAllocBP& c)
{
	const size_t N = kMaxBoxes, P = kMaxPairs;
	const size_t NP = std::max(N, P);
	const size_t tiles = (NP + size_t(pxh::radix_tile()) - 1) / size_t(pxh::radix_tile());
	const size_t scanN = std::max({size_t(pxh::PAIR_CLASSES) * N + 1, size_t(256) * 1024, 256 * tiles, P + 1});
	Carver k{base};
	k.take(c.bmin, N); k.take(c.bmax, N); k.take(c.smin, N); k.take(c.smax, N);
	k.take(c.gridF, 1); k.take(c.gridI, 1); k.take(c.gridAcc, 5);
	k.take(c.scene, N); k.take(c.group, N); k.take(c.bodyStart, 2);
	k.take(c.rowOff, 2); k.take(c.rowStart, 2 * N + 2);
	k.take(c.keys, N); k.take(c.keys2, N);
	k.take(c.vals, NP); k.take(c.vals2, NP);
	k.take(c.hist, 256 * tiles);
	k.take(c.scanTmp, size_t(pxh::scan_tmp_size(int64_t(scanN))) + 1024);
	k.take(c.counts, std::max(size_t(pxh::PAIR_CLASSES) * N + 1, P + 1));
	k.take(c.flags, P + 1); k.take(c.flags2, P + 1); k.take(c.rows, P + 1);
	k.take(c.pkPrev, P); k.take(c.pkCur, P); k.take(c.pkTmp, P);
	k.take(c.pairs, P); k.take(c.created, P); k.take(c.deleted, P);
	c.pairCap = P;
	return k.off;
}

struct DeviceState
{
	bool ready = false;
	unsigned char* ptr = nullptr;   // this GPU's scratch, from hipGetSymbolAddress
	NoAllocBP ctx;
	uint32_t hostCounts[4] = {};
};

std::mutex gLock[kMaxDevices];
DeviceState gState[kMaxDevices];

int fail(int code, const std::string& what)
{
	pxh::set_last_error(what);
	return code;
}

}  // namespace

extern "C" int32_t pxh_bp_update_noalloc(const float* bounds, const uint32_t* groups, const float* distances,
                                         int32_t n, uint32_t* created, int32_t max_created, int32_t* n_created,
                                         uint32_t* deleted, int32_t max_deleted, int32_t* n_deleted, void* stream)
{
	if (n < 0 || (n > 0 && !bounds) || !n_created || !n_deleted || max_created < 0 || max_deleted < 0 ||
	    (max_created > 0 && !created) || (max_deleted > 0 && !deleted))
		return fail(PXH_ERR_INVALID, "pxh_bp_update_noalloc: bad arguments");
	if (n > kMaxBoxes)
		return fail(PXH_ERR_CAPACITY, "pxh_bp_update_noalloc: " + std::to_string(n) +
		                                  " boxes exceed the maximum of " + std::to_string(kMaxBoxes) +
		                                  " (PXH_BP_NOALLOC_MAX_BOXES)");
	int dev = 0;
	PXH_CHECK(hipGetDevice(&dev));
	if (dev < 0 || dev >= kMaxDevices)
		return fail(PXH_ERR_UNSUPPORTED, "pxh_bp_update_noalloc: device ordinal out of range");
	std::lock_guard<std::mutex> guard(gLock[dev]);
	DeviceState& ds = gState[dev];
	if (!ds.ready)
	{
		// set-up on this device: the device address of its copy of scratch
		void* ptr = nullptr;
		PXH_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		if (!ptr)
			return fail(PXH_ERR_DEVICE, "pxh_bp_update_noalloc: no device address for the scratch array");
		NoAllocBP c;
		if (carve(nullptr, c) > kScratchBytes)
			return fail(PXH_ERR_CAPACITY, "pxh_bp_update_noalloc: buffer layout exceeds the scratch array");
		ds.ptr = static_cast<unsigned char*>(ptr);
		carve(ds.ptr, ds.ctx);
		ds.ready = true;
	}
	NoAllocBP& bp = ds.ctx;
	hipStream_t st = static_cast<hipStream_t>(stream);
	int required = 0;
	int r = pxh::bp_update(&bp, bounds, groups, distances, n, st, ds.hostCounts, false, &required);
	if (r == PXH_ERR_CAPACITY)
		return fail(PXH_ERR_CAPACITY, "pxh_bp_update_noalloc: " + std::to_string(required) +
		                                  " overlapping pairs exceed the maximum of " + std::to_string(kMaxPairs) +
		                                  " (PXH_BP_NOALLOC_MAX_PAIRS)");
	if (r)
		return r;
	if (bp.nCreated > max_created || bp.nDeleted > max_deleted)
	{
		// drop this update: the next one diffs against the last successful one
		const int nc = bp.nCreated, nd = bp.nDeleted;
		std::swap(bp.pkPrev, bp.pkCur);
		bp.nCur = bp.nPrev;
		bp.nCreated = bp.nDeleted = 0;
		*n_created = nc;
		*n_deleted = nd;
		return fail(PXH_ERR_CAPACITY, "pxh_bp_update_noalloc: " + std::to_string(nc) + " created / " +
		                                  std::to_string(nd) + " deleted pairs exceed the output capacities " +
		                                  std::to_string(max_created) + " / " + std::to_string(max_deleted));
	}
	if (bp.nCreated > 0)
		PXH_CHECK(hipMemcpyAsync(created, bp.created.p, sizeof(uint2) * size_t(bp.nCreated), hipMemcpyDeviceToDevice,
		                         st));
	if (bp.nDeleted > 0)
		PXH_CHECK(hipMemcpyAsync(deleted, bp.deleted.p, sizeof(uint2) * size_t(bp.nDeleted), hipMemcpyDeviceToDevice,
		                         st));
	PXH_CHECK(hipStreamSynchronize(st));
	*n_created = bp.nCreated;
	*n_deleted = bp.nDeleted;
	return PXH_OK;
}
