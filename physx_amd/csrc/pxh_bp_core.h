// Broadphase launch sequence shared by the scene step, the standalone
// broadphase (pxh_bp_*) and its allocation-free variant. The templates take a
// context whose buffers expose a device pointer `p` and `ensure(n)`: growable
// device buffers in the scene and the standalone broadphase, fixed views into
// one preallocated device array in the allocation-free entry point (where
// ensure fails instead of growing). Either way the kernels and their order are
// the same, so the pair sets are the same bit for bit.
#pragma once
#include "pxh_kernels.h"
#include "../../include/pxhip.h"

#include <algorithm>
#include <string>

namespace pxh {

// records the calling thread's last error (pxh_last_error)
void set_last_error(const std::string& what);

}  // namespace pxh

#define PXH_CHECK(call)                                                                         \
	do {                                                                                        \
		hipError_t e_ = (call);                                                                 \
		if (e_ != hipSuccess)                                                                   \
		{                                                                                       \
			pxh::set_last_error(std::string(#call) + ": " + hipGetErrorString(e_));             \
			return PXH_ERR_DEVICE;                                                              \
		}                                                                                       \
	} while (0)

namespace pxh {

inline int bitsFor(uint64_t v)
{
	int b = 0;
	while (b < 64 && (v >> b))
		b++;
	return b;
}

// k device counts into host[0..k), waiting for the stream
inline int readCounts(hipStream_t st, uint32_t* host, const uint32_t* const* src, int k)
{
	for (int i = 0; i < k; i++)
		PXH_CHECK(hipMemcpyAsync(host + i, src[i], sizeof(uint32_t), hipMemcpyDeviceToHost, st));
	PXH_CHECK(hipStreamSynchronize(st));
	return PXH_OK;
}

// the write pass of the sweep (again after the pair buffer has grown)
template <typename Ctx>
int bp_write_pairs(Ctx* s, int n, const int* d_scene, const int* d_group, const uint32_t* d_rowOff, hipStream_t st,
                   size_t pairCap, uint64_t groupLut = GROUP_LUT_ALL)
{
	PXH_CHECK(s->pairs.ensure(std::max<size_t>(pairCap, 1)));
	launch_sweep(true, n, s->sortedKeys, s->smin.p, s->smax.p, d_scene, d_group, groupLut, s->gridI.p, d_rowOff,
	             s->rowStart.p, nullptr, s->counts.p, s->pairs.p, uint32_t(pairCap), st);
	return PXH_OK;
}

// From inflated bounds (bmin/bmax, bmin.w = geometry type) of n boxes in
// nWorlds worlds whose row table d_rowOff (nWorlds + 1 entries, totalRows in
// all) is already on the device, produce the class-ordered pair list without a
// host round trip: the pair count stays on the device at
// counts[PAIR_CLASSES * n] (class c ends at counts[(c + 1) * n]) and at most
// pairCap pairs are written.
template <typename Ctx>
int broadphase(Ctx* s, int n, int nWorlds, const int* d_scene, const int* d_group, const uint32_t* d_rowOff,
               uint32_t totalRows, hipStream_t st, size_t pairCap, uint64_t groupLut = GROUP_LUT_ALL)
{
	PXH_CHECK(s->gridF.ensure(nWorlds));
	PXH_CHECK(s->gridI.ensure(nWorlds));
	PXH_CHECK(s->scanTmp.ensure(scan_tmp_size(std::max<int64_t>(int64_t(PAIR_CLASSES) * n + 1, 256 * 1024)) + 1024));
	PXH_CHECK(s->gridAcc.ensure(size_t(5) * nWorlds));
	launch_scene_grid(n, nWorlds, d_scene, s->bmin.p, s->bmax.p, s->gridAcc.p, s->gridF.p, s->gridI.p, d_rowOff, st);
	PXH_CHECK(s->keys.ensure(n));
	PXH_CHECK(s->keys2.ensure(n));
	PXH_CHECK(s->vals.ensure(n));
	PXH_CHECK(s->vals2.ensure(n));
	launch_keys(n, s->bmin.p, s->bmax.p, d_scene, s->gridF.p, s->gridI.p, d_rowOff, s->keys.p, s->vals.p, st);
	const int nblocks = (n + radix_tile() - 1) / radix_tile();
	PXH_CHECK(s->hist.ensure(size_t(256) * std::max(nblocks, 1)));
	bool alt = false;
	const int bitEnd = 32 + ((bitsFor(totalRows) + 7) / 8) * 8;
	launch_radix_sort(s->keys.p, s->vals.p, s->keys2.p, s->vals2.p, n, 0, bitEnd, s->hist.p, s->scanTmp.p, st, alt);
	s->sortedKeys = alt ? s->keys2.p : s->keys.p;
	const uint32_t* sv = alt ? s->vals2.p : s->vals.p;

	PXH_CHECK(s->smin.ensure(n));
	PXH_CHECK(s->smax.ensure(n));
	PXH_CHECK(s->rowStart.ensure(size_t(totalRows) + 2));
	launch_gather_sorted(n, sv, s->bmin.p, s->bmax.p, s->smin.p, s->smax.p, st);
	launch_row_start(n, s->sortedKeys, s->rowStart.p, totalRows, st);

	const int64_t nc = int64_t(PAIR_CLASSES) * n + 1;
	PXH_CHECK(s->counts.ensure(nc));
	PXH_CHECK(hipMemsetAsync(s->counts.p + PAIR_CLASSES * int64_t(n), 0, sizeof(uint32_t), st));
	launch_sweep(false, n, s->sortedKeys, s->smin.p, s->smax.p, d_scene, d_group, groupLut, s->gridI.p, d_rowOff,
	             s->rowStart.p, s->counts.p, nullptr, nullptr, 0u, st);
	launch_scan(s->counts.p, nc, s->scanTmp.p, st);
	return bp_write_pairs(s, n, d_scene, d_group, d_rowOff, st, pairCap, groupLut);
}

// Standalone broadphase update (PxBroadPhase::update, PxBroadPhase.h:576) on a
// context holding the previous update's sorted pair keys (pkPrev, nPrev):
// inflates the n bounds, finds the overlapping pairs (groups filtered by
// groupLut), sorts their keys into pkCur (nCur) and writes created = cur \ prev
// and deleted = prev \ cur as sorted (id0, id1) pairs (nCreated, nDeleted),
// ids being the boxes' handles (ids, device, every id below idBound) or their
// indices (ids null). hostCounts: at least two host-visible uint32.
// growPairs: whether the pair buffer may grow past pairCap when the pairs
// outnumber it; when it may not, the update stops with PXH_ERR_CAPACITY and
// *required set to the pair count.
template <typename Ctx>
int bp_update(Ctx* bp, const float* bounds, const uint32_t* groups, const float* distances, int n, hipStream_t st,
              uint32_t* hostCounts, bool growPairs, int* required, const uint32_t* ids = nullptr,
              uint32_t idBound = 0, uint64_t groupLut = GROUP_LUT_ALL)
{
	const size_t nn = std::max(n, 1);
	PXH_CHECK(bp->bmin.ensure(nn));
	PXH_CHECK(bp->bmax.ensure(nn));
	PXH_CHECK(bp->scene.ensure(nn));
	PXH_CHECK(bp->group.ensure(nn));
	PXH_CHECK(bp->bodyStart.ensure(2));
	launch_bp_inflate(n, bounds, distances, groups, bp->bmin.p, bp->bmax.p, bp->scene.p, bp->group.p, bp->bodyStart.p,
	                  st);
	int P = 0;
	if (n > 0)
	{
		// one world of at most 2n rows
		const uint32_t rows = std::min<uint32_t>(2u * uint32_t(n), 1u << 20);
		PXH_CHECK(bp->rowOff.ensure(2));
		launch_row_off_single(bp->rowOff.p, rows, st);
		if (growPairs)
			bp->pairCap = std::max(bp->pairCap, size_t(4) * n);
		int r = broadphase(bp, n, 1, bp->scene.p, bp->group.p, bp->rowOff.p, rows, st, bp->pairCap, groupLut);
		if (r)
			return r;
		const uint32_t* src[1] = {bp->counts.p + PAIR_CLASSES * int64_t(n)};
		r = readCounts(st, hostCounts, src, 1);
		if (r)
			return r;
		P = int(hostCounts[0]);
		if (size_t(P) > bp->pairCap)
		{
			if (!growPairs)
			{
				if (required)
					*required = P;
				set_last_error("broadphase: " + std::to_string(P) + " overlapping pairs exceed the capacity of " +
				               std::to_string(bp->pairCap));
				return PXH_ERR_CAPACITY;
			}
			bp->pairCap = size_t(P) + size_t(P) / 4;
			r = bp_write_pairs(bp, n, bp->scene.p, bp->group.p, bp->rowOff.p, st, bp->pairCap, groupLut);
			if (r)
				return r;
		}
	}
	// sorted pair keys of this update
	std::swap(bp->pkPrev, bp->pkCur);
	bp->nPrev = bp->nCur;
	PXH_CHECK(bp->pkCur.ensure(std::max(P, 1)));
	PXH_CHECK(bp->pkTmp.ensure(std::max(P, 1)));
	PXH_CHECK(bp->vals.ensure(std::max(P, 1)));
	PXH_CHECK(bp->vals2.ensure(std::max(P, 1)));
	launch_pair_keys(P, bp->pairs.p, ids, bp->pkCur.p, bp->vals.p, st);
	const int nblocks = (P + radix_tile() - 1) / radix_tile();
	PXH_CHECK(bp->hist.ensure(size_t(256) * std::max(nblocks, 1)));
	PXH_CHECK(bp->scanTmp.ensure(
	    scan_tmp_size(std::max<int64_t>(std::max<int64_t>(int64_t(256) * std::max(nblocks, 1), P + 1),
	                                    int64_t(bp->nCur) + 1)) +
	    1024));
	bool alt = false;
	const int bits = 32 + ((bitsFor(uint64_t(ids ? std::max<uint32_t>(idBound, 1u) : uint32_t(std::max(n, 1)))) + 7) / 8) * 8;
	launch_radix_sort(bp->pkCur.p, bp->vals.p, bp->pkTmp.p, bp->vals2.p, P, 0, bits, bp->hist.p, bp->scanTmp.p, st,
	                  alt);
	if (alt)
		std::swap(bp->pkCur, bp->pkTmp);
	bp->nCur = P;
	// created = cur \ prev, deleted = prev \ cur
	const int np = bp->nPrev;
	PXH_CHECK(bp->flags.ensure(size_t(P) + 1));
	PXH_CHECK(bp->flags2.ensure(size_t(np) + 1));
	PXH_CHECK(bp->created.ensure(std::max(P, 1)));
	PXH_CHECK(bp->deleted.ensure(std::max(np, 1)));
	launch_set_diff(P, bp->pkCur.p, np, bp->pkPrev.p, bp->flags.p, st);
	PXH_CHECK(hipMemsetAsync(bp->flags.p + P, 0, sizeof(uint32_t), st));
	PXH_CHECK(bp->counts.ensure(size_t(P) + 1));
	PXH_CHECK(hipMemcpyAsync(bp->counts.p, bp->flags.p, sizeof(uint32_t) * (P + 1), hipMemcpyDeviceToDevice, st));
	launch_scan(bp->counts.p, int64_t(P) + 1, bp->scanTmp.p, st);
	launch_compact_keys(P, bp->pkCur.p, bp->flags.p, bp->counts.p, bp->created.p, st);
	PXH_CHECK(hipMemcpyAsync(hostCounts, bp->counts.p + P, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
	launch_set_diff(np, bp->pkPrev.p, P, bp->pkCur.p, bp->flags2.p, st);
	PXH_CHECK(hipMemsetAsync(bp->flags2.p + np, 0, sizeof(uint32_t), st));
	PXH_CHECK(bp->rows.ensure(size_t(np) + 1));
	PXH_CHECK(hipMemcpyAsync(bp->rows.p, bp->flags2.p, sizeof(uint32_t) * (np + 1), hipMemcpyDeviceToDevice, st));
	launch_scan(bp->rows.p, int64_t(np) + 1, bp->scanTmp.p, st);
	launch_compact_keys(np, bp->pkPrev.p, bp->flags2.p, bp->rows.p, bp->deleted.p, st);
	PXH_CHECK(hipMemcpyAsync(hostCounts + 1, bp->rows.p + np, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
	PXH_CHECK(hipStreamSynchronize(st));
	PXH_CHECK(hipGetLastError());
	bp->nCreated = int(hostCounts[0]);
	bp->nDeleted = int(hostCounts[1]);
	bp->n = n;
	return PXH_OK;
}

}  // namespace pxh
