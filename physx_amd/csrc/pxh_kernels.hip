// HIP kernels of the rigid-body step for gfx950 (wave64). One lane per body,
// pair or constraint; all arrays are structure-of-arrays in HBM so a wave's
// loads are coalesced float4 rows.
//
// Stage map (reference stage -> kernels here):
//   shape bounds (updateTransformCacheAndBoundArrayLaunch)  k_bounds
//   broadphase (gpubroadphase/src/CUDA/broadphase.cu)        k_scene_grid, k_keys, radix sort, k_row_start,
//                                                            k_gather_sorted, k_sweep<count|write>
//   narrowphase (cudaSphere.cu, cudaBox.cu, cudaGJKEPA.cu)  k_np_pairs, k_np_planes
//   partitioning (PxgIncrementalPartition, host in ref)     k_con_build, k_color_min, k_color_pick, colour sort
//   pre-integration (preIntegrationTGS.cu)                  k_preintegrate
//   contact prep (constraintBlockPrepTGS.cu)                k_prep
//   solve (solverMultiBlockTGS.cu solveBlockUnified)        k_solve
//   integration (integrationTGS.cu)                         k_integrate, k_finalize
#include "pxh_kernels.h"
#include <hip/hip_ext.h>
#include "pxh_np.h"
#include "pxh_convex.h"
#include "pxh_joint.h"
#include "pxh_mesh.h"

#include <algorithm>

namespace pxh {

// ---------------------------------------------------------------------------
// wave / block scans
// ---------------------------------------------------------------------------
PXH_D uint32_t waveInclScan(uint32_t v)
{
	const int lane = threadIdx.x & 63;
#pragma unroll
	for (int o = 1; o < 64; o <<= 1)
	{
		const uint32_t t = __shfl_up(v, o, 64);
		if (lane >= o)
			v += t;
	}
	return v;
}

// exclusive scan over the block (blockDim.x <= 1024, multiple of 64); returns
// the exclusive prefix, writes the block total
PXH_D uint32_t blockExclScan(uint32_t v, uint32_t* sh, uint32_t& total)
{
	const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
	const uint32_t inc = waveInclScan(v);
	if (lane == 63)
		sh[wave] = inc;
	__syncthreads();
	if (threadIdx.x < 64)
	{
		const uint32_t w = threadIdx.x < nw ? sh[threadIdx.x] : 0u;
		const uint32_t wi = waveInclScan(w);
		if (threadIdx.x < nw)
			sh[threadIdx.x] = wi - w;
		if (threadIdx.x == nw - 1)
			sh[32] = wi;
	}
	__syncthreads();
	const uint32_t r = sh[wave] + inc - v;
	total = sh[32];
	__syncthreads();
	return r;
}

constexpr int SCAN_THREADS = 256, SCAN_ITEMS = 8, SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

// in-place exclusive scan of each 2048-element tile; tile sums to blockSums
__global__ void __launch_bounds__(SCAN_THREADS) k_scan_tiles(uint32_t* data, int64_t n, uint32_t* blockSums)
{
	__shared__ uint32_t sh[33];
	__shared__ uint32_t tile[SCAN_TILE];
	const int64_t base = int64_t(blockIdx.x) * SCAN_TILE;
	for (int k = 0; k < SCAN_ITEMS; k++)
	{
		const int64_t g = base + k * SCAN_THREADS + threadIdx.x;
		tile[k * SCAN_THREADS + threadIdx.x] = g < n ? data[g] : 0u;
	}
	__syncthreads();
	uint32_t loc[SCAN_ITEMS];
	uint32_t sum = 0;
	for (int k = 0; k < SCAN_ITEMS; k++)
	{
		loc[k] = tile[threadIdx.x * SCAN_ITEMS + k];
		sum += loc[k];
	}
	uint32_t total;
	uint32_t pre = blockExclScan(sum, sh, total);
	for (int k = 0; k < SCAN_ITEMS; k++)
	{
		tile[threadIdx.x * SCAN_ITEMS + k] = pre;
		pre += loc[k];
	}
	__syncthreads();
	for (int k = 0; k < SCAN_ITEMS; k++)
	{
		const int64_t g = base + k * SCAN_THREADS + threadIdx.x;
		if (g < n)
			data[g] = tile[k * SCAN_THREADS + threadIdx.x];
	}
	if (threadIdx.x == 0 && blockSums)
		blockSums[blockIdx.x] = total;
}

__global__ void k_scan_add(uint32_t* data, int64_t n, const uint32_t* blockOffsets)
{
	const int64_t g = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (g < n)
		data[g] += blockOffsets[g / SCAN_TILE];
}

// ---------------------------------------------------------------------------
// LSD radix sort of (u64 key, u32 value), 8-bit digits, stable
// ---------------------------------------------------------------------------
// 512-key tiles: a 262k-key sort spreads over 512 workgroups, two per CU
constexpr int RS_THREADS = 256, RS_ITEMS = 2, RS_TILE = RS_THREADS * RS_ITEMS;
int radix_tile() { return RS_TILE; }

__global__ void __launch_bounds__(RS_THREADS) k_rs_hist(const uint64_t* keys, int n, int shift, uint32_t* hist, int nblocks)
{
	__shared__ uint32_t cnt[256];
	cnt[threadIdx.x] = 0;
	__syncthreads();
	const int base = blockIdx.x * RS_TILE;
	for (int k = 0; k < RS_ITEMS; k++)
	{
		const int g = base + k * RS_THREADS + threadIdx.x;
		if (g < n)
			atomicAdd(&cnt[uint32_t(keys[g] >> shift) & 0xFFu], 1u);
	}
	__syncthreads();
	hist[threadIdx.x * nblocks + blockIdx.x] = cnt[threadIdx.x];
}

__global__ void __launch_bounds__(RS_THREADS) k_rs_scatter(const uint64_t* kin, const uint32_t* vin, uint64_t* kout,
                                                            uint32_t* vout, int n, int shift, const uint32_t* hist,
                                                            int nblocks)
{
	__shared__ uint64_t sk[RS_TILE];
	__shared__ uint32_t sv[RS_TILE];
	__shared__ uint32_t sh[33];
	__shared__ uint32_t dstart[256];
	const int base = blockIdx.x * RS_TILE;
	const int nvalid = min(RS_TILE, n - base);
	for (int k = 0; k < RS_ITEMS; k++)
	{
		const int l = k * RS_THREADS + threadIdx.x;
		const int g = base + l;
		sk[l] = g < n ? kin[g] : ~0ull;
		sv[l] = g < n ? vin[g] : 0u;
	}
	__syncthreads();
	// stable split of the tile by each of the 8 bits, least significant first
	uint64_t k8[RS_ITEMS];
	uint32_t v8[RS_ITEMS];
	for (int k = 0; k < RS_ITEMS; k++)
	{
		k8[k] = sk[threadIdx.x * RS_ITEMS + k];
		v8[k] = sv[threadIdx.x * RS_ITEMS + k];
	}
	for (int b = 0; b < 8; b++)
	{
		const int bit = shift + b;
		uint32_t zeros = 0;
		for (int k = 0; k < RS_ITEMS; k++)
			zeros += ((k8[k] >> bit) & 1ull) ? 0u : 1u;
		uint32_t totalZeros;
		const uint32_t zb = blockExclScan(zeros, sh, totalZeros);
		uint32_t z = zb, o = totalZeros + threadIdx.x * RS_ITEMS - zb;
		__syncthreads();
		for (int k = 0; k < RS_ITEMS; k++)
		{
			const bool one = (k8[k] >> bit) & 1ull;
			const uint32_t dst = one ? o++ : z++;
			sk[dst] = k8[k];
			sv[dst] = v8[k];
		}
		__syncthreads();
		for (int k = 0; k < RS_ITEMS; k++)
		{
			k8[k] = sk[threadIdx.x * RS_ITEMS + k];
			v8[k] = sv[threadIdx.x * RS_ITEMS + k];
		}
		__syncthreads();
	}
	// digit starts inside the sorted tile
	for (int k = 0; k < RS_ITEMS; k++)
	{
		const int p = threadIdx.x * RS_ITEMS + k;
		const uint32_t d = uint32_t(k8[k] >> shift) & 0xFFu;
		const uint32_t dp = p == 0 ? 0xFFFFFFFFu : (uint32_t(sk[p - 1] >> shift) & 0xFFu);
		if (p < nvalid && d != dp)
			dstart[d] = p;
	}
	__syncthreads();
	for (int k = 0; k < RS_ITEMS; k++)
	{
		const int p = k * RS_THREADS + threadIdx.x;
		if (p < nvalid)
		{
			const uint64_t key = sk[p];
			const uint32_t d = uint32_t(key >> shift) & 0xFFu;
			const uint32_t dst = hist[d * nblocks + blockIdx.x] + (p - dstart[d]);
			kout[dst] = key;
			vout[dst] = sv[p];
		}
	}
}

// ---------------------------------------------------------------------------
// bounds and broadphase
// ---------------------------------------------------------------------------

// a body's shape bounds, centre - w to centre + w (Gu::computeBounds,
// geomutils/src/GuBounds.cpp:354-395)
PXH_D void shapeBounds(int i, const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, f3& mn,
                       f3& mx)
{
	const float4 g = geom[i];
	f3 p = mk(pos[i]);
	f3 w;
	if (int(g.w) == GEOM_SPHERE)
		w = mk(g.x, g.x, g.x);
	else if (int(g.w) == GEOM_CAPSULE)
	{
		// computeCapsuleBounds (geomutils/include/GuBounds.h:54-62)
		const m33 R = qmat(mkq(rot[i]));
		w = mk(fabsf(R.c0.x) * g.y + g.x, fabsf(R.c0.y) * g.y + g.x, fabsf(R.c0.z) * g.y + g.x);
	}
	else
	{
		const m33 R = qmat(mkq(rot[i]));
		f3 e = mk(g.x, g.y, g.z);
		if (int(g.w) == GEOM_CONVEXMESH)
		{
			// hull bounds: the local box rotated about the pose
			// (geomutils/src/GuBounds.cpp:136-140 transformNoEmptyTest)
			const int h = int(g.x);
			e = mk(hs.be[h]);
			p = mmul(R, mk(hs.bc[h])) + p;
		}
		const f3 c0 = R.c0 * e.x, c1 = R.c1 * e.y, c2 = R.c2 * e.z;
		w = mk(fabsf(c0.x) + fabsf(c1.x) + fabsf(c2.x), fabsf(c0.y) + fabsf(c1.y) + fabsf(c2.y),
		       fabsf(c0.z) + fabsf(c1.z) + fabsf(c2.z));
	}
	mn = p - w;
	mx = p + w;
}

// shape bounds plus the broadphase's contact-distance inflation
// (lowlevelaabb/src/BpBroadPhaseABP.cpp:1187-1192)
__global__ void k_bounds(int n, const float4* pos, const float4* rot, const float4* geom, HullSet hs, float co,
                         float4* bmin, float4* bmax)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	f3 mn, mx;
	shapeBounds(i, pos, rot, geom, hs, mn, mx);
	bmin[i] = make_float4(mn.x - co, mn.y - co, mn.z - co, geom[i].w);
	bmax[i] = make_float4(mx.x + co, mx.y + co, mx.z + co, 0.0f);
}

// Per-world row grid for the sweep: rows are (y,z) cells no smaller than the
// widest inflated box, so overlapping boxes sit in the same or adjacent rows.
// Pass 1 reduces each world's centre range and widest box with order-
// preserving integer atomics (one per wave when the wave holds one world).
PXH_D float decode_float(uint32_t u) { return __uint_as_float((u & 0x80000000u) ? (u ^ 0x80000000u) : ~u); }

constexpr int GRID_ITEMS = 16;
__global__ void __launch_bounds__(256) k_grid_reduce(int n, const float4* bmin, const float4* bmax, const int* scene,
                                                      uint32_t* acc)
{
	__shared__ uint32_t red[5][256];
	const int base = blockIdx.x * 256 * GRID_ITEMS;
	const int first = base + threadIdx.x * GRID_ITEMS;
	const int blockLast = min(n, base + 256 * GRID_ITEMS) - 1;
	// one world in the whole block: reduce in LDS, one atomic set per block
	const bool blockUniform = scene[base] == scene[blockLast];
	uint32_t v[5] = {0xFFFFFFFFu, 0u, 0xFFFFFFFFu, 0u, 0u};
	int cur = -1;
	for (int i = first; i < min(n, first + GRID_ITEMS); i++)
	{
		const int s = scene[i];
		if (s != cur && cur >= 0 && !blockUniform)
		{
			uint32_t* w = acc + 5 * int64_t(cur);
			atomicMin(&w[0], v[0]); atomicMax(&w[1], v[1]); atomicMin(&w[2], v[2]); atomicMax(&w[3], v[3]);
			atomicMax(&w[4], v[4]);
			v[0] = 0xFFFFFFFFu; v[1] = 0u; v[2] = 0xFFFFFFFFu; v[3] = 0u; v[4] = 0u;
		}
		cur = s;
		const float4 a = bmin[i], b = bmax[i];
		const uint32_t cy = encode_float(0.5f * (a.y + b.y)), cz = encode_float(0.5f * (a.z + b.z));
		v[0] = min(v[0], cy); v[1] = max(v[1], cy); v[2] = min(v[2], cz); v[3] = max(v[3], cz);
		v[4] = max(v[4], encode_float(fmaxf(b.y - a.y, b.z - a.z)));
	}
	if (!blockUniform)
	{
		if (cur >= 0)
		{
			uint32_t* w = acc + 5 * int64_t(cur);
			atomicMin(&w[0], v[0]); atomicMax(&w[1], v[1]); atomicMin(&w[2], v[2]); atomicMax(&w[3], v[3]);
			atomicMax(&w[4], v[4]);
		}
		return;
	}
	for (int k = 0; k < 5; k++)
		red[k][threadIdx.x] = v[k];
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1)
	{
		if (threadIdx.x < o)
		{
			red[0][threadIdx.x] = min(red[0][threadIdx.x], red[0][threadIdx.x + o]);
			red[1][threadIdx.x] = max(red[1][threadIdx.x], red[1][threadIdx.x + o]);
			red[2][threadIdx.x] = min(red[2][threadIdx.x], red[2][threadIdx.x + o]);
			red[3][threadIdx.x] = max(red[3][threadIdx.x], red[3][threadIdx.x + o]);
			red[4][threadIdx.x] = max(red[4][threadIdx.x], red[4][threadIdx.x + o]);
		}
		__syncthreads();
	}
	if (threadIdx.x == 0)
	{
		uint32_t* w = acc + 5 * int64_t(scene[base]);
		atomicMin(&w[0], red[0][0]); atomicMax(&w[1], red[1][0]); atomicMin(&w[2], red[2][0]);
		atomicMax(&w[3], red[3][0]); atomicMax(&w[4], red[4][0]);
	}
}

__global__ void k_grid_init(int nScenes, uint32_t* acc)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s < nScenes)
	{
		acc[5 * s + 0] = 0xFFFFFFFFu; acc[5 * s + 1] = 0u;
		acc[5 * s + 2] = 0xFFFFFFFFu; acc[5 * s + 3] = 0u;
		acc[5 * s + 4] = 0u;
	}
}

// Each world's rows are capped by its slice of rowOff (the host sizes the row
// table up front, so no count comes back before the sort): the cell grows
// until ny * nz fits. Cells only need to be at least as wide as the widest
// box; wider cells give the same pairs with longer sweeps.
__global__ void k_grid_finish(int nScenes, const uint32_t* acc, float4* gridF, int2* gridI, const uint32_t* rowOff)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= nScenes)
		return;
	const uint32_t* w = acc + 5 * int64_t(s);
	const bool empty = w[0] == 0xFFFFFFFFu && w[1] == 0u;
	const float ly = empty ? 0.0f : decode_float(w[0]), hy = empty ? 0.0f : decode_float(w[1]);
	const float lz = empty ? 0.0f : decode_float(w[2]), hz = empty ? 0.0f : decode_float(w[3]);
	const float wm = empty ? 0.0f : decode_float(w[4]);
	float cell = wm * 1.01f + 1e-6f;
	cell = fmaxf(cell, fmaxf(hy - ly, hz - lz) * (1.0f / 1000.0f));
	const int64_t cap = int64_t(rowOff[s + 1] - rowOff[s]);
	int ny = 1, nz = 1;
	for (int it = 0; it < 96; it++)
	{
		ny = max(1, min(int((hy - ly) / cell) + 1, 1024));
		nz = max(1, min(int((hz - lz) / cell) + 1, 1024));
		if (int64_t(ny) * nz <= cap)
			break;
		cell *= 1.25f;
	}
	if (int64_t(ny) * nz > cap)
	{
		ny = 1;
		nz = 1;
		cell = fmaxf(cell, fmaxf(hy - ly, hz - lz) + 1.0f);
	}
	gridF[s] = make_float4(ly, lz, 1.0f / cell, cell);
	gridI[s] = make_int2(ny, nz);
}

// row table of a single world with cap rows
__global__ void k_row_off_single(uint32_t* rowOff, uint32_t cap)
{
	if (threadIdx.x == 0 && blockIdx.x == 0)
	{
		rowOff[0] = 0u;
		rowOff[1] = cap;
	}
}

__global__ void k_keys(int n, const float4* bmin, const float4* bmax, const int* scene, const float4* gridF,
                       const int2* gridI, const uint32_t* rowOff, uint64_t* keys, uint32_t* vals)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float4 a = bmin[i], b = bmax[i];
	const int s = scene[i];
	const float4 gf = gridF[s];
	const int2 gi = gridI[s];
	const float cy = 0.5f * (a.y + b.y), cz = 0.5f * (a.z + b.z);
	const int iy = min(max(int(floorf((cy - gf.x) * gf.z)), 0), gi.x - 1);
	const int iz = min(max(int(floorf((cz - gf.y) * gf.z)), 0), gi.y - 1);
	const uint32_t row = rowOff[s] + uint32_t(iz * gi.x + iy);
	keys[i] = (uint64_t(row) << 32) | encode_float(a.x);
	vals[i] = uint32_t(i);
}

__global__ void k_gather_sorted(int n, const uint32_t* sortedVals, const float4* bmin, const float4* bmax,
                                float4* smin, float4* smax)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const uint32_t i = sortedVals[s];
	smin[s] = bmin[i];
	float4 b = bmax[i];
	b.w = __uint_as_float(i);
	smax[s] = b;
}

// rowStart[r] = first sorted position of row r; rowStart[totalRows] = n.
// Keys past totalRows count as totalRows (rowStart has totalRows + 2 entries).
// Lane s fills the rows from the previous key's row (exclusive) to its own;
// the rows after the last key are filled by k_row_tail, one lane per row.
__global__ void k_row_start(int n, const uint64_t* sortedKeys, uint32_t* rowStart, uint32_t totalRows)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const uint32_t r = min(uint32_t(sortedKeys[s] >> 32), totalRows);
	const int64_t rp = s == 0 ? -1 : int64_t(min(uint32_t(sortedKeys[s - 1] >> 32), totalRows));
	for (int64_t rr = rp + 1; rr <= int64_t(r); rr++)
		rowStart[rr] = uint32_t(s);
}

__global__ void k_row_tail(int n, const uint64_t* sortedKeys, uint32_t* rowStart, uint32_t totalRows)
{
	const int64_t rr = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (rr > int64_t(totalRows) || n <= 0)
		return;
	const uint32_t last = min(uint32_t(sortedKeys[n - 1] >> 32), totalRows);
	if (rr > int64_t(last))
		rowStart[rr] = uint32_t(n);
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v)
{
	const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

// pair class from the two geometry types, spheres and capsules counting as
// round shapes: 0 round-round, 1 round-box, 2 box-box, 3 round-convex,
// 4 box-convex, 5 convex-convex
PXH_D int pairClass(int t0, int t1)
{
	const int nc = (t0 == GEOM_CONVEXMESH) + (t1 == GEOM_CONVEXMESH);
	const int nb = (t0 == GEOM_BOX) + (t1 == GEOM_BOX);
	return nc == 0 ? nb : (nc == 2 ? 5 : 3 + nb);
}

// Sweep of the sorted rows: lane s (body i) reports each overlapping body j in
// its own and the 8 neighbouring rows whose (minX, index) sorts after its own,
// so every pair is found exactly once. The overlap test is the reference ABP's
// inclusive test on inflated bounds (BpBroadPhaseABP.cpp:2246-2262, :2474).
// WRITE=false counts pairs per class into counts[cls*n + s]; WRITE=true writes
// them at the scanned offsets.
// Broadphase group filter (Bp::groupFiltering, lowlevelaabb/include/BpFiltering.h:99-109):
// boxes of equal groups never pair; otherwise bit (type0 * 8 + type1) of lut,
// type = group & 7 (BpFilter's 7 x 7 table, lowlevelaabb/src/BpFiltering.cpp,
// in an 8 x 8 bit mask; all bits set: only equal groups are filtered)
PXH_D bool groupsPair(int g0, int g1, uint64_t lut)
{
	return g0 != g1 && ((lut >> ((g0 & 7) * 8 + (g1 & 7))) & 1ull);
}

template <bool WRITE>
__global__ void k_sweep(int n, const uint64_t* sortedKeys, const float4* smin, const float4* smax, const int* scene,
                        const int* group, uint64_t groupLut, const int2* gridI, const uint32_t* rowOff,
                        const uint32_t* rowStart, uint32_t* counts, const uint32_t* offsets, uint2* pairs,
                        uint32_t pairCap)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= n)
		return;
	const uint64_t key = sortedKeys[s];
	const float4 a0 = smin[s], a1 = smax[s];
	const uint32_t i = __float_as_uint(a1.w);
	const int ti = int(a0.w);
	const int sc = scene[i];
	const int gi_ = group ? group[i] : -1;
	const int2 gi = gridI[sc];
	const uint32_t r = uint32_t(key >> 32);
	const uint32_t ro = rowOff[sc];
	const int lr = int(r - ro);
	const int cy = lr % gi.x, cz = lr / gi.x;
	uint32_t cnt[PAIR_CLASSES];
	uint32_t off[PAIR_CLASSES];
	for (int c = 0; c < PAIR_CLASSES; c++)
	{
		cnt[c] = 0u;
		off[c] = WRITE ? offsets[int64_t(c) * n + s] : 0u;
	}
	for (int dz = -1; dz <= 1; dz++)
	{
		const int z = cz + dz;
		if (z < 0 || z >= gi.y)
			continue;
		for (int dy = -1; dy <= 1; dy++)
		{
			const int y = cy + dy;
			if (y < 0 || y >= gi.x)
				continue;
			const uint32_t rr = ro + uint32_t(z * gi.x + y);
			uint32_t lo = rowStart[rr];
			const uint32_t hi = rowStart[rr + 1];
			if (rr == r)
				lo = uint32_t(s) + 1u;
			else
			{
				// first position whose (minX key, body index) follows ours
				uint32_t l = lo, h = hi;
				while (l < h)
				{
					const uint32_t m = (l + h) >> 1;
					const uint64_t km = sortedKeys[m];
					const bool after = (uint32_t(km) > uint32_t(key)) ||
					                   (uint32_t(km) == uint32_t(key) && __float_as_uint(smax[m].w) > i);
					if (after) h = m;
					else l = m + 1;
				}
				lo = l;
			}
			for (uint32_t j = lo; j < hi; j++)
			{
				const float4 b0 = smin[j];
				if (b0.x > a1.x)
					break;
				const float4 b1 = smax[j];
				if (b1.y < a0.y || a1.y < b0.y || b1.z < a0.z || a1.z < b0.z)
					continue;
				const uint32_t jj = __float_as_uint(b1.w);
				if (group && !groupsPair(gi_, group[jj], groupLut))
					continue;
				const int tj = int(b0.w);
				const int c = pairClass(ti, tj);
				if (WRITE)
				{
					// shape order of the reference's contact managers: the lower
					// geometry type first (contact-method table order); two
					// shapes of one type put the later-added body first
					// (shouldSwapBodies, simulationcontroller/src/ScNPhaseCore.cpp:182-249)
					uint32_t pa, pb;
					if (ti != tj) { pa = ti < tj ? i : jj; pb = ti < tj ? jj : i; }
					else { pa = max(i, jj); pb = min(i, jj); }
					// pairs past the buffer's capacity are counted, not written
					if (off[c] < pairCap)
						pairs[off[c]] = make_uint2(pa, pb);
					off[c]++;
				}
				else
					cnt[c]++;
			}
		}
	}
	if (!WRITE)
		for (int c = 0; c < PAIR_CLASSES; c++)
			counts[int64_t(c) * n + s] = cnt[c];
}

// ---------------------------------------------------------------------------
// narrowphase
// ---------------------------------------------------------------------------
PXH_D void storeContacts(int slot, const Contacts& c, int cnt, float4* ctN, float4* ctP)
{
	ctN[slot] = make_float4(c.n.x, c.n.y, c.n.z, __int_as_float(cnt));
	for (int k = 0; k < cnt; k++)
		ctP[4 * int64_t(slot) + k] = make_float4(c.p[k].x, c.p[k].y, c.p[k].z, c.sep[k]);
}

// Body-body pairs; the shape order is the reference contact-method table's
// (lower geometry type first), so body a of a sphere-box pair is the sphere.
// Lanes past the pair count on the device (*dP) exit; pair p fills slot
// slotBase + p.
__global__ void k_np_pairs(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                           const float4* pos, const float4* rot, const float4* geom, float cd, float tol, float4* ctN,
                           float4* ctP)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= cap || uint32_t(p) >= *dConvex || uint32_t(p) >= *dP)
		return;
	const uint2 ab = pairs[p];
	const float4 g0 = geom[ab.x], g1 = geom[ab.y];
	const int t0 = int(g0.w), t1 = int(g1.w);
	Contacts c;
	c.count = 0;
	int cnt = 0;
	const float4 p0 = pos[ab.x], p1 = pos[ab.y];
	if (t0 == GEOM_SPHERE && t1 == GEOM_SPHERE)
		cnt = sphereSphere(g0.x, mk(p0), g1.x, mk(p1), cd, c);
	else
	{
		xf x0, x1;
		x0.p = mk(p0); x0.q = t0 == GEOM_SPHERE ? mkq(make_float4(0.0f, 0.0f, 0.0f, 1.0f)) : mkq(rot[ab.x]);
		x1.p = mk(p1); x1.q = mkq(rot[ab.y]);
		if (t0 == GEOM_SPHERE && t1 == GEOM_CAPSULE)
			cnt = sphereCapsule(g0.x, x0.p, g1.x, g1.y, x1, cd, c);
		else if (t0 == GEOM_SPHERE)
			cnt = sphereBox(g0.x, x0.p, mk(g1.x, g1.y, g1.z), x1, cd, c);
		else if (t1 == GEOM_CAPSULE)
			cnt = capsuleCapsule(g0.x, g0.y, x0, g1.x, g1.y, x1, cd, c);
		else if (t0 == GEOM_CAPSULE)
			cnt = capsulePoly(g0.x, g0.y, x0, boxPoly(mk(g1.x, g1.y, g1.z)), x1, cd, tol, c);
		else
			cnt = boxBox(mk(g0.x, g0.y, g0.z), x0, mk(g1.x, g1.y, g1.z), x1, cd, tol, c);
	}
	storeContacts(slotBase + p, c, cnt, ctN, ctP);
}

// Body-plane pairs (every body against every plane of its world; a plane's
// bounds are unbounded so the broadphase always reports these pairs). The body
// is constraint body a; normals point from the plane to the body.
__global__ void k_np_planes(int Q, int slotBase, const int* qBody, const int* qPlane, const float4* pos,
                            const float4* rot, const float4* geom, const float4* planes, float cd, float4* ctN,
                            float4* ctP)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= Q)
		return;
	const int i = qBody[q];
	const float4 g = geom[i];
	if (int(g.w) == GEOM_CONVEXMESH)
		return;
	const float4 pl = planes[qPlane[q]];
	Contacts c;
	c.count = 0;
	int cnt;
	if (int(g.w) == GEOM_SPHERE)
		cnt = spherePlane(g.x, mk(pos[i]), mk(pl), pl.w, cd, c);
	else
	{
		xf x;
		x.p = mk(pos[i]); x.q = mkq(rot[i]);
		if (int(g.w) == GEOM_CAPSULE)
			cnt = planeCapsule(mk(pl), pl.w, g.x, g.y, x, cd, c);
		else
			cnt = planeBox(mk(pl), pl.w, mk(g.x, g.y, g.z), x, cd, c);
		c.n = -c.n;
	}
	storeContacts(slotBase + q, c, cnt, ctN, ctP);
}

// Sphere- and capsule-convex pairs: the class segments from *dConvex to *dP
// hold the convex pairs; lane q takes pair *dConvex + q when its first shape
// is round. For a box- or convex-convex pair it makes the bounding-sphere
// test and lists the pair for k_np_convex_wave (list[1] counts, entries from
// list[2 + cap]) or writes its empty slot.
__global__ void k_np_convex(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                            const float4* pos, const float4* rot, const float4* geom, HullSet hs, float cd, float tol,
                            float4* ctN, float4* ctP, uint32_t* list)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	const uint32_t p = *dConvex + uint32_t(q);
	const bool live = q < cap && p < *dP && p < uint32_t(cap);
	uint2 ab = make_uint2(0u, 0u);
	float4 g0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), g1 = g0;
	int t0 = -1;
	if (live)
	{
		ab = pairs[p];
		g0 = geom[ab.x];
		g1 = geom[ab.y];
		t0 = int(g0.w);
	}
	const bool round = t0 == GEOM_SPHERE || t0 == GEOM_CAPSULE;
	// Polytope pairs: those whose bounding spheres come within the contact
	// distance go to the search list (one atomic per wave), the others get
	// their empty contact slot here.
	bool search = false;
	if (live && !round)
	{
		const f3 p0 = mk(pos[ab.x]), p1 = mk(pos[ab.y]);
		const float r0 = t0 == GEOM_BOX ? sqrtf(g0.x * g0.x + g0.y * g0.y + g0.z * g0.z) : hs.be[int(g0.x)].w;
		const float r1 = hs.be[int(g1.x)].w;
		const f3 d = p1 - p0;
		const float reach = r0 + r1 + cd;
		search = dot(d, d) <= reach * reach;
		if (!search)
			ctN[slotBase + int(p)] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
	}
	const uint64_t m = __ballot(search);
	if (m)
	{
		const int lane = threadIdx.x & 63, lead = __ffsll((unsigned long long)m) - 1;
		uint32_t at = 0u;
		if (lane == lead)
			at = atomicAdd(list + 1, uint32_t(__popcll(m)));
		at = __shfl(at, lead);
		if (search)
			list[2 + cap + at + uint32_t(__popcll(m & ((1ull << lane) - 1ull)))] = p;
	}
	if (!live || !round)
		return;
	xf x1;
	x1.p = mk(pos[ab.y]); x1.q = mkq(rot[ab.y]);
	Contacts c;
	c.count = 0;
	int cnt;
	if (t0 == GEOM_SPHERE)
		cnt = sphereConvex(g0.x, mk(pos[ab.x]), hullPoly(hs, int(g1.x)), x1, cd, c);
	else
	{
		xf x0;
		x0.p = mk(pos[ab.x]); x0.q = mkq(rot[ab.x]);
		cnt = capsulePoly(g0.x, g0.y, x0, hullPoly(hs, int(g1.x)), x1, cd, tol, c);
	}
	storeContacts(slotBase + int(p), c, cnt, ctN, ctP);
}

// Box- and convex-convex pairs, separation phase: one wavefront per pair
// (grid-stride over the convex segments) finds the closest features or the
// minimum translation for each pair of the search list k_np_convex wrote
// (list[1] entries from list[2 + cap]). A pair within the contact distance
// stores (dist, n), cA, cB at wit[3p .. 3p+2] and is appended to the
// touching list (list[0] entries from list[2]); any other pair gets its empty
// contact slot here. Polytopes past the LDS limits are searched by lane 0
// alone.
constexpr int CVX_WAVES = 4;
__global__ void __launch_bounds__(64 * CVX_WAVES, 4)   // at least 4 waves per SIMD
k_np_convex_wave(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                 const float4* pos, const float4* rot, const float4* geom, HullSet hs, float cd, float4* wit,
                 uint32_t* list, float4* ctN)
{
	__shared__ CvxLds lds[CVX_WAVES];
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	const uint32_t n = min(list[1], uint32_t(cap));
	for (uint32_t i = blockIdx.x * CVX_WAVES + w; i < n; i += gridDim.x * CVX_WAVES)
	{
		const uint32_t p = list[2 + cap + i];
		const uint2 ab = pairs[p];
		const float4 g0 = geom[ab.x], g1 = geom[ab.y];
		const int t0 = int(g0.w);
		float4 sep = make_float4(2.0f * cd + 1.0f, 1.0f, 0.0f, 0.0f);
		f3 cA = mk(0.0f, 0.0f, 0.0f), cB = cA;
		xf x0, x1;
		x0.p = mk(pos[ab.x]); x0.q = mkq(rot[ab.x]);
		x1.p = mk(pos[ab.y]); x1.q = mkq(rot[ab.y]);
		const Poly B = hullPoly(hs, int(g1.x));
		const Poly A = t0 == GEOM_BOX ? boxPoly(mk(g0.x, g0.y, g0.z)) : hullPoly(hs, int(g0.x));
		if (fitsWave(A, B))
			convexSeparationWave(A, x0, B, x1, cd, lds[w], lane, sep, cA, cB);
		else if (lane == 0)
		{
			const mxf aToB = mxfOf(xinvmul(x1, x0));
			f3 n = mk(1.0f, 0.0f, 0.0f);
			polySeparation(A, aToB, B, cd, sep.x, n, cA, cB);
			sep.y = n.x; sep.z = n.y; sep.w = n.z;
		}
		if (lane == 0)
		{
			if (sep.x <= cd)
			{
				wit[3 * int64_t(p)] = sep;
				wit[3 * int64_t(p) + 1] = f4(cA, 0.0f);
				wit[3 * int64_t(p) + 2] = f4(cB, 0.0f);
				list[2 + atomicAdd(list, 1u)] = p;
			}
			else
				ctN[slotBase + int(p)] = make_float4(0.0f, 0.0f, 0.0f, __int_as_float(0));
		}
	}
}

// Box- and convex-convex pairs, contact phase: one lane per listed (touching)
// pair builds the contacts from the separation record (witness polygons,
// clipping, reduction); the list order does not matter, pair p fills slot
// slotBase + p.
__global__ void k_np_convex_finish(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase,
                                   const uint2* pairs, const float4* pos, const float4* rot, const float4* geom,
                                   HullSet hs, float cd, float tol, const float4* wit, const uint32_t* list,
                                   float4* ctN, float4* ctP)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= cap || uint32_t(q) >= list[0])
		return;
	const uint32_t p = list[2 + q];
	const uint2 ab = pairs[p];
	const float4 g0 = geom[ab.x], g1 = geom[ab.y];
	const int t0 = int(g0.w);
	if (t0 == GEOM_SPHERE || t0 == GEOM_CAPSULE)
		return;
	const float4 sep = wit[3 * int64_t(p)];
	Contacts c;
	c.count = 0;
	int cnt = 0;
	if (sep.x <= cd)
	{
		xf x0, x1;
		x0.p = mk(pos[ab.x]); x0.q = mkq(rot[ab.x]);
		x1.p = mk(pos[ab.y]); x1.q = mkq(rot[ab.y]);
		const Poly B = hullPoly(hs, int(g1.x));
		const Poly A = t0 == GEOM_BOX ? boxPoly(mk(g0.x, g0.y, g0.z)) : hullPoly(hs, int(g0.x));
		const mxf aToB = mxfOf(xinvmul(x1, x0));
		cnt = convexContactsFrom(A, x0, B, x1, aToB, sep.x, mk(sep.y, sep.z, sep.w), mk(wit[3 * int64_t(p) + 1]),
		                         mk(wit[3 * int64_t(p) + 2]), cd, tol, c);
	}
	storeContacts(slotBase + int(p), c, cnt, ctN, ctP);
}

// Body-plane pairs of convex bodies (the others are left to k_np_planes); the
// hull is constraint body a, so normals point from the plane to the hull.
__global__ void k_np_planes_convex(int Q, int slotBase, const int* qBody, const int* qPlane, const float4* pos,
                                   const float4* rot, const float4* geom, const float4* planes, HullSet hs, float cd,
                                   float tol, float4* ctN, float4* ctP)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= Q)
		return;
	const int i = qBody[q];
	const float4 g = geom[i];
	if (int(g.w) != GEOM_CONVEXMESH)
		return;
	const float4 pl = planes[qPlane[q]];
	xf x;
	x.p = mk(pos[i]); x.q = mkq(rot[i]);
	Contacts c;
	c.count = 0;
	const int cnt = planeConvex(mk(pl), pl.w, hullPoly(hs, int(g.x)), x, cd, tol, c);
	c.n = -c.n;
	storeContacts(slotBase + q, c, cnt, ctN, ctP);
}

// Contact data of the last step, one 128-byte record per constraint in
// constraint order (row o of the colour order holds constraint order[o]):
// body0, body1 (or -1 - plane), contact count | normal (shape1 -> shape0) |
// 4 x (point, separation) | 4 applied normal impulses | 0. Lane 0 stores the
// number of records written.
__global__ void k_copy_contacts(int C, int maxPairs, const int* order, const int* conA, const int* conB,
                                const int* conSlot, const float4* ctN, const float4* ctP, const float* nF,
                                float4* out, uint32_t* nbPairs)
{
	const int o = blockIdx.x * blockDim.x + threadIdx.x;
	if (o == 0 && nbPairs)
		*nbPairs = uint32_t(min(C, maxPairs));
	if (o >= C)
		return;
	const int c = order[o];
	if (c >= maxPairs)
		return;
	const int slot = conSlot[c];
	const float4 n = ctN[slot];
	const int cnt = __float_as_int(n.w);
	float4* r = out + 8 * int64_t(c);
	r[0] = make_float4(__int_as_float(conA[c]), __int_as_float(conB[c]), __int_as_float(cnt), 0.0f);
	r[1] = make_float4(n.x, n.y, n.z, 0.0f);
	float f[4];
	for (int k = 0; k < 4; k++)
	{
		r[2 + k] = k < cnt ? ctP[4 * int64_t(slot) + k] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		f[k] = k < cnt ? nF[int64_t(k) * C + o] : 0.0f;
	}
	r[6] = make_float4(f[0], f[1], f[2], f[3]);
	r[7] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// slots: [0, Q) body-plane pairs, then body pairs up to the device count *dP
// slots: body-plane [0, Q), body pairs [Q, pairEnd) up to the count on the
// device, body-mesh patches [pairEnd, S)
PXH_D bool liveSlot(int s, int Q, int pairEnd, const uint32_t* dP)
{
	return s < Q || s >= pairEnd || uint32_t(s - Q) < *dP;
}

__global__ void k_touch_flags(int S, int Q, int pairEnd, const uint32_t* dP, const float4* ctN, uint32_t* flags)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s < S)
		flags[s] = liveSlot(s, Q, pairEnd, dP) && __float_as_int(ctN[s].w) > 0 ? 1u : 0u;
}

// Pair identity across steps, world-relative so it survives changes to other
// worlds: world (16 bits) | local index of a (24 bits) | local index of b or
// 0x800000 | plane index (24 bits).
PXH_D uint64_t worldPairKey(int w, uint32_t la, uint32_t lb)
{
	const uint32_t b24 = (lb & 0x80000000u) ? (0x800000u | (lb & 0x7FFFFFu)) : (lb & 0x7FFFFFu);
	return (uint64_t(uint32_t(w) & 0xFFFFu) << 48) | (uint64_t(la & 0xFFFFFFu) << 24) | b24;
}

// constraint list: touching pair slots in slot order, with the scene-local
// ordering key used by the colouring
__global__ void k_con_build(int S, int Q, int pairEnd, const uint32_t* dP, const float4* ctN, const uint32_t* scanned,
                            const uint2* pairs, const int* qBody, const int* qPlane, const int* meshBody, int nPlanes,
                            const int* local, const int* planeLocal, const int* world, const int* worldTag, int* conA,
                            int* conB, int* conSlot, uint64_t* conKey, uint64_t* conPKey, int* conColor)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= S || !(liveSlot(s, Q, pairEnd, dP) && __float_as_int(ctN[s].w) > 0))
		return;
	const int c = int(scanned[s]);
	int a, b;
	uint32_t la, lb;
	if (s >= pairEnd)
	{
		// a mesh patch: static index plane count + patch, local id 0x400000 + patch
		const int k = s - pairEnd, p = k % MESH_MAXP;
		a = meshBody[k / MESH_MAXP];
		b = -1 - (nPlanes + p);
		la = uint32_t(local[a]);
		lb = 0x80000000u | (0x400000u + uint32_t(p));
	}
	else if (s >= Q)
	{
		const uint2 ab = pairs[s - Q];
		a = int(ab.x); b = int(ab.y);
		la = uint32_t(local[a]); lb = uint32_t(local[b]);
	}
	else
	{
		a = qBody[s];
		const int pl = qPlane[s];
		b = -1 - pl;
		la = uint32_t(local[a]);
		lb = 0x80000000u | uint32_t(planeLocal[pl]);
	}
	conA[c] = a;
	conB[c] = b;
	conSlot[c] = s;
	conKey[c] = mix64((uint64_t(la) << 32) | lb);
	conPKey[c] = worldPairKey(worldTag[world[a]], la, lb);
	conColor[c] = b >= 0 ? -1 : -2;  // -2: body-plane, placed after the body-body rounds
}

// ---------------------------------------------------------------------------
// colouring: a body-body constraint joins colour r when its key is the minimum
// over the uncoloured constraints of both its bodies (deterministic, and a
// function of each world's own content); body-plane constraints then follow
// every colour of their body's body-body constraints
// ---------------------------------------------------------------------------
__global__ void k_fill_u64(uint64_t* p, int n, uint64_t v)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		p[i] = v;
}

// Colouring kernels run over the still-uncoloured constraints: list[i] when a
// compacted list is given, else i, for i below the count *mDev on the device
// (or the host bound m when mDev is null). They loop grid-stride so the grid
// can be sized for the bound while few constraints remain.
PXH_D int listed(const int* list, int i) { return list ? list[i] : i; }
PXH_D int listCount(int m, const uint32_t* mDev) { return mDev ? min(m, int(*mDev)) : m; }
#define PXH_GRID_STRIDE(i, count) \
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (count); i += gridDim.x * blockDim.x)

__global__ void k_color_reset(int m, const uint32_t* mDev, const int* list, const int* conA, const int* conB,
                              unsigned long long* best)
{
	const int mm = listCount(m, mDev);
	PXH_GRID_STRIDE(i, mm)
	{
		const int c = listed(list, i);
		best[conA[c]] = ~0ull;
		const int b = conB[c];
		if (b >= 0)
			best[b] = ~0ull;
	}
}

__global__ void k_color_min(int m, const uint32_t* mDev, const int* list, const int* conA, const int* conB,
                            const uint64_t* conKey, const int* conColor, unsigned long long* best)
{
	const int mm = listCount(m, mDev);
	PXH_GRID_STRIDE(i, mm)
	{
		const int c = listed(list, i);
		if (conColor[c] != -1)
			continue;
		const unsigned long long k = conKey[c];
		atomicMin(&best[conA[c]], k);
		const int b = conB[c];
		if (b >= 0)
			atomicMin(&best[b], k);
	}
}

// A local minimum takes the smallest colour used by neither of its bodies
// (first fit on 64-bit per-body colour masks; colours past 63 fall back to
// one colour per round, 64 + round). Minima of one round share no body, so
// their mask updates never race with each other's reads.
__global__ void k_color_pick(int m, const uint32_t* mDev, const int* list, int round, const int* conA,
                             const int* conB, const uint64_t* conKey, int* conColor, const unsigned long long* best,
                             unsigned long long* mask)
{
	const int mm = listCount(m, mDev);
	PXH_GRID_STRIDE(i, mm)
	{
		const int c = listed(list, i);
		if (conColor[c] != -1)
			continue;
		const unsigned long long k = conKey[c];
		const int a = conA[c], b = conB[c];
		if (best[a] == k && (b < 0 || best[b] == k))
		{
			const unsigned long long used = mask[a] | (b >= 0 ? mask[b] : 0ull);
			const int col = used == ~0ull ? 64 + round : __ffsll((long long)~used) - 1;
			conColor[c] = col;
			if (col < 64)
			{
				atomicOr(&mask[a], 1ull << col);
				if (b >= 0)
					atomicOr(&mask[b], 1ull << col);
			}
		}
	}
}

// flags[i] for i <= m (flags[m] = 0 ends the scan)
__global__ void k_color_uncolored(int m, const uint32_t* mDev, const int* list, const int* conColor, uint32_t* flags)
{
	const int mm = listCount(m, mDev);
	PXH_GRID_STRIDE(i, m + 1)
		flags[i] = i < mm && conColor[listed(list, i)] == -1 ? 1u : 0u;
}

__global__ void k_color_compact(int m, const int* list, const uint32_t* flags, const uint32_t* scanned, int* out)
{
	PXH_GRID_STRIDE(i, m)
		if (flags[i])
			out[scanned[i]] = listed(list, i);
}

// Body-plane constraints go after every colour of their body's body-body
// constraints, one colour each in constraint order (the reference places
// static constraints at maxSolverNormalProgress + k:
// lowleveldynamics/src/DyConstraintPartition.cpp:168-179, 604-616).
__global__ void k_body_next(int C, const int* conA, const int* conB, const int* conColor, uint32_t* bodyNext)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= C || conB[c] < 0)
		return;
	const uint32_t nx = uint32_t(conColor[c] + 1);
	atomicMax(&bodyNext[conA[c]], nx);
	atomicMax(&bodyNext[conB[c]], nx);
}

// A body's body-plane constraints are consecutive in the constraint list
// (slots are body-major); k counts the ones before c.
__global__ void k_static_color(int C, const int* conA, const int* conB, const uint32_t* bodyNext, int* conColor)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= C || conB[c] >= 0)
		return;
	const int a = conA[c];
	int k = 0;
	for (int e = c - 1; e >= 0 && conA[e] == a && conB[e] < 0; e--)
		k++;
	conColor[c] = int(bodyNext[a]) + k;
}

// open-addressing table of the previous step's pair keys -> row (keys are
// unique; the table has at least twice as many slots as keys)
PXH_D uint32_t patchSlot(uint64_t key, uint32_t mask) { return uint32_t(mix64(key)) & mask; }

__global__ void k_patch_insert(int m, const uint64_t* keys, unsigned long long* htKey, int* htVal, uint32_t mask)
{
	const int o = blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= m)
		return;
	const unsigned long long key = keys[o];
	uint32_t slot = patchSlot(key, mask);
	for (uint32_t probe = 0; probe <= mask; probe++)
	{
		if (atomicCAS(&htKey[slot], ~0ull, key) == ~0ull)
		{
			htVal[slot] = o;
			return;
		}
		slot = (slot + 1) & mask;
	}
}

PXH_D int patchFind(const PatchIO& io, uint64_t key)
{
	if (!io.htKey)
		return -1;
	uint32_t slot = patchSlot(key, io.htMask);
	for (uint32_t probe = 0; probe <= io.htMask; probe++)
	{
		const uint64_t k = io.htKey[slot];
		if (k == key)
			return io.htVal[slot];
		if (k == ~0ull)
			return -1;
		slot = (slot + 1) & io.htMask;
	}
	return -1;
}


// A body-body constraint that existed in the previous step keeps its colour
// (below 64): those constraints were coloured together then, so they never
// share a colour at a body now. The reference GPU solver keeps its partitions
// from step to step the same way (gpusolver/src/PxgConstraintPartition.cpp:1940).
__global__ void k_con_prev(int C, const int* conA, const int* conB, const int* conSlot, PatchIO io, int* conColor,
                           unsigned long long* mask)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= C)
		return;
	const int a = conA[c], b = conB[c];
	const uint64_t pkey = io.conPKey[c];
	const int po = patchFind(io, pkey);
	if (io.prevRow)
		io.prevRow[c] = po;
	io.conRec[2 * int64_t(c)] = make_int4(a, b, conSlot[c], po);
	io.conRec[2 * int64_t(c) + 1] = make_int4(int(uint32_t(pkey)), int(uint32_t(pkey >> 32)), 0, 0);
	if (b >= 0 && po >= 0)
	{
		const int pc = io.prevColor[po];
		if (pc >= 0 && pc < 64)
		{
			conColor[c] = pc;
			atomicOr(&mask[a], 1ull << pc);
			atomicOr(&mask[b], 1ull << pc);
		}
	}
}

// block-level maximum then one atomic per block (grid-stride)
__global__ void __launch_bounds__(256) k_max_color(int C, const int* conColor, uint32_t* maxColor)
{
	__shared__ int red[256];
	int v = 0;
	for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < C; c += gridDim.x * blockDim.x)
		v = max(v, conColor[c]);
	red[threadIdx.x] = v;
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1)
	{
		if (threadIdx.x < o)
			red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + o]);
		__syncthreads();
	}
	if (threadIdx.x == 0)
		atomicMax(maxColor, uint32_t(red[0]));
}

// colour-ordered constraint list by a stable radix sort of (colour, index)
// A constraint not coloured yet (-1, when the rounds have not finished) sorts
// to colour `limit`, past every real colour.
__global__ void k_color_keys(int C, const int* conColor, uint32_t limit, uint64_t* keys, uint32_t* vals)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < C)
	{
		const int col = conColor[c];
		const uint32_t k = col < 0 ? limit : min(uint32_t(col), limit);
		keys[c] = (uint64_t(k) << 32) | uint32_t(c);
		vals[c] = uint32_t(c);
	}
}
// rows ordered by (world, colour): key world * CB + colour, CB = limit + 1
__global__ void k_color_keys_world(int C, const int* conColor, const int* conA, const int* bodyWorld,
                                   uint32_t limit, uint64_t* keys, uint32_t* vals)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c < C)
	{
		const int col = conColor[c];
		const uint32_t k = col < 0 ? limit : min(uint32_t(col), limit);
		const uint64_t w = uint64_t(bodyWorld[conA[c]]);
		keys[c] = ((w * (limit + 1) + k) << 32) | uint32_t(c);
		vals[c] = uint32_t(c);
	}
}

// ---------------------------------------------------------------------------
// solver
// ---------------------------------------------------------------------------

// bodyCoreComputeUnconstrainedVelocity (lowleveldynamics/src/DyBodyCoreIntegrator.h:40-75)
// + copyToSolverBodyDataStep (lowleveldynamics/src/DyTGSDynamics.cpp:154-242)
__global__ void k_preintegrate(int n, const float4* pos, const float4* rot, const float4* lin, const float4* ang,
                               const float4* invI, float3 gdt, float linDampDt, float angDampDt, float maxLinSq,
                               float maxAngSq, int pgs, float4* vel, float4* svP, float4* svDQ, float4* svI0,
                               float4* svI1, float4* svOrigLin, float4* rec, float dt, float4* extLin, float4* extAng,
                               float4* prevLin, float4* prevAng, const uint32_t* sleepFlags)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	const float4 l = lin[i];
	f3 lv = mk(l);
	f3 av = mk(ang[i]);
	// velocities before the step, for the acceleration reads
	if (prevLin)
	{
		prevLin[i] = f4(lv, 0.0f);
		prevAng[i] = f4(av, 0.0f);
	}
	// external accelerations (forces and torques set through the Direct GPU
	// API) change the velocities first, then gravity (BodySim::updateForces,
	// simulationcontroller/src/ScBodySim.cpp:656-713, before
	// bodyCoreComputeUnconstrainedVelocity); they last one step
	if (extLin)
	{
		const float4 el = extLin[i], ea = extAng[i];
		if (el.x != 0.0f || el.y != 0.0f || el.z != 0.0f)
			lv = lv + mk(el) * dt;
		if (ea.x != 0.0f || ea.y != 0.0f || ea.z != 0.0f)
			av = av + mk(ea) * dt;
		extLin[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		extAng[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	}
	// sleeping bodies are not in the solver: no gravity, velocities stay zero
	if (!sleepFlags || !(sleepFlags[i] & SLEEP_ASLEEP))
		lv = lv + mk(gdt.x, gdt.y, gdt.z);
	const float ld = 1.0f - linDampDt, ad = 1.0f - angDampDt;
	lv = lv * (ld >= 0.0f ? ld : 0.0f);
	av = av * (ad >= 0.0f ? ad : 0.0f);
	const float l2 = dot(lv, lv);
	if (l2 > maxLinSq) lv = lv * sqrtf(maxLinSq / l2);
	const float a2 = dot(av, av);
	if (a2 > maxAngSq) av = av * sqrtf(maxAngSq / a2);
	const m33 R = qmat(mkq(rot[i]));
	const float4 ii = invI[i];
	const f3 si = mk(sqrtf(ii.x), sqrtf(ii.y), sqrtf(ii.z));
	const f3 sI = mk(si.x == 0.0f ? 0.0f : 1.0f / si.x, si.y == 0.0f ? 0.0f : 1.0f / si.y,
	                 si.z == 0.0f ? 0.0f : 1.0f / si.z);
	m33 A, B;
	A.c0 = R.c0 * si.x; A.c1 = R.c1 * si.y; A.c2 = R.c2 * si.z;
	B.c0 = R.c0 * sI.x; B.c1 = R.c1 * sI.y; B.c2 = R.c2 * sI.z;
	const m33 Rt = mtrans(R);
	const m33 S = mmm(A, Rt);
	const m33 SI = mmm(B, Rt);
	float4* v = vel + 4 * int64_t(i);
	if (pgs)
	{
		// PGS solves for velocity changes from the pre-integrated velocities,
		// kept in the record's second half (copyToSolverBodyData,
		// lowleveldynamics/src/DyRigidBodyToSolverBody.cpp:37-113)
		v[0] = make_float4(0.0f, 0.0f, 0.0f, l.w);
		v[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		v[2] = f4(lv, 0.0f);
		v[3] = f4(av, 0.0f);
	}
	else
	{
		v[0] = f4(lv, l.w);
		v[1] = f4(mmul(SI, av), 0.0f);
		v[2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		v[3] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	}
	svP[i] = pos[i];
	svDQ[i] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
	// symmetric sqrt inverse inertia: (xx yy zz xy) (xz yz)
	const float4 i0 = make_float4(S.c0.x, S.c1.y, S.c2.z, S.c1.x);
	const float4 i1 = make_float4(S.c2.x, S.c2.y, 0.0f, 0.0f);
	svI0[i] = i0;
	svI1[i] = i1;
	svOrigLin[i] = f4(lv, 0.0f);
	float4* r = rec + PREP_REC * int64_t(i);
	const float4 p = pos[i];
	r[0] = make_float4(p.x, p.y, p.z, l.w);
	r[1] = rot[i];
	r[2] = i0;
	r[3] = i1;
	r[4] = f4(lv, 0.0f);
	r[5] = f4(av, 0.0f);
}

PXH_D m33 loadSym(float4 a, float4 b)
{
	m33 m;
	m.c0 = mk(a.x, a.w, b.x);
	m.c1 = mk(a.w, a.y, b.y);
	m.c2 = mk(b.x, b.y, a.z);
	return m;
}

// Friction patch of the constraint at row o (constraint c, bodies a, b,
// normal n, frames f0/f1, contact points cp/cs): the previous step's patch of
// this pair is kept when it is unbroken, has anchors, its body normals still
// agree, no anchor drifted apart along the normal, and its world normal
// matches this step's normal (getFrictionPatches, correlatePatches,
// growPatches: lowleveldynamics/src/DyContactPrepShared.h:53-127,
// DyFrictionCorrelation.cpp:136-316). Writes the patch for the next step and
// returns the anchor count, the anchors in pa0/pa1 (body space). Shared by the
// TGS and PGS preparation.
PXH_D int frictionPatch(int o, int c, int po, uint64_t pkey, f3 n, const xf& f0, const xf& f1, const f3 (&cp)[4],
                        const float (&cs)[4], int cnt, float restDist, float frictionOffsetThreshold,
                        const PatchIO& io, f3 (&pa0)[2], f3 (&pa1)[2])
{
	// friction patch: the previous step's patch of this pair is kept when it is
	// unbroken, has anchors, its body normals still agree, no anchor drifted
	// apart along the normal, and its world normal matches this step's normal
	f3 pn0, pn1;
	int nKept = 0;
	{
		bool kept = false;
		if (po >= 0 && !io.prevBroken[po])
		{
			const float4 q0 = io.prev[6 * int64_t(po) + 0];
			const int pc = __float_as_int(q0.w);
			if (pc != 0)
			{
				const float4 q1 = io.prev[6 * int64_t(po) + 1];
				pn0 = mk(q0); pn1 = mk(q1);
				pa0[0] = mk(io.prev[6 * int64_t(po) + 2]); pa0[1] = mk(io.prev[6 * int64_t(po) + 3]);
				pa1[0] = mk(io.prev[6 * int64_t(po) + 4]); pa1[1] = mk(io.prev[6 * int64_t(po) + 5]);
				const xf b10 = xinvmul(f0, f1);
				if (dot(pn0, qrot(b10.q, pn1)) > 0.999f)
				{
					bool separated = false;
#pragma unroll
					for (int j = 0; j < 2; j++)
						if (j < pc && !(fabsf(dot(pa0[j] - xtrans(b10, pa1[j]), pn0)) < io.corrDist))
							separated = true;
					if (!separated && !(dot(n, qrot(f0.q, pn0)) < 0.999f))
					{
						kept = true;
						nKept = pc;
					}
				}
			}
		}
		if (!kept)
		{
			pn0 = qrotinv(f0.q, n);
			pn1 = qrotinv(f1.q, n);
		}
	}
	// anchors: two kept anchors stay unless closer than half the diagonal of
	// this step's contact bounds; one kept anchor is completed from the
	// contacts; otherwise up to two points closer than the offset threshold
	if (nKept == 2)
	{
		f3 mn = cp[0], mx = cp[0];
#pragma unroll
		for (int j = 1; j < 4; j++)
		{
			if (j >= cnt)
				break;
			mn = mk(fminf(mn.x, cp[j].x), fminf(mn.y, cp[j].y), fminf(mn.z, cp[j].z));
			mx = mk(fmaxf(mx.x, cp[j].x), fmaxf(mx.y, cp[j].y), fmaxf(mx.z, cp[j].z));
		}
		const f3 dg = mx - mn, da = pa0[0] - pa0[1];
		if (!(dot(da, da) * 4.0f >= dot(dg, dg)))
			nKept = 0;
	}
	int nAnch = nKept;
	if (nKept != 2)
	{
		f3 wa0 = mk(0, 0, 0), wa1 = mk(0, 0, 0);
		nAnch = 0;
		if (nKept == 1)
		{
			wa0 = xtrans(f0, pa0[0]);
			nAnch = 1;
		}
		float pd2 = 0.0f;
		const float thr = frictionOffsetThreshold + restDist;
#pragma unroll
		for (int j = 0; j < 4; j++)
		{
			if (j >= cnt || !(cs[j] < thr))
				continue;
			const f3 wp = cp[j];
			if (nAnch == 0) { wa0 = wp; nAnch = 1; }
			else if (nAnch == 1)
			{
				const f3 d = wp - wa0;
				pd2 = dot(d, d);
				if (pd2 > 1e-8f) { wa1 = wp; nAnch = 2; }
			}
			else
			{
				const f3 e0 = wp - wa0, e1 = wp - wa1;
				const float d0 = dot(e0, e0), d1 = dot(e1, e1);
				if (d0 > d1) { if (d0 > pd2) { wa1 = wp; pd2 = d0; } }
				else if (d1 > pd2) { wa0 = wp; pd2 = d1; }
			}
		}
		if (nKept == 0 && nAnch > 0)
		{
			pa0[0] = xtransinv(f0, wa0);
			pa1[0] = xtransinv(f1, wa0);
		}
		if (nAnch == 2)
		{
			pa0[1] = xtransinv(f0, wa1);
			pa1[1] = xtransinv(f1, wa1);
		}
		if (nAnch == 0)
		{
			pa0[0] = mk(0, 0, 0);
			pa1[0] = pa0[0];
		}
		if (nAnch < 2)
		{
			pa0[1] = mk(0, 0, 0);
			pa1[1] = pa0[1];
		}
	}
	io.curKey[o] = pkey;
	io.curColor[o] = io.conColor[c];
	io.cur[6 * int64_t(o) + 0] = f4(pn0, __int_as_float(nAnch));
	io.cur[6 * int64_t(o) + 1] = f4(pn1, 0.0f);
	io.cur[6 * int64_t(o) + 2] = f4(pa0[0], 0.0f);
	io.cur[6 * int64_t(o) + 3] = f4(pa0[1], 0.0f);
	io.cur[6 * int64_t(o) + 4] = f4(pa1[0], 0.0f);
	io.cur[6 * int64_t(o) + 5] = f4(pa1[1], 0.0f);
	return nAnch;
}

PXH_D void prepBody(int o, int C, const int* order, const float4* ctN, const float4* ctP, const float4* rec,
                    float invStepDt, float scale, float restDist, float frictionOffsetThreshold, Rows& rows,
                    const PatchIO& io, uint32_t& nCntOut, uint32_t& nDynOut, uint32_t& nFricOut);

// Per-block sums of four per-thread counts (256-thread blocks) stored at
// partial[4 * block]: no atomics on shared counters from every wave (with
// thousands of waves those serialise); k_sum_counters adds the blocks up.
PXH_D void blockCounters4(uint32_t v0, uint32_t v1, uint32_t v2, uint32_t v3, unsigned long long* partial)
{
	__shared__ uint32_t red[4][4];
	const uint32_t s0 = waveInclScan(v0), s1 = waveInclScan(v1), s2 = waveInclScan(v2), s3 = waveInclScan(v3);
	const int w = threadIdx.x >> 6;
	if ((threadIdx.x & 63) == 63)
	{
		red[w][0] = s0; red[w][1] = s1; red[w][2] = s2; red[w][3] = s3;
	}
	__syncthreads();
	if (threadIdx.x < 4)
	{
		unsigned long long t = 0;
		for (int k = 0; k < int(blockDim.x >> 6); k++)
			t += red[k][threadIdx.x];
		partial[4 * int64_t(blockIdx.x) + threadIdx.x] = t;
	}
}

// counters[k] = sum over blocks of partial[4 * block + k] (one block)
__global__ void __launch_bounds__(256) k_sum_counters(int nBlocks, const unsigned long long* partial,
                                                      unsigned long long* counters)
{
	__shared__ unsigned long long red[4][256];
	unsigned long long t[4] = {0, 0, 0, 0};
	for (int b = threadIdx.x; b < nBlocks; b += blockDim.x)
		for (int k = 0; k < 4; k++)
			t[k] += partial[4 * int64_t(b) + k];
	for (int k = 0; k < 4; k++)
		red[k][threadIdx.x] = t[k];
	__syncthreads();
	for (int o = 128; o > 0; o >>= 1)
	{
		if (threadIdx.x < o)
			for (int k = 0; k < 4; k++)
				red[k][threadIdx.x] += red[k][threadIdx.x + o];
		__syncthreads();
	}
	if (threadIdx.x < 4)
		counters[threadIdx.x] = red[threadIdx.x][0];
}

// Contact rows of one constraint at its colour-ordered position o
// (setupFinalizeSolverConstraints / constructContactConstraintStep,
//  lowleveldynamics/src/DyTGSContactPrep.cpp:310-404, 426-812; friction patches
//  getFrictionPatches / correlatePatches / growPatches, DyContactPrepShared.h:53-127,
//  DyFrictionCorrelation.cpp:136-316)
__global__ void k_prep(int C, const int* order, const float4* ctN, const float4* ctP, const float4* rec,
                       float invStepDt, float scale, float restDist, float frictionOffsetThreshold, Rows rows,
                       PatchIO io, unsigned long long* counters)
{
	const int o = blockIdx.x * blockDim.x + threadIdx.x;
	// per-step counters: contact points, friction rows, dynamic-dynamic constraints
	uint32_t myCnt = 0, myFric = 0, myDyn = 0;
	prepBody(o, C, order, ctN, ctP, rec, invStepDt, scale, restDist, frictionOffsetThreshold, rows, io, myCnt, myDyn,
	         myFric);
	blockCounters4(myCnt, myFric, myDyn, myDyn ? myFric : 0u, counters);
}

PXH_D void prepBody(int o, int C, const int* order, const float4* ctN, const float4* ctP, const float4* rec,
                    float invStepDt, float scale, float restDist, float frictionOffsetThreshold, Rows& rows,
                    const PatchIO& io, uint32_t& nCntOut, uint32_t& nDynOut, uint32_t& nFricOut)
{
	if (o >= C)
		return;
	const int c = order[o];
	const int4 cr = io.conRec[2 * int64_t(c)], ck = io.conRec[2 * int64_t(c) + 1];
	const int a = cr.x, b = cr.y, slot = cr.z, po = cr.w;
	const uint64_t pkey = (uint64_t(uint32_t(ck.y)) << 32) | uint32_t(ck.x);
	const float4 nn = ctN[slot];
	const int cnt = __float_as_int(nn.w);
	const f3 n = mk(nn);
	const bool dynB = b >= 0;
	nCntOut = uint32_t(cnt);
	nDynOut = dynB ? 1u : 0u;
	const float4* ra4 = rec + PREP_REC * int64_t(a);
	const float4 pa4 = ra4[0];
	const m33 Sa = loadSym(ra4[2], ra4[3]);
	const f3 linA = mk(ra4[4]);
	m33 Sb;
	Sb.c0 = mk(0, 0, 0); Sb.c1 = Sb.c0; Sb.c2 = Sb.c0;
	float invMassB = 0.0f;
	xf f0, f1;
	f0.p = mk(pa4); f0.q = mkq(ra4[1]);
	f1.p = mk(0, 0, 0); f1.q.x = 0; f1.q.y = 0; f1.q.z = 0; f1.q.w = 1;
	f3 linB = mk(0, 0, 0);
	if (dynB)
	{
		const float4* rb4 = rec + PREP_REC * int64_t(b);
		const float4 pb4 = rb4[0];
		Sb = loadSym(rb4[2], rb4[3]);
		invMassB = pb4.w;
		f1.p = mk(pb4); f1.q = mkq(rb4[1]);
		linB = mk(rb4[4]);
	}
	const float invMassA = pa4.w;
	const float nlen2 = dot(n, n);
	const float imn0 = invMassA * nlen2, imn1 = -invMassB * nlen2;
	const float invDtp8 = invStepDt * scale;
	// loops run to the fixed bound 4 with guards so the point arrays stay in
	// registers (a loop to cnt would index them dynamically, in scratch)
	f3 cp[4];
	float cs[4];
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		cp[k] = mk(0, 0, 0);
		cs[k] = 0.0f;
		if (k >= cnt)
			continue;
		const float4 q = ctP[4 * int64_t(slot) + k];
		cp[k] = mk(q);
		cs[k] = q.w;
		const f3 ra = cp[k] - f0.p, rb = cp[k] - f1.p;
		const f3 raXnI = mmul(Sa, cross(ra, n)), rbXnI = mmul(Sb, cross(rb, n));
		const float unit = (imn0 + dot(raXnI, raXnI)) + (dot(rbXnI, rbXnI) - imn1);
		const float pen = q.w - restDist;
		const float recip = unit > 0.0f ? 1.0f / unit : 0.0f;
		rows.nA[k * int64_t(C) + o] = f4(raXnI, pen);
		// the bias coefficient is -invStepDt for a separated point, else -invStepDt*scale:
		// the sign of the stored reciprocal response carries that choice
		rows.nB[k * int64_t(C) + o] = f4(rbXnI, pen > 0.0f ? -recip : recip);
		rows.nF[k * int64_t(C) + o] = 0.0f;
	}
	f3 pa0[2], pa1[2];
	const int nAnch =
	    frictionPatch(o, c, po, pkey, n, f0, f1, cp, cs, cnt, restDist, frictionOffsetThreshold, io, pa0, pa1);
	nFricOut = uint32_t(2 * nAnch);
	(void)invDtp8;
	rows.hI[o] = make_int4(a, b, cnt | ((2 * nAnch) << 8), 0);
	rows.hF[o] = make_float4(n.x, n.y, n.z, 0.0f);
	if (nAnch == 0)
		return;
	const f3 lvr = linA - linB;
	const f3 fb = (0.70710678f > fabsf(n.x)) ? mk(0.0f, -n.z, n.y) : mk(-n.y, n.x, 0.0f);
	f3 t0 = lvr - n * dot(n, lvr);
	t0 = dot(t0, t0) > 0.0001f ? t0 : fb;
	t0 = normalize(t0);
	const f3 t1 = normalize(cross(n, t0));
	const f3 relTr = f0.p - f1.p;
#pragma unroll
	for (int j = 0; j < 2; j++)
	{
		if (j >= nAnch)
			break;
		const f3 ra = qrot(f0.q, pa0[j]), rb = qrot(f1.q, pa1[j]);
		const f3 err = (ra - rb) + relTr;
		for (int d = 0; d < 2; d++)
		{
			const f3 t = d == 0 ? t0 : t1;
			const f3 raXnI = mmul(Sa, cross(ra, t)), rbXnI = mmul(Sb, cross(rb, t));
			const float unit = (imn0 + dot(raXnI, raXnI)) + (dot(rbXnI, rbXnI) - imn1);
			const int r = 2 * j + d;
			rows.fT[r * int64_t(C) + o] = f4(t, dot(err, t));
			rows.fA[r * int64_t(C) + o] = f4(raXnI, unit > 0.0f ? scale / unit : 0.0f);
			if (dynB)
				rows.fB[r * int64_t(C) + o] = f4(rbXnI, 0.0f);
			rows.fF[r * int64_t(C) + o] = 0.0f;
		}
	}
}

// one colour: each lane solves one constraint (solveContact /
// solveDynamicContactsStep, lowleveldynamics/src/DyTGSContactPrep.cpp:1490-1843).
// All rows of the lane are loaded before any arithmetic so a wave keeps every
// gather in flight at once; 64-lane blocks spread a colour over all CUs.
constexpr int SOLVE_THREADS = 64;
constexpr int SOLVE_WORLD_THREADS = 64;

// One TGS contact row o. Body velocity records are at vel[4 * (index -
// bodyBase)] (bodyBase 0: the scene-wide array; a world's first body: that
// world's records staged in LDS); the arithmetic is the same either way.
__device__ __forceinline__ void solveContactRowTGS(int o, int C, const Rows& rows, float4* vel, int bodyBase,
                                                   float minPen, float muS, float muD, float invStepDt,
                                                   float biasScale, float scale, uint32_t* broken)
{
	const int4 hi = rows.hI[o];
	const float4 hf = rows.hF[o];
	const int a = hi.x - bodyBase, b = hi.y < 0 ? hi.y : hi.y - bodyBase;
	const int cnt = hi.z & 0xFF, nF = hi.z >> 8;
	const bool dynB = hi.y >= 0;
	// gathers: both bodies
	const float4* va = vel + 4 * int64_t(a);
	const float4 la4 = va[0], aa4 = va[1], dla4 = va[2], am04 = va[3];
	float4 lb4 = make_float4(0, 0, 0, 0), ab4 = lb4, am14 = lb4, dlb4 = lb4;
	if (dynB)
	{
		const float4* vb = vel + 4 * int64_t(b);
		lb4 = vb[0]; ab4 = vb[1]; dlb4 = vb[2]; am14 = vb[3];
	}
	// rows
	float4 nA[4], nB[4], fT[4], fA[4], fB[4];
	float nFc[4], fFc[4];
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		const int64_t ri = k * int64_t(C) + o;
		if (k < cnt)
		{
			nA[k] = rows.nA[ri]; nB[k] = rows.nB[ri]; nFc[k] = rows.nF[ri];
		}
		if (k < nF)
		{
			fT[k] = rows.fT[ri]; fA[k] = rows.fA[ri]; fFc[k] = rows.fF[ri];
			fB[k] = dynB ? rows.fB[ri] : make_float4(0, 0, 0, 0);
		}
	}
	const f3 n = mk(hf);
	const float invMassA = la4.w, invMassB = lb4.w;
	f3 l0 = mk(la4), a0 = mk(aa4), l1 = mk(lb4), a1 = mk(ab4);
	const f3 am0 = mk(am04), am1 = mk(am14);
	const f3 rel = mk(dla4) - mk(dlb4);
	const f3 nim0 = n * invMassA, nim1 = n * invMassB;
	const float dv = dot(rel, n);
	const float biasSep = -invStepDt, biasPen = -invStepDt * scale;
	float acc = 0.0f;
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		if (k >= cnt)
			break;
		const f3 raXnI = mk(nA[k]), rbXnI = mk(nB[k]);
		const float rs = nB[k].w;
		const float recip = fabsf(rs);
		const float biasCoef = rs < 0.0f ? biasSep : biasPen;
		const float dang = dot(am0, raXnI) - dot(am1, rbXnI);
		const float sep = fmaxf(minPen, nA[k].w + (dv + dang));
		const float bias = fminf(1e32f, biasCoef * sep);
		const f3 v0 = mulv(l0, n) + mulv(a0, raXnI);
		const f3 v1 = mulv(l1, n) + mulv(a1, rbXnI);
		const f3 dvv = v0 - v1;
		const float nv = dvv.x + dvv.y + dvv.z;
		const float lambda = bias * recip - nv * recip;
		const float af = nFc[k];
		const float nf = fminf(af + fmaxf(lambda, -af), 1e32f);
		const float df = nf - af;
		l0 = nim0 * df + l0;
		l1 = l1 - nim1 * df;
		a0 = raXnI * df + a0;
		a1 = a1 - rbXnI * df;
		nFc[k] = nf;
		acc = acc + nf;
	}
	bool anyClamp = false;
	if (nF)
	{
		const float maxF = muS * acc, maxD = muD * acc;
		const float fscale = nF == 4 ? 0.5f : 1.0f;
#pragma unroll
		for (int i = 0; i < 4; i += 2)
		{
			if (i >= nF)
				break;
			float ti[2];
#pragma unroll
			for (int d = 0; d < 2; d++)
			{
				const f3 t = mk(fT[i + d]), raI = mk(fA[i + d]), rbI = mk(fB[i + d]);
				const float dvf = (dot(raI, am0) - dot(rbI, am1)) + dot(t, rel);
				const float bias = (fT[i + d].w + dvf) * biasScale;
				const f3 v0 = mulv(l0, t) + mulv(a0, raI);
				const f3 v1 = mulv(l1, t) + mulv(a1, rbI);
				const f3 dd = v0 - v1;
				const float nv = dd.x + dd.y + dd.z;
				const float vm = fA[i + d].w;
				ti[d] = (fFc[i + d] - bias * vm) - nv * vm;
			}
			const float tot = sqrtf(ti[0] * ti[0] + ti[1] * ti[1]);
			const bool clamp = tot > fscale * maxF;
			const float tc = clamp ? fminf(fscale * maxD, tot) : tot;
			const float ratio = tot > 0.0f ? tc / tot : 0.0f;
			const float n0 = ti[0] * ratio, n1 = ti[1] * ratio;
			const float d0 = n0 - fFc[i], d1 = n1 - fFc[i + 1];
			const f3 t0 = mk(fT[i]), t1 = mk(fT[i + 1]);
			l0 = (t0 * invMassA) * d0 + ((t1 * invMassA) * d1 + l0);
			l1 = (l1 - (t1 * invMassB) * d1) - (t0 * invMassB) * d0;
			a0 = mk(fA[i]) * d0 + (mk(fA[i + 1]) * d1 + a0);
			a1 = (a1 - mk(fB[i + 1]) * d1) - mk(fB[i]) * d0;
			fFc[i] = n0;
			fFc[i + 1] = n1;
			anyClamp |= clamp;
		}
	}
	// the broken flag of the friction patch is the last pass's clamp state
	// (DyTGSContactPrep.cpp:1837, 1907)
	if (broken)
		broken[o] = anyClamp ? 1u : 0u;
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		const int64_t ri = k * int64_t(C) + o;
		if (k < cnt)
			rows.nF[ri] = nFc[k];
		if (k < nF)
			rows.fF[ri] = fFc[k];
	}
	vel[4 * int64_t(a)] = f4(l0, invMassA);
	vel[4 * int64_t(a) + 1] = f4(a0, 0.0f);
	if (dynB)
	{
		vel[4 * int64_t(b)] = f4(l1, invMassB);
		vel[4 * int64_t(b) + 1] = f4(a1, 0.0f);
	}
}

__global__ void __launch_bounds__(SOLVE_THREADS) k_solve(int begin, int end, int C, Rows rows, float4* vel,
                                                          float minPen, float muS, float muD, float invStepDt,
                                                          float biasScale, float scale, uint32_t* broken)
{
	const int o = begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= end)
		return;
	solveContactRowTGS(o, C, rows, vel, 0, minPen, muS, muD, invStepDt, biasScale, scale, broken);
}

// ---------------------------------------------------------------------------
// PGS (DynamicsContext, lowleveldynamics/src/DyDynamics.cpp; DySolverControl.cpp)
// Rows (PGS use): hI.w = dynamic friction (scaled for the anchor count, float
// bits), hF.w = static friction (scaled); nA = raXnI | biased error; nB =
// rbXnI | velocity multiplier; unbiased error in its own array; fT = t |
// bias - target velocity; fA = raXnI | velocity multiplier; fB = rbXnI |
// -target velocity.
// ---------------------------------------------------------------------------

// setupFinalizeSolverConstraints (lowleveldynamics/src/DyContactPrep.cpp:49-357)
// with constructContactConstraint (DyContactPrepShared.h:298-393)
__global__ void k_prep_pgs(int C, const int* order, const float4* ctN, const float4* ctP, const float4* rec,
                           float invDt, float muS,
                           float muD, float restDist, float frictionOffsetThreshold, Rows rows, float* rowsU,
                           PatchIO io, unsigned long long* counters)
{
	const int o = blockIdx.x * blockDim.x + threadIdx.x;
	uint32_t myCnt = 0, myFric = 0, myDyn = 0;
	if (o < C)
	{
		const int c = order[o];
		const int4 cr = io.conRec[2 * int64_t(c)], ck = io.conRec[2 * int64_t(c) + 1];
		const int a = cr.x, b = cr.y, slot = cr.z, po = cr.w;
		const uint64_t pkey = (uint64_t(uint32_t(ck.y)) << 32) | uint32_t(ck.x);
		const float4 nn = ctN[slot];
		const int cnt = __float_as_int(nn.w);
		const f3 n = mk(nn);
		const bool dynB = b >= 0;
		myCnt = uint32_t(cnt);
		myDyn = dynB ? 1u : 0u;
		const float4* ra4 = rec + PREP_REC * int64_t(a);
		const float4 pa4 = ra4[0];
		const m33 Sa = loadSym(ra4[2], ra4[3]);
		m33 Sb;
		Sb.c0 = mk(0, 0, 0); Sb.c1 = Sb.c0; Sb.c2 = Sb.c0;
		xf f0, f1;
		f0.p = mk(pa4); f0.q = mkq(ra4[1]);
		f1.p = mk(0, 0, 0); f1.q.x = 0; f1.q.y = 0; f1.q.z = 0; f1.q.w = 1;
		const float invMassA = pa4.w;
		const f3 v0 = mk(ra4[4]), w0 = mk(ra4[5]);
		float invMassB = 0.0f;
		f3 v1 = mk(0, 0, 0), w1 = mk(0, 0, 0);
		if (dynB)
		{
			const float4* rb4 = rec + PREP_REC * int64_t(b);
			const float4 pb4 = rb4[0];
			Sb = loadSym(rb4[2], rb4[3]);
			invMassB = pb4.w;
			v1 = mk(rb4[4]);
			w1 = mk(rb4[5]);
			f1.p = mk(pb4); f1.q = mkq(rb4[1]);
		}
		const float nl2 = dot(n, n);
		const float imn0 = invMassA * nl2, imn1 = -invMassB * nl2;
		const float invDtp8 = invDt * 0.8f;
		const float norVel = (n.x * v0.x - n.x * v1.x) + (n.y * v0.y - n.y * v1.y) + (n.z * v0.z - n.z * v1.z);
		f3 cp[4];
		float cs[4];
#pragma unroll
		for (int k = 0; k < 4; k++)
		{
			cp[k] = mk(0, 0, 0);
			cs[k] = 0.0f;
			if (k >= cnt)
				continue;
			const float4 q = ctP[4 * int64_t(slot) + k];
			cp[k] = mk(q);
			cs[k] = q.w;
			const f3 ra = cp[k] - f0.p, rb = cp[k] - f1.p;
			const f3 raXn = cross(ra, n), rbXn = cross(rb, n);
			const float vrel = norVel + (dot(raXn, w0) - dot(rbXn, w1));
			const f3 raXnI = mmul(Sa, raXn), rbXnI = mmul(Sb, rbXn);
			const float unit = (imn0 + dot(raXnI, raXnI)) + (dot(rbXnI, rbXnI) - imn1);
			const float pen = q.w - restDist;
			const float target = 0.0f - vrel;
			const float recip = unit > 0.0f ? 1.0f / unit : 0.0f;
			const float penScaled = pen >= 0.0f ? pen * invDt : pen * invDtp8;
			const float scaledBias = recip * fmaxf(-1e32f, penScaled);
			const int64_t ri = k * int64_t(C) + o;
			rows.nA[ri] = f4(raXnI, target * recip + -scaledBias);
			rows.nB[ri] = f4(rbXnI, recip);
			rowsU[ri] = target * recip + -fmaxf(scaledBias, 0.0f);
			rows.nF[ri] = 0.0f;
		}
		f3 pa0[2], pa1[2];
		const int nAnch =
		    frictionPatch(o, c, po, pkey, n, f0, f1, cp, cs, cnt, restDist, frictionOffsetThreshold, io, pa0, pa1);
		myFric = uint32_t(2 * nAnch);
		const float fc = nAnch == 2 ? 0.5f : 1.0f;
		rows.hI[o] = make_int4(a, b, cnt | ((2 * nAnch) << 8), __float_as_int(muD * fc));
		rows.hF[o] = make_float4(n.x, n.y, n.z, muS * fc);
		if (nAnch > 0)
		{
			const f3 lvr = v0 - v1;
			const f3 fb = (0.70710678f > fabsf(n.x)) ? mk(0.0f, -n.z, n.y) : mk(-n.y, n.x, 0.0f);
			f3 t0 = lvr - n * dot(n, lvr);
			t0 = dot(t0, t0) > 0.0001f ? t0 : fb;
			t0 = normalize(t0);
			const f3 t1 = cross(n, t0);
#pragma unroll
			for (int j = 0; j < 2; j++)
			{
				if (j >= nAnch)
					break;
				const f3 ra = qrot(f0.q, pa0[j]), rb = qrot(f1.q, pa1[j]);
				const f3 err = (ra + f0.p) - (rb + f1.p);
#pragma unroll
				for (int d = 0; d < 2; d++)
				{
					const f3 t = d == 0 ? t0 : t1;
					const f3 raXn = cross(ra, t), rbXn = cross(rb, t);
					const f3 raXnI = mmul(Sa, raXn), rbXnI = mmul(Sb, rbXn);
					const float resp = (invMassA + dot(raXnI, raXnI)) + (dot(rbXnI, rbXnI) - -invMassB);
					const float vrel = (dot(t, v0) + dot(raXn, w0)) - (dot(t, v1) + dot(rbXn, w1));
					const float target = 0.0f - vrel;
					const float bias = dot(t, err) * invDt;
					const int64_t ri = (2 * j + d) * int64_t(C) + o;
					rows.fT[ri] = f4(t, bias - target);
					rows.fA[ri] = f4(raXnI, resp > 0.0f ? 0.8f / resp : 0.0f);
					rows.fB[ri] = f4(rbXnI, 0.0f - target);
					rows.fF[ri] = 0.0f;
				}
			}
		}
	}
	blockCounters4(myCnt, myFric, myDyn, myDyn ? myFric : 0u, counters);
}

// one colour of PGS (solveContact, lowleveldynamics/src/DySolverConstraints.cpp:221-371;
// normal rows solveDynamicContacts, DySolverConstraintsShared.h:49-105) on the
// velocity changes in vel[4i], vel[4i+1]. velocityPass: rows concluded after the
// last position pass use their unbiased error and friction without bias
// (concludeContact, DySolverConstraints.cpp:508-549).
__global__ void __launch_bounds__(SOLVE_THREADS) k_solve_pgs(int begin, int end, int C, Rows rows,
                                                              const float* rowsU, float4* vel, int doFriction,
                                                              int velocityPass, uint32_t* broken)
{
	const int o = begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (o >= end)
		return;
	const int4 hi = rows.hI[o];
	const float4 hf = rows.hF[o];
	const int a = hi.x, b = hi.y;
	const int cnt = hi.z & 0xFF, nF = doFriction ? hi.z >> 8 : 0;
	const bool dynB = b >= 0;
	const float4* va = vel + 4 * int64_t(a);
	const float4 la4 = va[0], aa4 = va[1];
	float4 lb4 = make_float4(0, 0, 0, 0), ab4 = lb4;
	if (dynB)
	{
		const float4* vb = vel + 4 * int64_t(b);
		lb4 = vb[0]; ab4 = vb[1];
	}
	float4 nA[4], nB[4], fT[4], fA[4], fB[4];
	float nFc[4], fFc[4], err[4];
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		const int64_t ri = k * int64_t(C) + o;
		if (k < cnt)
		{
			nA[k] = rows.nA[ri]; nB[k] = rows.nB[ri]; nFc[k] = rows.nF[ri];
			err[k] = velocityPass ? rowsU[ri] : nA[k].w;
		}
		if (k < nF)
		{
			fT[k] = rows.fT[ri]; fA[k] = rows.fA[ri]; fB[k] = rows.fB[ri]; fFc[k] = rows.fF[ri];
		}
	}
	const f3 n = mk(hf);
	const float invMassA = la4.w, invMassB = lb4.w;
	f3 l0 = mk(la4), a0 = mk(aa4), l1 = mk(lb4), a1 = mk(ab4);
	const f3 dl0 = n * invMassA, dl1 = n * invMassB;
	float acc = 0.0f;
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		if (k >= cnt)
			break;
		const f3 raXnI = mk(nA[k]), rbXnI = mk(nB[k]);
		const f3 v0 = mulv(l0, n) + mulv(a0, raXnI);
		const f3 v1 = mulv(l1, n) + mulv(a1, rbXnI);
		const f3 dv = v0 - v1;
		const float nv = dv.x + dv.y + dv.z;
		const float af = nFc[k];
		const float d0 = fmaxf(err[k] - nv * nB[k].w, -af);
		const float nf = fminf(af + d0, 1e32f);
		const float df = nf - af;
		l0 = dl0 * df + l0;
		l1 = l1 - dl1 * df;
		a0 = raXnI * df + a0;
		a1 = a1 - rbXnI * df;
		nFc[k] = nf;
		acc = acc + nf;
	}
	bool anyClamp = false;
	if (nF)
	{
		const float maxF = hf.w * acc, maxD = __int_as_float(hi.w) * acc;
#pragma unroll
		for (int r = 0; r < 4; r++)
		{
			if (r >= nF)
				break;
			const f3 t = mk(fT[r]), raI = mk(fA[r]), rbI = mk(fB[r]);
			const f3 v0 = mulv(l0, t) + mulv(a0, raI);
			const f3 v1 = mulv(l1, t) + mulv(a1, rbI);
			const f3 dv = v0 - v1;
			const float nv = dv.x + dv.y + dv.z;
			const float vm = fA[r].w;
			const float af = fFc[r];
			const float tmp = af - (velocityPass ? fB[r].w : fT[r].w) * vm;
			const float total = tmp - nv * vm;
			const bool clamp = fabsf(total) > maxF;
			const float tc = fminf(maxD, fmaxf(-maxD, total));
			const float nf = clamp ? tc : total;
			anyClamp |= clamp;
			const float df = nf - af;
			l0 = (t * invMassA) * df + l0;
			l1 = l1 - (t * invMassB) * df;
			a0 = raI * df + a0;
			a1 = a1 - rbI * df;
			fFc[r] = nf;
		}
	}
	if (broken)
		broken[o] = anyClamp ? 1u : 0u;
#pragma unroll
	for (int k = 0; k < 4; k++)
	{
		const int64_t ri = k * int64_t(C) + o;
		if (k < cnt)
			rows.nF[ri] = nFc[k];
		if (k < nF)
			rows.fF[ri] = fFc[k];
	}
	vel[4 * int64_t(a)] = f4(l0, invMassA);
	vel[4 * int64_t(a) + 1] = f4(a0, 0.0f);
	if (dynB)
	{
		vel[4 * int64_t(b)] = f4(l1, invMassB);
		vel[4 * int64_t(b) + 1] = f4(a1, 0.0f);
	}
}

// velocity changes after the position passes move the bodies (saveMotionVelocities,
// lowleveldynamics/src/DySolverCore.cpp:40-55)
__global__ void k_pgs_save_motion(int n, const float4* vel, float4* motionLin, float4* motionAng)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	motionLin[i] = vel[4 * int64_t(i)];
	motionAng[i] = vel[4 * int64_t(i) + 1];
}

// updateWakeCounter without stabilization and the readiness it reports
// (lowleveldynamics/src/DySleep.cpp:173-235) for body i after its step: ml,
// ma its motion velocities (world frame), q its new orientation, invMass and
// invI (mass-space inverse inertia) its mass properties
PXH_D void sleepCheck(const SleepState& ss, int i, float dt, f3 ml, f3 ma, qt q, float invMass, float4 invI)
{
	const float resetTime = 20.0f * 0.02f;
	float wc = ss.wc[i];
	if (wc < resetTime * 0.5f || wc < dt)
	{
		const f3 I = mk(invI.x > 0.0f ? 1.0f / invI.x : 1.0f, invI.y > 0.0f ? 1.0f / invI.y : 1.0f,
		                invI.z > 0.0f ? 1.0f / invI.z : 1.0f);
		const f3 l = mk(ss.accL[i]) + ml;
		const f3 a = mk(ss.accA[i]) + qrotinv(q, ma);
		const float im = invMass == 0.0f ? 1.0f : invMass;
		const float angular = (a.x * a.x * I.x + a.y * a.y * I.y + a.z * a.z * I.z) * im;
		const float linear = l.x * l.x + l.y * l.y + l.z * l.z;
		const float energy = 0.5f * (angular + linear);
		const float cluster = float(1 + ss.counted[i]);
		const float threshold = cluster * ss.threshold;
		if (energy >= threshold)
		{
			// woken by its own motion: filter reset, new wake counter; from a
			// zero counter (eACTIVATE_THIS_FRAME) no longer ready for sleep
			ss.accL[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			ss.accA[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
			const float factor = threshold == 0.0f ? 2.0f : fminf(energy / threshold, 2.0f);
			ss.wc[i] = factor * 0.5f * resetTime + dt * (cluster - 1.0f);
			if (wc == 0.0f)
				ss.flags[i] &= ~SLEEP_READY;
			return;
		}
		ss.accL[i] = f4(l, 0.0f);
		ss.accA[i] = f4(a, 0.0f);
	}
	wc = fmaxf(wc - dt, 0.0f);
	ss.wc[i] = wc;
	if (wc == 0.0f)
	{
		// eDEACTIVATE_THIS_FRAME: ready for sleep, filter reset
		ss.flags[i] |= SLEEP_READY;
		ss.accL[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
		ss.accA[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
	}
}

// integrateCore (lowleveldynamics/src/DyBodyCoreIntegrator.h:83-175): poses
// from the motion velocities, body velocities from the final changes; the
// sleep check on the motion velocities (DyDynamics.cpp:1403-1420)
__global__ void k_pgs_finalize(int n, float dt, float4* pos, float4* rot, float4* lin, float4* ang,
                               const float4* invI, const float4* vel, const float4* motionLin,
                               const float4* motionAng, const float4* svI0, const float4* svI1, SleepState ss)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (ss.flags && (ss.flags[i] & SLEEP_ASLEEP))
		return;   // sleeping: pose and zero velocities kept
	const float4* v = vel + 4 * int64_t(i);
	const f3 vD = mk(v[2]), wD = mk(v[3]);
	const m33 S = loadSym(svI0[i], svI1[i]);
	const f3 lm = vD + mk(motionLin[i]);
	pos[i] = f4(mk(pos[i]) + lm * dt, pos[i].w);
	f3 am = wD + mmul(S, mk(motionAng[i]));
	float w = dot(am, am);
	if (w != 0.0f)
	{
		w = sqrtf(w);
		if (w > 1e7f)
		{
			am = normalize(am) * 1e7f;
			w = 1e7f;
		}
		const float hv = dt * w * 0.5f;
		float s, c;
		sincosf(hv, &s, &c);
		s /= w;
		const f3 pqr = am * s;
		qt qv; qv.x = pqr.x; qv.y = pqr.y; qv.z = pqr.z; qv.w = 0.0f;
		const qt q0 = mkq(rot[i]);
		qt res = qmul(qv, q0);
		res.x += q0.x * c; res.y += q0.y * c; res.z += q0.z * c; res.w += q0.w * c;
		rot[i] = f4(qnormalize(res));
	}
	const float4 l = v[0];
	lin[i] = f4(vD + mk(l), l.w);
	ang[i] = f4(wD + mmul(S, mk(v[1])), 0.0f);
	if (ss.flags)
		sleepCheck(ss, i, dt, lm, am, mkq(rot[i]), l.w, invI[i]);
}

// integrateCoreStep (lowleveldynamics/src/DyTGSDynamics.cpp:1403-1476) of
// body i whose velocity record is at v (scene-wide array or LDS)
__device__ __forceinline__ void integrateBodyTGS(int i, float dt, float4* v, float4* svP, float4* svDQ,
                                                 const float4* svI0, const float4* svI1)
{
	const f3 lv = mk(v[0]);
	const f3 av = mk(v[1]);
	const f3 delta = lv * dt;
	const f3 amv = mmul(loadSym(svI0[i], svI1[i]), av);
	const float w2 = dot(amv, amv);
	svP[i] = f4(mk(svP[i]) + delta, 0.0f);
	if (w2 != 0.0f)
	{
		const float w = sqrtf(w2);
		const float half = dt * w * 0.5f;
		float s, q;
		sincosf(half, &s, &q);
		s /= w;
		const f3 pqr = amv * s;
		qt qv; qv.x = pqr.x; qv.y = pqr.y; qv.z = pqr.z; qv.w = 0.0f;
		const qt dq = mkq(svDQ[i]);
		qt res = qmul(qv, dq);
		res.x += dq.x * q; res.y += dq.y * q; res.z += dq.z * q; res.w += dq.w * q;
		svDQ[i] = f4(qnormalize(res));
	}
	v[3] = f4(mk(v[3]) + av * dt, 0.0f);
	v[2] = f4(mk(v[2]) + delta, 0.0f);
}

__global__ void k_integrate(int n, float dt, float4* vel, float4* svP, float4* svDQ, const float4* svI0,
                            const float4* svI1)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	integrateBodyTGS(i, dt, vel + 4 * int64_t(i), svP, svDQ, svI0, svI1);
}

// The whole TGS iteration loop of independent worlds (the batched form).
// Persistent workgroups take worlds from a queue ordered by size (largest
// first); a workgroup stages its world's velocity records in LDS (worlds of
// at most ldsBodies bodies; larger ones work on the scene-wide array), then
// for every position and velocity iteration solves the world's colours in
// order, one barrier between colours, and after each position iteration
// integrates the world's bodies, as the per-colour launches and k_integrate
// do for all worlds at once. Rows are ordered by (world, colour):
// segStart[w * CB + c] is the first row of colour c of world w. Every wave
// leaves once the queue is empty. Same arithmetic in the same order as the
// per-colour path (cf. solveWholeIslandTGS, gpusolver/src/CUDA/
// solverMultiBlockTGS.cu:1686, which runs a small island in one block).
__global__ void __launch_bounds__(SOLVE_WORLD_THREADS) k_solve_worlds(
    int nWorlds, const int* worldOrder, uint32_t* queueHead, const int* bodyStart, const uint32_t* segStart, int CB,
    int C, Rows rows, float4* vel, float4* svP, float4* svDQ, const float4* svI0, const float4* svI1, int posIters,
    int velIters, float stepDt, float muS, float muD, float invStepDt, float biasScale, float scale, uint32_t* broken,
    int ldsBodies)
{
	extern __shared__ float4 lds[];
	__shared__ int sWorld;
	const int tid = threadIdx.x;
	for (;;)
	{
		if (tid == 0)
		{
			const uint32_t q = atomicAdd(queueHead, 1u);
			sWorld = q < uint32_t(nWorlds) ? worldOrder[q] : -1;
		}
		__syncthreads();
		const int w = sWorld;
		__syncthreads();
		if (w < 0)
			return;
		const int b0 = bodyStart[w], nb = bodyStart[w + 1] - b0;
		const bool inLds = nb <= ldsBodies;
		float4* V = inLds ? lds : vel + 4 * int64_t(b0);
		if (inLds)
			for (int i = tid; i < 4 * nb; i += SOLVE_WORLD_THREADS)
				lds[i] = vel[4 * int64_t(b0) + i];
		__syncthreads();
		const uint32_t* seg = segStart + int64_t(w) * CB;
		const uint32_t wEnd = seg[CB];
		for (int it = 0; it < posIters + velIters; it++)
		{
			const bool posIter = it < posIters;
			const float minPen = posIter ? -3.402823466e+38f : 0.0f;
			uint32_t* brk = it == posIters + velIters - 1 ? broken : nullptr;
			for (int c = 0; c < CB; c++)
			{
				const uint32_t s = seg[c];
				if (s >= wEnd)
					break;
				const uint32_t e = seg[c + 1];
				if (e == s)
					continue;
				for (uint32_t o = s + tid; o < e; o += SOLVE_WORLD_THREADS)
					solveContactRowTGS(int(o), C, rows, V, b0, minPen, muS, muD, invStepDt, biasScale, scale, brk);
				__syncthreads();
			}
			if (posIter)
			{
				for (int i = tid; i < nb; i += SOLVE_WORLD_THREADS)
					integrateBodyTGS(b0 + i, stepDt, V + 4 * int64_t(i), svP, svDQ, svI0, svI1);
				__syncthreads();
			}
		}
		if (inLds)
			for (int i = tid; i < 4 * nb; i += SOLVE_WORLD_THREADS)
				vel[4 * int64_t(b0) + i] = lds[i];
		__syncthreads();
	}
}

// ---------------------------------------------------------------------------
// joints (pxh_joint.h)
// ---------------------------------------------------------------------------
// jointed bodies do not collide (PxConstraintFlag::eCOLLISION_ENABLED is off
// by default, physx/include/PxConstraint.h): their pair slots lose their
// contacts before the constraint list is built
// Body-mesh narrowphase, one lane per body of a world with a mesh (k-th of
// meshBody): candidate triangles through the mesh's tree (depth-first nodes,
// each with the index after its subtree), taken in triangle order, then the
// patches of pxh_mesh.h into slots slotBase + k * MESH_MAXP + patch (deepest
// first; empty slots count 0). More than MESH_MAX_TRIS candidates set
// *overflow and leave the body without mesh contacts.
__global__ void __launch_bounds__(64) k_np_mesh(int nMB, const int* meshBody, const int* world, const int4* worldMesh,
                                                MeshSet M, const float4* pos, const float4* rot, const float4* geom,
                                                HullSet hs, float cd, int slotBase, float4* ctN, float4* ctP,
                                                uint32_t* overflow)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= nMB)
		return;
	const int i = meshBody[k];
	const int base = slotBase + k * MESH_MAXP;
	for (int p = 0; p < MESH_MAXP; p++)
		ctN[base + p].w = __int_as_float(0);
	const int4 wm = worldMesh[world[i]];
	if (wm.x < 0)
		return;
	f3 mn, mx;
	shapeBounds(i, pos, rot, geom, hs, mn, mx);
	mn = mn - mk(cd, cd, cd);
	mx = mx + mk(cd, cd, cd);
	int idx[MESH_MAX_TRIS];
	int nc = 0;
	for (int nd = wm.x; nd < wm.x + wm.y;)
	{
		const float4 lo = M.nodeLo[nd], hi = M.nodeHi[nd];
		const int4 info = M.nodeInfo[nd];
		if (lo.x > mx.x || mn.x > hi.x || lo.y > mx.y || mn.y > hi.y || lo.z > mx.z || mn.z > hi.z)
		{
			nd = info.z;
			continue;
		}
		for (int j = 0; j < info.y; j++)
		{
			const int t = M.leafTri[info.x + j];
			const float4 tl = M.lo[t], th = M.hi[t];
			if (tl.x > mx.x || mn.x > th.x || tl.y > mx.y || mn.y > th.y || tl.z > mx.z || mn.z > th.z)
				continue;
			if (nc == MESH_MAX_TRIS)
			{
				atomicOr(overflow, 1u);
				return;
			}
			idx[nc++] = t;
		}
		nd++;
	}
	if (nc == 0)
		return;
	for (int a = 1; a < nc; a++)   // triangle order
	{
		const int x = idx[a];
		int b = a - 1;
		while (b >= 0 && idx[b] > x)
		{
			idx[b + 1] = idx[b];
			b--;
		}
		idx[b + 1] = x;
	}
	MPatches P;
	P.np = 0;
	const float4 g = geom[i];
	const int type = int(g.w);
	xf x;
	x.p = mk(pos[i]);
	x.q = mkq(rot[i]);
	if (type == GEOM_SPHERE || type == GEOM_CAPSULE)
	{
		MeshCache C;
		MeshDef def[MESH_MAX_TRIS];
		if (type == GEOM_SPHERE)
			meshSphere(M, idx, nc, x.p, g.x, cd, C, def, P);
		else
		{
			const f3 tmp = qrot(x.q, mk(1.0f, 0.0f, 0.0f)) * g.y;
			meshCapsule(M, idx, nc, x.p + tmp, x.p - tmp, x.p, g.x, cd, C, def, P);
		}
	}
	else
		meshPoly(M, idx, nc, type == GEOM_CONVEXMESH ? hullPoly(hs, int(g.x)) : boxPoly(mk(g.x, g.y, g.z)), x, cd, P);
	int order[MESH_NPATCH];
	for (int a = 0; a < P.np; a++)
	{
		int b = a - 1;
		while (b >= 0 && P.p[order[b]].maxPen > P.p[a].maxPen)
		{
			order[b + 1] = order[b];
			b--;
		}
		order[b + 1] = a;
	}
	const int keep = P.np < MESH_MAXP ? P.np : MESH_MAXP;
	for (int a = 0; a < keep; a++)
	{
		MPatch& q = P.p[order[a]];
		int cnt = meshReduce4(q.c, q.count, q.n);
		if (type == GEOM_SPHERE)
		{
			int d = 0;
			for (int j = 1; j < cnt; j++)
				if (q.c[d].s > q.c[j].s) d = j;
			q.c[0] = q.c[d];
			cnt = 1;
		}
		ctN[base + a] = f4(q.n, __int_as_float(cnt));
		for (int j = 0; j < cnt; j++)
			ctP[4 * int64_t(base + a) + j] = f4(q.c[j].p, q.c[j].s);
	}
}

// whether the bodies of a pair are joined: its (lower, higher) key is among
// the sorted joint keys
PXH_D bool pairJointed(uint2 ab, const uint64_t* keys, int nKeys)
{
	const uint32_t lo = ab.x < ab.y ? ab.x : ab.y, hi = ab.x < ab.y ? ab.y : ab.x;
	const uint64_t k = (uint64_t(lo) << 32) | hi;
	int a = 0, b = nKeys;
	while (a < b)
	{
		const int m = (a + b) >> 1;
		if (keys[m] < k)
			a = m + 1;
		else
			b = m;
	}
	return a < nKeys && keys[a] == k;
}

__global__ void k_joint_filter(int cap, int Q, const uint32_t* dP, const uint2* pairs, const uint64_t* keys, int nKeys,
                               float4* ctN)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= cap || uint32_t(i) >= *dP)
		return;
	if (pairJointed(pairs[i], keys, nKeys))
		ctN[Q + i].w = __int_as_float(0);
}

// ---------------------------------------------------------------------------
// sleeping. The observable order of the reference CPU pipeline
// (simulationcontroller/src/ScPipeline.cpp:1472-1776, lowlevel/software/src/
// PxsIslandSim.cpp:1630-1720): the narrowphase runs on every pair; a touching
// pair of an awake and a sleeping body wakes the sleeping body's island; an
// island whose bodies were all ready for sleep after the last step sleeps
// before the solve; the solve and integration skip sleeping bodies.
// ---------------------------------------------------------------------------
PXH_D bool asleepBody(const uint32_t* flags, uint32_t i) { return (flags[i] & SLEEP_ASLEEP) != 0u; }

__global__ void k_sleep_count_init(int n, SleepState ss)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		ss.counted[i] = ss.jcount[i];   // each joint counts (BodySim::onConstraintAttach, ScBodySim.h:159)
}

// Interactions a body counts: broadphase pairs that are not filtered out
// (ScShapeInteraction.cpp:84-93, ScPipeline.cpp:1222-1228) and planes whose
// bounds its inflated bounds overlap: the half-space behind the plane when its
// normal is a coordinate axis, else the whole space (computePlaneBounds,
// geomutils/src/GuBounds.cpp:210-250), inflated by the plane's contact offset.
__global__ void k_sleep_count(int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody,
                              const int* qPlane, const float4* planes, const float4* bmin, const float4* bmax, float co,
                              const uint64_t* keys, int nKeys, int nMB, const int* meshBody, const int* world,
                              const float4* meshLo, const float4* meshHi, SleepState ss)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= Q + cap)
	{
		// a mesh whose broadphase bounds (its world bounds inflated by its
		// contact offset) overlap the body's
		const int k = s - Q - cap;
		if (k >= nMB)
			return;
		const int i = meshBody[k];
		const float4 lo = meshLo[world[i]], hi = meshHi[world[i]], a = bmin[i], b = bmax[i];
		if (!(lo.x - co > b.x || a.x > hi.x + co || lo.y - co > b.y || a.y > hi.y + co || lo.z - co > b.z ||
		      a.z > hi.z + co))
			atomicAdd(&ss.counted[i], 1);
		return;
	}
	if (s < Q)
	{
		const int i = qBody[s];
		const float4 pl = planes[qPlane[s]];
		const float ax = fabsf(pl.x), ay = fabsf(pl.y), az = fabsf(pl.z);
		const float e = 1e-6f, one = 1.0f - e;
		int axis = -1;
		if (ax > one && ay < e && az < e) axis = 0;
		else if (ax < e && ay > one && az < e) axis = 1;
		else if (ax < e && ay < e && az > one) axis = 2;
		bool counted = true;
		if (axis >= 0)
		{
			const float4 lo4 = bmin[i], hi4 = bmax[i];
			const float lo = axis == 0 ? lo4.x : axis == 1 ? lo4.y : lo4.z;
			const float hi = axis == 0 ? hi4.x : axis == 1 ? hi4.y : hi4.z;
			const float c = axis == 0 ? pl.x : axis == 1 ? pl.y : pl.z;
			counted = c > 0.0f ? lo <= -pl.w + co : hi >= pl.w - co;
		}
		if (counted)
			atomicAdd(&ss.counted[i], 1);
		return;
	}
	const int k = s - Q;
	if (k >= cap || uint32_t(k) >= *dP)
		return;
	const uint2 ab = pairs[k];
	if (nKeys > 0 && pairJointed(ab, keys, nKeys))
		return;   // jointed bodies: the pair is filtered out, no interaction
	atomicAdd(&ss.counted[ab.x], 1);
	atomicAdd(&ss.counted[ab.y], 1);
}

// a touching pair of an awake and a sleeping body names the sleeping body's
// island (setEdgesConnected, secondPassIslandGen: ScPipeline.cpp:1726-1776)
__global__ void k_sleep_wake_mark(int Q, int cap, const uint32_t* dP, const uint2* pairs, const float4* ctN,
                                  SleepState ss)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= cap || uint32_t(k) >= *dP || __float_as_int(ctN[Q + k].w) <= 0)
		return;
	const uint2 ab = pairs[k];
	const bool sa = asleepBody(ss.flags, ab.x), sb = asleepBody(ss.flags, ab.y);
	if (sa != sb)
		ss.wakeIsl[ss.isl[sa ? ab.x : ab.y]] = 1u;
}

// bodies of named islands wake (wakeObjectsUp), keeping their wake counter,
// not ready for sleep (IslandSim::activateNode, PxsIslandSim.cpp:400-432)
__global__ void k_sleep_wake_apply(int n, SleepState ss)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n && asleepBody(ss.flags, i) && ss.wakeIsl[ss.isl[i]])
		ss.flags[i] = 0u;
}

// pairs of sleeping bodies lose their contacts (inactive contact managers)
__global__ void k_sleep_drop(int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody, int nMB,
                             const int* meshBody, float4* ctN, const uint32_t* flags)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s >= Q + cap)
	{
		const int k = s - Q - cap;
		if (k < nMB * MESH_MAXP && asleepBody(flags, meshBody[k / MESH_MAXP]))
			ctN[s].w = __int_as_float(0);
		return;
	}
	if (s < Q)
	{
		if (asleepBody(flags, qBody[s]))
			ctN[s].w = __int_as_float(0);
		return;
	}
	const int k = s - Q;
	if (k >= cap || uint32_t(k) >= *dP)
		return;
	const uint2 ab = pairs[k];
	if (asleepBody(flags, ab.x) && asleepBody(flags, ab.y))
		ctN[s].w = __int_as_float(0);
}

__global__ void k_sleep_isl_init(int n, SleepState ss)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
	{
		ss.parent[i] = i;
		ss.notReady[i] = 0u;
		ss.wakeIsl[i] = 0u;
	}
}

// Island union-find over the body indices, roots the smallest index of their
// tree. Parents only ever move to smaller indices, so a stale read is still an
// ancestor; roots are linked only by compare-and-swap, and path halving writes
// only to non-roots.
PXH_D int islFind(int* par, int x)
{
	int cur = par[x];
	if (cur == x)
		return x;
	int prev = x, next;
	while (cur > (next = par[cur]))
	{
		par[prev] = next;
		prev = cur;
		cur = next;
	}
	return cur;
}

PXH_D void islUnion(int* par, int a, int b)
{
	int ra = islFind(par, a), rb = islFind(par, b);
	while (ra != rb)
	{
		if (ra < rb)
		{
			const int t = ra;
			ra = rb;
			rb = t;
		}
		// link the larger root under the smaller one
		const int old = atomicCAS(&par[ra], ra, rb);
		if (old == ra)
			return;
		ra = islFind(par, old);
		rb = islFind(par, rb);
	}
}

// touching body pairs (both awake after the wake-up) and joints of two awake
// bodies join islands
__global__ void k_sleep_union(int Q, int cap, const uint32_t* dP, const uint2* pairs, const float4* ctN, int J,
                              const int4* jInfo, SleepState ss)
{
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s < cap)
	{
		if (uint32_t(s) < *dP && __float_as_int(ctN[Q + s].w) > 0)
		{
			const uint2 ab = pairs[s];
			islUnion(ss.parent, int(ab.x), int(ab.y));
		}
		return;
	}
	const int j = s - cap;
	if (j >= J)
		return;
	const int4 ji = jInfo[j];
	if (ji.y >= 0 && ji.z >= 0 && !asleepBody(ss.flags, ji.y) && !asleepBody(ss.flags, ji.z))
		islUnion(ss.parent, ji.y, ji.z);
}

// awake bodies take their island's label; a body not ready for sleep keeps
// its island awake
__global__ void k_sleep_label(int n, SleepState ss)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || asleepBody(ss.flags, i))
		return;
	const int r = islFind(ss.parent, i);
	ss.isl[i] = r;
	if (!(ss.flags[i] & SLEEP_READY))
		ss.notReady[r] = 1u;
}

// islands whose bodies are all ready go to sleep, velocities zeroed
// (IslandSim::processLostEdges deactivation, PxsIslandSim.cpp:1690-1720;
// BodySim::deactivate, ScBodySim.cpp:469-490)
__global__ void k_sleep_deactivate(int n, SleepState ss, float4* lin, float4* ang)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || asleepBody(ss.flags, i) || ss.notReady[ss.isl[i]])
		return;
	ss.flags[i] |= SLEEP_ASLEEP;
	lin[i] = make_float4(0.0f, 0.0f, 0.0f, lin[i].w);
	ang[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

PXH_D void wakeBody(const SleepState& ss, uint32_t i);

__global__ void k_sleep_wake_bodies(int k, const int* bodies, SleepState ss)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j < k)
		wakeBody(ss, uint32_t(bodies[j]));
}

__global__ void k_sleep_read(int n, SleepState ss, float* wc, int* asleep)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
	{
		wc[i] = ss.wc[i];
		asleep[i] = asleepBody(ss.flags, i) ? 1 : 0;
	}
}

PXH_D xf bodyPose(int i, const float4* pos, const float4* rot)
{
	xf t;
	if (i < 0)
	{
		t.q.x = t.q.y = t.q.z = 0.0f;
		t.q.w = 1.0f;
		t.p = mk(0.0f, 0.0f, 0.0f);
		return t;
	}
	t.q = mkq(rot[i]);
	t.p = mk(pos[i]);
	return t;
}

PXH_D m33 bodySqrtInvI(int i, const float4* svI0, const float4* svI1)
{
	if (i < 0)
	{
		m33 z;
		z.c0 = z.c1 = z.c2 = mk(0.0f, 0.0f, 0.0f);
		return z;
	}
	return loadSym(svI0[i], svI1[i]);
}

PXH_D xf jointFrame(const float4* f)
{
	xf t;
	t.p = mk(f[0]);
	t.q = mkq(f[1]);
	return t;
}

// a joint whose bodies all sleep is out of the solve
PXH_D bool jointAsleep(int4 ji, const uint32_t* sleepFlags)
{
	return sleepFlags && (ji.y < 0 || (sleepFlags[ji.y] & SLEEP_ASLEEP)) &&
	       (ji.z < 0 || (sleepFlags[ji.z] & SLEEP_ASLEEP));
}

__global__ void k_joint_prep(int J, const int4* info, const float4* frame, const float4* param, const float4* pos,
                             const float4* rot, const float4* vel, const float4* rec, const float4* svI0,
                             const float4* svI1, float stepDt, float invStepDt, float invDt, float erp,
                             float lengthScale, float4* head, float4* rows, const uint32_t* sleepFlags)
{
	const int j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= J)
		return;
	const int4 ji = info[j];
	if (jointAsleep(ji, sleepFlags))
		return;
	const xf bA = bodyPose(ji.y, pos, rot), bB = bodyPose(ji.z, pos, rot);
	const m33 sI0 = bodySqrtInvI(ji.y, svI0, svI1), sI1 = bodySqrtInvI(ji.z, svI0, svI1);
	const float im0 = ji.y >= 0 ? vel[4 * int64_t(ji.y)].w : 0.0f;
	const float im1 = ji.z >= 0 ? vel[4 * int64_t(ji.z)].w : 0.0f;
	// the bodies' velocities at the start of the solve (world frame)
	const f3 zero = mk(0.0f, 0.0f, 0.0f);
	const f3 v0 = ji.y >= 0 ? mk(rec[PREP_REC * int64_t(ji.y) + 4]) : zero;
	const f3 w0 = ji.y >= 0 ? mk(rec[PREP_REC * int64_t(ji.y) + 5]) : zero;
	const f3 v1 = ji.z >= 0 ? mk(rec[PREP_REC * int64_t(ji.z) + 4]) : zero;
	const f3 w1 = ji.z >= 0 ? mk(rec[PREP_REC * int64_t(ji.z) + 5]) : zero;
	const float4* jp = param + JOINT_PARAM * int64_t(j);
	prepJoint(ji.x, jp[0], jp[1], jp[2], bA, bB, jointFrame(frame + 4 * int64_t(j)),
	          jointFrame(frame + 4 * int64_t(j) + 2), sI0, sI1, im0, im1, v0, w0, v1, w1, stepDt, invStepDt, invDt,
	          erp, lengthScale, head + JOINT_HEAD * int64_t(j), rows + JOINT_ROWS * int64_t(j));
}

__global__ void k_joint_solve(int begin, int end, const int4* info, const float4* head, float4* rows, float4* vel,
                              const float4* svDQ, const float4* svI0, const float4* svI1, float elapsed, int conclude,
                              const uint32_t* sleepFlags)
{
	const int j = begin + blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= end)
		return;
	const int4 ji = info[j];
	if (jointAsleep(ji, sleepFlags))
		return;
	const int a = ji.y, b = ji.z;
	f3 l0 = mk(0.0f, 0.0f, 0.0f), a0 = l0, dl0 = l0, da0 = l0, l1 = l0, a1 = l0, dl1 = l0, da1 = l0;
	qt q0, q1;
	q0.x = q0.y = q0.z = 0.0f; q0.w = 1.0f;
	q1 = q0;
	float w0 = 0.0f, w1 = 0.0f;
	if (a >= 0)
	{
		const float4* v = vel + 4 * int64_t(a);
		const float4 v0 = v[0];
		l0 = mk(v0); w0 = v0.w; a0 = mk(v[1]); dl0 = mk(v[2]); da0 = mk(v[3]);
		q0 = mkq(svDQ[a]);
	}
	if (b >= 0)
	{
		const float4* v = vel + 4 * int64_t(b);
		const float4 v0 = v[0];
		l1 = mk(v0); w1 = v0.w; a1 = mk(v[1]); dl1 = mk(v[2]); da1 = mk(v[3]);
		q1 = mkq(svDQ[b]);
	}
	const m33 sI0 = bodySqrtInvI(a, svI0, svI1), sI1 = bodySqrtInvI(b, svI0, svI1);
	const float4* h = head + JOINT_HEAD * int64_t(j);
	solveJoint(__float_as_int(h[8].x), h, rows + JOINT_ROWS * int64_t(j), l0, a0, dl0, da0,
	           q0, sI0, l1, a1, dl1, da1, q1, sI1, elapsed, conclude != 0);
	if (a >= 0)
	{
		vel[4 * int64_t(a)] = f4(l0, w0);
		vel[4 * int64_t(a) + 1] = f4(a0, 0.0f);
	}
	if (b >= 0)
	{
		vel[4 * int64_t(b)] = f4(l1, w1);
		vel[4 * int64_t(b) + 1] = f4(a1, 0.0f);
	}
}

// copyBackBodies (lowleveldynamics/src/DyTGSDynamics.cpp:1549-1580), with
// the sleep check on the step's mean velocities (DyTGSDynamics.cpp:1561, 1577)
__global__ void k_finalize(int n, float dt, float4* pos, float4* rot, float4* lin, float4* ang, const float4* invI,
                           const float4* vel, const float4* svP, const float4* svDQ, const float4* svI0,
                           const float4* svI1, SleepState ss)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	if (ss.flags && (ss.flags[i] & SLEEP_ASLEEP))
		return;   // sleeping: pose and zero velocities kept
	const qt q = qnormalize(qmul(mkq(svDQ[i]), mkq(rot[i])));
	rot[i] = f4(q);
	pos[i] = svP[i];
	const float4* v = vel + 4 * int64_t(i);
	const float4 l = v[0];
	lin[i] = l;
	const m33 S = loadSym(svI0[i], svI1[i]);
	ang[i] = f4(mmul(S, mk(v[1])), 0.0f);
	if (ss.flags)
	{
		const float invDt = 1.0f / dt;
		sleepCheck(ss, i, dt, mk(v[2]) * invDt, mmul(S, mk(v[3]) * invDt), q, l.w, invI[i]);
	}
}

// ---------------------------------------------------------------------------
// Direct GPU API gather/scatter (updateBodiesAndShapes.cu:999-1253)
// ---------------------------------------------------------------------------
// what: PxRigidDynamicGPUAPIReadType (0 pose, 1 linear velocity, 2 angular
// velocity, 3 linear acceleration, 4 angular acceleration: the velocity change
// of the last step over its dt, getRigidDynamicLinearAcceleration,
// gpusimulationcontroller/src/CUDA/updateBodiesAndShapes.cu:1063-1106)
__global__ void k_get_data(int n, const uint32_t* idx, int what, const float4* pos, const float4* rot,
                           const float4* lin, const float4* ang, const float4* prevLin, const float4* prevAng,
                           float invDt, float* out)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const uint32_t i = idx ? idx[k] : uint32_t(k);
	if (what == 0)
	{
		const float4 q = rot[i], p = pos[i];
		float* o = out + 7 * int64_t(k);
		o[0] = q.x; o[1] = q.y; o[2] = q.z; o[3] = q.w; o[4] = p.x; o[5] = p.y; o[6] = p.z;
	}
	else if (what <= 2)
	{
		const float4 v = what == 1 ? lin[i] : ang[i];
		float* o = out + 3 * int64_t(k);
		o[0] = v.x; o[1] = v.y; o[2] = v.z;
	}
	else
	{
		const f3 a = (what == 3 ? mk(lin[i]) - mk(prevLin[i]) : mk(ang[i]) - mk(prevAng[i])) * invDt;
		float* o = out + 3 * int64_t(k);
		o[0] = a.x; o[1] = a.y; o[2] = a.z;
	}
}

// what: PxRigidDynamicGPUAPIWriteType (0 pose, 1 linear velocity, 2 angular
// velocity, 3 force, 4 torque). A force becomes the linear acceleration
// F / m, a torque the angular acceleration I_world^-1 tau with the body's
// current orientation (PxForceMode::eFORCE, NpRigidBodyTemplate.h:507-528;
// setRigidDynamicForce / setRigidDynamicTorque, updateBodiesAndShapes.cu:1229-1278)
// PxRigidDynamic::wakeUp (BodyCore::wakeUp with the scene's wake counter
// reset value, 20 * 0.02 s): awake now, not ready for sleep; the rest of a
// sleeping body's island wakes at the next step (IslandSim::activateNode)
PXH_D void wakeBody(const SleepState& ss, uint32_t i)
{
	ss.wc[i] = 20.0f * 0.02f;
	const uint32_t f = ss.flags[i];
	ss.flags[i] = 0u;
	if (f & SLEEP_ASLEEP)
		ss.wakeIsl[ss.isl[i]] = 1u;
}

__global__ void k_set_data(int n, const uint32_t* idx, int what, float4* pos, float4* rot, float4* lin, float4* ang,
                           const float4* invI, float4* extLin, float4* extAng, const float* in, SleepState ss)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const uint32_t i = idx ? idx[k] : uint32_t(k);
	if (what == 0)
	{
		const float* s = in + 7 * int64_t(k);
		rot[i] = make_float4(s[0], s[1], s[2], s[3]);
		pos[i] = make_float4(s[4], s[5], s[6], 0.0f);
		if (ss.flags)
			wakeBody(ss, i);   // autowake of setGlobalPose
		return;
	}
	const float* s = in + 3 * int64_t(k);
	const f3 v = mk(s[0], s[1], s[2]);
	// velocities always wake the body, forces and torques when nonzero
	// (autowake of setLinearVelocity / addForce)
	if (ss.flags && (what <= 2 || v.x != 0.0f || v.y != 0.0f || v.z != 0.0f))
		wakeBody(ss, i);
	if (what == 1)
		lin[i] = make_float4(v.x, v.y, v.z, lin[i].w);
	else if (what == 2)
		ang[i] = make_float4(v.x, v.y, v.z, 0.0f);
	else if (what == 3)
		extLin[i] = f4(v * lin[i].w, 0.0f);
	else
	{
		// R diag(invI) R^T (Cm::transformInertiaTensor)
		const m33 R = qmat(mkq(rot[i]));
		const float4 ii = invI[i];
		m33 A;
		A.c0 = R.c0 * ii.x; A.c1 = R.c1 * ii.y; A.c2 = R.c2 * ii.z;
		extAng[i] = f4(mmul(mmm(A, mtrans(R)), v), 0.0f);
	}
}

// ---------------------------------------------------------------------------
// standalone broadphase: found/lost pairs between consecutive pair sets, both
// sorted by (id0, id1) (BpBroadPhase.h:159-185 created/deleted semantics)
// ---------------------------------------------------------------------------
// canonical pair key (lower index first)
PXH_D uint64_t pkey(uint2 p) { return (uint64_t(min(p.x, p.y)) << 32) | max(p.x, p.y); }

// pair keys in the caller's handle space (ids[box], or the box index)
__global__ void k_pair_keys(int P, const uint2* pairs, const uint32_t* ids, uint64_t* keys, uint32_t* vals)
{
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p < P)
	{
		uint2 q = pairs[p];
		if (ids)
			q = make_uint2(ids[q.x], ids[q.y]);
		keys[p] = pkey(q);
		vals[p] = uint32_t(p);
	}
}

// flag[i] = 1 if key a[i] is absent from sorted b
__global__ void k_set_diff_flags(int na, const uint64_t* a, int nb, const uint64_t* b, uint32_t* flag)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= na)
		return;
	const uint64_t k = a[i];
	int l = 0, h = nb;
	while (l < h)
	{
		const int m = (l + h) >> 1;
		if (b[m] < k) l = m + 1;
		else h = m;
	}
	flag[i] = (l < nb && b[l] == k) ? 0u : 1u;
}

// touch events (prepareLostFoundPairs_Stage1, gpunarrowphase/src/CUDA/cudaGJKEPA.cu:1468):
// a constraint without a row in the previous step starts touching; the
// previous rows it finds are marked, the unmarked ones stopped touching
__global__ void k_found_flags(int C, const int* prevRow, uint32_t* found, uint32_t* matched)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= C)
		return;
	const int po = prevRow ? prevRow[c] : -1;
	found[c] = po < 0 ? 1u : 0u;
	if (po >= 0)
		matched[po] = 1u;
}
__global__ void k_lost_flags(int R, const uint32_t* matched, uint32_t* lost)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r < R)
		lost[r] = matched[r] ? 0u : 1u;
}

__global__ void k_compact_keys(int na, const uint64_t* a, const uint32_t* flag, const uint32_t* scanned, uint2* out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < na && flag[i])
		out[scanned[i]] = make_uint2(uint32_t(a[i] >> 32), uint32_t(a[i]));
}


// batched immediate-mode contacts (PxGenerateContacts on a fresh cache)
__global__ void k_np_immediate(int n, const int* type0, const float* geom0, const float* pose0, const int* type1,
                               const float* geom1, const float* pose1, float cd, float tol, float* out, int* counts)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= n)
		return;
	const int t0 = type0[k], t1 = type1[k];
	const float* g0 = geom0 + 4 * int64_t(k);
	const float* g1 = geom1 + 4 * int64_t(k);
	const float* q0 = pose0 + 7 * int64_t(k);
	const float* q1 = pose1 + 7 * int64_t(k);
	xf x0, x1;
	x0.p = mk(q0[0], q0[1], q0[2]); x0.q.x = q0[3]; x0.q.y = q0[4]; x0.q.z = q0[5]; x0.q.w = q0[6];
	x1.p = mk(q1[0], q1[1], q1[2]); x1.q.x = q1[3]; x1.q.y = q1[4]; x1.q.z = q1[5]; x1.q.w = q1[6];
	Contacts c;
	c.count = 0;
	int cnt = -1;
	if (t0 == GEOM_SPHERE && t1 == GEOM_SPHERE)
		cnt = sphereSphere(g0[0], x0.p, g1[0], x1.p, cd, c);
	else if (t0 == GEOM_SPHERE && t1 == GEOM_PLANE)
	{
		const f3 pn = qrot(x1.q, mk(1.0f, 0.0f, 0.0f));
		cnt = spherePlane(g0[0], x0.p, pn, -dot(pn, x1.p), cd, c);
	}
	else if (t0 == GEOM_SPHERE && t1 == GEOM_BOX)
		cnt = sphereBox(g0[0], x0.p, mk(g1[0], g1[1], g1[2]), x1, cd, c);
	else if (t0 == GEOM_PLANE && t1 == GEOM_BOX)
	{
		const f3 pn = qrot(x0.q, mk(1.0f, 0.0f, 0.0f));
		cnt = planeBox(pn, -dot(pn, x0.p), mk(g1[0], g1[1], g1[2]), x1, cd, c);
	}
	else if (t0 == GEOM_BOX && t1 == GEOM_BOX)
		cnt = boxBox(mk(g0[0], g0[1], g0[2]), x0, mk(g1[0], g1[1], g1[2]), x1, cd, tol, c);
	else if (t0 == GEOM_SPHERE && t1 == GEOM_CAPSULE)
		cnt = sphereCapsule(g0[0], x0.p, g1[0], g1[1], x1, cd, c);
	else if (t0 == GEOM_CAPSULE && t1 == GEOM_CAPSULE)
		cnt = capsuleCapsule(g0[0], g0[1], x0, g1[0], g1[1], x1, cd, c);
	else if (t0 == GEOM_PLANE && t1 == GEOM_CAPSULE)
	{
		const f3 pn = qrot(x0.q, mk(1.0f, 0.0f, 0.0f));
		cnt = planeCapsule(pn, -dot(pn, x0.p), g1[0], g1[1], x1, cd, c);
	}
	else if (t0 == GEOM_CAPSULE && t1 == GEOM_BOX)
		cnt = capsulePoly(g0[0], g0[1], x0, boxPoly(mk(g1[0], g1[1], g1[2])), x1, cd, tol, c);
	counts[k] = cnt;
	for (int i = 0; i < (cnt > 0 ? cnt : 0); i++)
	{
		float* o = out + (4 * int64_t(k) + i) * 8;
		o[0] = c.p[i].x; o[1] = c.p[i].y; o[2] = c.p[i].z; o[3] = c.sep[i];
		o[4] = c.n.x; o[5] = c.n.y; o[6] = c.n.z; o[7] = 0.0f;
	}
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int nblk(int64_t n, int t) { return int((n + t - 1) / t); }

void launch_scan(uint32_t* data, int64_t n, uint32_t* tmp, hipStream_t st)
{
	// tmp must hold ceil(n/2048) + ceil(n/2048^2) + ... entries
	if (n <= 0)
		return;
	const int64_t nb = (n + SCAN_TILE - 1) / SCAN_TILE;
	k_scan_tiles<<<int(nb), SCAN_THREADS, 0, st>>>(data, n, nb > 1 ? tmp : nullptr);
	if (nb > 1)
	{
		launch_scan(tmp, nb, tmp + nb, st);
		k_scan_add<<<nblk(n, 256), 256, 0, st>>>(data, n, tmp);
	}
}

int64_t scan_tmp_size(int64_t n)
{
	int64_t total = 0;
	while (n > 1)
	{
		n = (n + SCAN_TILE - 1) / SCAN_TILE;
		total += n;
	}
	return total + 1;
}

void launch_radix_sort(uint64_t* keys, uint32_t* vals, uint64_t* keys2, uint32_t* vals2, int n, int bitBegin,
                       int bitEnd, uint32_t* hist, uint32_t* scanTmp, hipStream_t st, bool& resultInAlt)
{
	resultInAlt = false;
	if (n <= 0)
		return;
	const int nb = nblk(n, RS_TILE);
	uint64_t* ka = keys;
	uint32_t* va = vals;
	uint64_t* kb = keys2;
	uint32_t* vb = vals2;
	for (int shift = bitBegin; shift < bitEnd; shift += 8)
	{
		k_rs_hist<<<nb, RS_THREADS, 0, st>>>(ka, n, shift, hist, nb);
		launch_scan(hist, int64_t(256) * nb, scanTmp, st);
		k_rs_scatter<<<nb, RS_THREADS, 0, st>>>(ka, va, kb, vb, n, shift, hist, nb);
		uint64_t* tk = ka; ka = kb; kb = tk;
		uint32_t* tv = va; va = vb; vb = tv;
		resultInAlt = !resultInAlt;
	}
}

#define PXH_LAUNCH(kernel, n, ...) \
	do { if ((n) > 0) kernel<<<nblk((n), 256), 256, 0, st>>>(__VA_ARGS__); } while (0)

void launch_bounds(int n, const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, float co,
                   float4* bmin, float4* bmax, hipStream_t st)
{
	PXH_LAUNCH(k_bounds, n, n, pos, rot, geom, hs, co, bmin, bmax);
}
void launch_scene_grid(int n, int nScenes, const int* scene, const float4* bmin, const float4* bmax, uint32_t* acc,
                       float4* gridF, int2* gridI, const uint32_t* rowOff, hipStream_t st)
{
	PXH_LAUNCH(k_grid_init, nScenes, nScenes, acc);
	if (n > 0)
		k_grid_reduce<<<nblk(n, 256 * GRID_ITEMS), 256, 0, st>>>(n, bmin, bmax, scene, acc);
	PXH_LAUNCH(k_grid_finish, nScenes, nScenes, acc, gridF, gridI, rowOff);
}
void launch_row_off_single(uint32_t* rowOff, uint32_t cap, hipStream_t st)
{
	k_row_off_single<<<1, 64, 0, st>>>(rowOff, cap);
}
void launch_keys(int n, const float4* bmin, const float4* bmax, const int* scene, const float4* gridF,
                 const int2* gridI, const uint32_t* rowOff, uint64_t* keys, uint32_t* vals, hipStream_t st)
{
	PXH_LAUNCH(k_keys, n, n, bmin, bmax, scene, gridF, gridI, rowOff, keys, vals);
}
void launch_gather_sorted(int n, const uint32_t* sv, const float4* bmin, const float4* bmax, float4* smin,
                          float4* smax, hipStream_t st)
{
	PXH_LAUNCH(k_gather_sorted, n, n, sv, bmin, bmax, smin, smax);
}
void launch_row_start(int n, const uint64_t* sk, uint32_t* rowStart, uint32_t totalRows, hipStream_t st)
{
	if (n == 0)
	{
		PXH_LAUNCH(k_fill_u32, int64_t(totalRows) + 1, rowStart, int64_t(totalRows) + 1, 0u);
		return;
	}
	PXH_LAUNCH(k_row_start, n, n, sk, rowStart, totalRows);
	PXH_LAUNCH(k_row_tail, int64_t(totalRows) + 1, n, sk, rowStart, totalRows);
}
void launch_sweep(bool write, int n, const uint64_t* sk, const float4* smin, const float4* smax, const int* scene,
                  const int* group, uint64_t groupLut, const int2* gridI, const uint32_t* rowOff,
                  const uint32_t* rowStart, uint32_t* counts, const uint32_t* offsets, uint2* pairs, uint32_t pairCap,
                  hipStream_t st)
{
	if (write)
		PXH_LAUNCH(k_sweep<true>, n, n, sk, smin, smax, scene, group, groupLut, gridI, rowOff, rowStart, counts,
		           offsets, pairs, pairCap);
	else
		PXH_LAUNCH(k_sweep<false>, n, n, sk, smin, smax, scene, group, groupLut, gridI, rowOff, rowStart, counts,
		           offsets, pairs, pairCap);
}
void launch_np_pairs(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                     const float4* pos, const float4* rot, const float4* geom, float cd, float tol, float4* ctN,
                     float4* ctP, hipStream_t st)
{
	PXH_LAUNCH(k_np_pairs, cap, cap, dP, dConvex, slotBase, pairs, pos, rot, geom, cd, tol, ctN, ctP);
}
void launch_np_convex(int cap, const uint32_t* dP, const uint32_t* dConvex, int slotBase, const uint2* pairs,
                      const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, float cd, float tol,
                      float4* wit, uint32_t* list, float4* ctN, float4* ctP, hipStream_t st)
{
	PXH_LAUNCH(k_np_convex, cap, cap, dP, dConvex, slotBase, pairs, pos, rot, geom, hs, cd, tol, ctN, ctP, list);
	if (cap > 0)
	{
		// grid-stride: enough waves to fill the chip several times over
		const int blocks = std::min((cap + CVX_WAVES - 1) / CVX_WAVES, 8192);
		k_np_convex_wave<<<blocks, 64 * CVX_WAVES, 0, st>>>(cap, dP, dConvex, slotBase, pairs, pos, rot, geom, hs, cd,
		                                                    wit, list, ctN);
	}
	PXH_LAUNCH(k_np_convex_finish, cap, cap, dP, dConvex, slotBase, pairs, pos, rot, geom, hs, cd, tol, wit, list, ctN,
	           ctP);
}
void launch_np_planes_convex(int Q, int slotBase, const int* qBody, const int* qPlane, const float4* pos,
                             const float4* rot, const float4* geom, const float4* planes, const HullSet& hs, float cd,
                             float tol, float4* ctN, float4* ctP, hipStream_t st)
{
	PXH_LAUNCH(k_np_planes_convex, Q, Q, slotBase, qBody, qPlane, pos, rot, geom, planes, hs, cd, tol, ctN, ctP);
}
void launch_np_planes(int Q, int slotBase, const int* qBody, const int* qPlane, const float4* pos, const float4* rot,
                      const float4* geom, const float4* planes, float cd, float4* ctN, float4* ctP, hipStream_t st)
{
	PXH_LAUNCH(k_np_planes, Q, Q, slotBase, qBody, qPlane, pos, rot, geom, planes, cd, ctN, ctP);
}
void launch_copy_contacts(int C, int maxPairs, const int* order, const int* conA, const int* conB, const int* conSlot,
                          const float4* ctN, const float4* ctP, const float* nF, float4* out, uint32_t* nbPairs,
                          hipStream_t st)
{
	const int n = std::max(C, 1);
	k_copy_contacts<<<nblk(n, 256), 256, 0, st>>>(C, maxPairs, order, conA, conB, conSlot, ctN, ctP, nF, out, nbPairs);
}
void launch_joint_filter(int cap, int Q, const uint32_t* dP, const uint2* pairs, const uint64_t* jointKeys, int nKeys,
                         float4* ctN, hipStream_t st)
{
	if (nKeys > 0)
		PXH_LAUNCH(k_joint_filter, cap, cap, Q, dP, pairs, jointKeys, nKeys, ctN);
}
void launch_sleep_count(int n, int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody,
                        const int* qPlane, const float4* planes, const float4* bmin, const float4* bmax, float co,
                        const uint64_t* jointKeys, int nKeys, int nMB, const int* meshBody, const int* world,
                        const float4* meshLo, const float4* meshHi, SleepState ss, hipStream_t st)
{
	PXH_LAUNCH(k_sleep_count_init, n, n, ss);
	PXH_LAUNCH(k_sleep_count, Q + cap + nMB, Q, cap, dP, pairs, qBody, qPlane, planes, bmin, bmax, co, jointKeys,
	           nKeys, nMB, meshBody, world, meshLo, meshHi, ss);
}
void launch_sleep_wake(int n, int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody, int nMB,
                       const int* meshBody, float4* ctN, SleepState ss, hipStream_t st)
{
	PXH_LAUNCH(k_sleep_wake_mark, cap, Q, cap, dP, pairs, ctN, ss);
	PXH_LAUNCH(k_sleep_wake_apply, n, n, ss);
	PXH_LAUNCH(k_sleep_drop, Q + cap + nMB * MESH_MAXP, Q, cap, dP, pairs, qBody, nMB, meshBody, ctN, ss.flags);
}
void launch_sleep_islands(int n, int Q, int cap, const uint32_t* dP, const uint2* pairs, const int* qBody, int nMB,
                          const int* meshBody, float4* ctN, int J, const int4* jInfo, float4* lin, float4* ang,
                          SleepState ss, hipStream_t st)
{
	PXH_LAUNCH(k_sleep_isl_init, n, n, ss);
	PXH_LAUNCH(k_sleep_union, cap + J, Q, cap, dP, pairs, ctN, J, jInfo, ss);
	PXH_LAUNCH(k_sleep_label, n, n, ss);
	PXH_LAUNCH(k_sleep_deactivate, n, n, ss, lin, ang);
	PXH_LAUNCH(k_sleep_drop, Q + cap + nMB * MESH_MAXP, Q, cap, dP, pairs, qBody, nMB, meshBody, ctN, ss.flags);
}
void launch_np_mesh(int nMB, const int* meshBody, const int* world, const int4* worldMesh, const MeshSet& M,
                    const float4* pos, const float4* rot, const float4* geom, const HullSet& hs, float cd,
                    int slotBase, float4* ctN, float4* ctP, uint32_t* overflow, hipStream_t st)
{
	if (nMB > 0)
		k_np_mesh<<<nblk(nMB, 64), 64, 0, st>>>(nMB, meshBody, world, worldMesh, M, pos, rot, geom, hs, cd, slotBase,
		                                        ctN, ctP, overflow);
}
void launch_sleep_wake_bodies(int k, const int* bodies, SleepState ss, hipStream_t st)
{
	PXH_LAUNCH(k_sleep_wake_bodies, k, k, bodies, ss);
}
void launch_sleep_read(int n, SleepState ss, float* wc, int* asleep, hipStream_t st)
{
	PXH_LAUNCH(k_sleep_read, n, n, ss, wc, asleep);
}
void launch_joint_prep(int J, const int4* info, const float4* frame, const float4* param, const float4* pos,
                       const float4* rot, SolverBodies sb, float stepDt, float invStepDt, float invDt, float erp,
                       float lengthScale, float4* head, float4* rows, const uint32_t* sleepFlags, hipStream_t st)
{
	PXH_LAUNCH(k_joint_prep, J, J, info, frame, param, pos, rot, sb.vel, sb.rec, sb.i0, sb.i1, stepDt, invStepDt,
	           invDt, erp, lengthScale, head, rows, sleepFlags);
}
void launch_joint_solve(int begin, int end, const int4* info, const float4* head, float4* rows, SolverBodies sb,
                        float elapsed, bool conclude, const uint32_t* sleepFlags, hipStream_t st)
{
	if (end <= begin)
		return;
	k_joint_solve<<<nblk(end - begin, 64), 64, 0, st>>>(begin, end, info, head, rows, sb.vel, sb.dq, sb.i0, sb.i1,
	                                                     elapsed, conclude ? 1 : 0, sleepFlags);
}
void launch_touch_flags(int S, int Q, int pairEnd, const uint32_t* dP, const float4* ctN, uint32_t* flags,
                        hipStream_t st)
{
	PXH_LAUNCH(k_touch_flags, S, S, Q, pairEnd, dP, ctN, flags);
}
void launch_con_build(int S, int Q, int pairEnd, const uint32_t* dP, const float4* ctN, const uint32_t* scanned,
                      const uint2* pairs, const int* qBody, const int* qPlane, const int* meshBody, int nPlanes,
                      const int* local, const int* planeLocal, const int* world, const int* worldTag, int* conA,
                      int* conB, int* conSlot, uint64_t* conKey, uint64_t* conPKey, int* conColor, hipStream_t st)
{
	PXH_LAUNCH(k_con_build, S, S, Q, pairEnd, dP, ctN, scanned, pairs, qBody, qPlane, meshBody, nPlanes, local,
	           planeLocal, world, worldTag, conA, conB, conSlot, conKey, conPKey, conColor);
}
// out[i] = in[idx[i]] for elements of `bytes` (4 or 16) bytes
__global__ void k_gather4(int n, const int* idx, const uint32_t* in, uint32_t* out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = in[idx[i]];
}
__global__ void k_gather16(int n, const int* idx, const float4* in, float4* out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i < n)
		out[i] = in[idx[i]];
}
void launch_gather(int n, const int* idx, const void* in, void* out, int bytes, hipStream_t st)
{
	if (bytes == 16)
		PXH_LAUNCH(k_gather16, n, n, idx, static_cast<const float4*>(in), static_cast<float4*>(out));
	else
		PXH_LAUNCH(k_gather4, n, n, idx, static_cast<const uint32_t*>(in), static_cast<uint32_t*>(out));
}
// colouring launches are grid-stride over at most 1024 blocks
static int colorBlocks(int m) { return std::max(1, std::min(nblk(m, 256), 1024)); }
void launch_color_round(int m, const uint32_t* mDev, const int* list, int round, const int* conA, const int* conB,
                        const uint64_t* conKey, int* conColor, uint64_t* best, uint64_t* mask, hipStream_t st)
{
	if (m <= 0)
		return;
	unsigned long long* bst = reinterpret_cast<unsigned long long*>(best);
	const int g = colorBlocks(m);
	k_color_reset<<<g, 256, 0, st>>>(m, mDev, list, conA, conB, bst);
	k_color_min<<<g, 256, 0, st>>>(m, mDev, list, conA, conB, conKey, conColor, bst);
	k_color_pick<<<g, 256, 0, st>>>(m, mDev, list, round, conA, conB, conKey, conColor, bst,
	                                reinterpret_cast<unsigned long long*>(mask));
}
void launch_color_uncolored(int m, const uint32_t* mDev, const int* list, const int* conColor, uint32_t* flags,
                            hipStream_t st)
{
	k_color_uncolored<<<colorBlocks(m + 1), 256, 0, st>>>(m, mDev, list, conColor, flags);
}
void launch_color_compact(int m, const int* list, const uint32_t* flags, const uint32_t* scanned, int* out,
                          hipStream_t st)
{
	if (m > 0)
		k_color_compact<<<colorBlocks(m), 256, 0, st>>>(m, list, flags, scanned, out);
}
void launch_fill_u64(uint64_t* p, int n, uint64_t v, hipStream_t st) { PXH_LAUNCH(k_fill_u64, n, p, n, v); }
void launch_static_colors(int n, int C, const int* conA, const int* conB, int* conColor, uint32_t* bodyNext,
                          hipStream_t st)
{
	PXH_LAUNCH(k_fill_u32, n, bodyNext, int64_t(n), 0u);
	PXH_LAUNCH(k_body_next, C, C, conA, conB, conColor, bodyNext);
	PXH_LAUNCH(k_static_color, C, C, conA, conB, bodyNext, conColor);
}
void launch_patch_table(int m, const uint64_t* keys, uint64_t* htKey, int* htVal, uint32_t htMask, hipStream_t st)
{
	PXH_LAUNCH(k_fill_u64, int(htMask) + 1, htKey, int(htMask) + 1, ~0ull);
	PXH_LAUNCH(k_patch_insert, m, m, keys, reinterpret_cast<unsigned long long*>(htKey), htVal, htMask);
}
void launch_con_prev(int C, const int* conA, const int* conB, const int* conSlot, PatchIO io, int* conColor,
                     uint64_t* mask, hipStream_t st)
{
	PXH_LAUNCH(k_con_prev, C, C, conA, conB, conSlot, io, conColor, reinterpret_cast<unsigned long long*>(mask));
}
void launch_color_keys(int C, const int* conColor, uint32_t limit, uint64_t* keys, uint32_t* vals, hipStream_t st)
{
	PXH_LAUNCH(k_color_keys, C, C, conColor, limit, keys, vals);
}
void launch_max_color(int C, const int* conColor, uint32_t* maxColor, hipStream_t st)
{
	if (C > 0)
		k_max_color<<<std::min(nblk(C, 256), 256), 256, 0, st>>>(C, conColor, maxColor);
}
void launch_preintegrate(int n, const float4* pos, const float4* rot, const float4* lin, const float4* ang,
                         const float4* invI, float3 gdt, float linDampDt, float angDampDt, float maxLinSq,
                         float maxAngSq, bool pgs, SolverBodies sb, float dt, float4* extLin, float4* extAng,
                         float4* prevLin, float4* prevAng, const uint32_t* sleepFlags, hipStream_t st)
{
	PXH_LAUNCH(k_preintegrate, n, n, pos, rot, lin, ang, invI, gdt, linDampDt, angDampDt, maxLinSq, maxAngSq, pgs ? 1 : 0, sb.vel,
	           sb.p, sb.dq, sb.i0, sb.i1, sb.origLin, sb.rec, dt, extLin, extAng, prevLin, prevAng, sleepFlags);
}
void launch_prep(int C, const int* order, const int* conA, const int* conB, const int* conSlot, const float4* ctN,
                 const float4* ctP, const float4* pos, const float4* rot, SolverBodies sb, float invStepDt,
                 float scale, float restDist, float fot, Rows rows, PatchIO patches, unsigned long long* counters,
                 hipStream_t st)
{
	(void)pos;
	(void)rot;
	(void)conA;
	(void)conB;
	(void)conSlot;
	PXH_LAUNCH(k_prep, C, C, order, ctN, ctP, sb.rec, invStepDt, scale, restDist, fot, rows, patches, counters + 4);
	k_sum_counters<<<1, 256, 0, st>>>(nblk(C, 256), counters + 4, counters);
}
void launch_solve(int begin, int end, int C, Rows rows, SolverBodies sb, float minPen, float muS, float muD,
                  float invStepDt, float biasScale, float scale, uint32_t* broken, hipStream_t st, hipEvent_t evStart,
                  hipEvent_t evStop)
{
	if (end <= begin)
		return;
	const dim3 grid(nblk(end - begin, SOLVE_THREADS)), block(SOLVE_THREADS);
	if (evStart)
		hipExtLaunchKernelGGL(k_solve, grid, block, 0, st, evStart, evStop, 0, begin, end, C, rows, sb.vel, minPen,
		                      muS, muD, invStepDt, biasScale, scale, broken);
	else
		k_solve<<<grid, block, 0, st>>>(begin, end, C, rows, sb.vel, minPen, muS, muD, invStepDt, biasScale, scale,
		                                broken);
}
void launch_color_keys_world(int C, const int* conColor, const int* conA, const int* bodyWorld, uint32_t limit,
                             uint64_t* keys, uint32_t* vals, hipStream_t st)
{
	PXH_LAUNCH(k_color_keys_world, C, C, conColor, conA, bodyWorld, limit, keys, vals);
}
int solve_world_threads() { return SOLVE_WORLD_THREADS; }
void launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st)
{
	PXH_LAUNCH(k_fill_u32, n, p, n, v);
}
void launch_solve_worlds(int grid, int nWorlds, const int* worldOrder, uint32_t* queueHead, const int* bodyStart,
                         const uint32_t* segStart, int CB, int C, Rows rows, SolverBodies sb, int posIters,
                         int velIters, float stepDt, float muS, float muD, float invStepDt, float biasScale,
                         float scale, uint32_t* broken, int ldsBodies, hipStream_t st, hipEvent_t evStart,
                         hipEvent_t evStop)
{
	if (nWorlds <= 0 || grid <= 0)
		return;
	// the queue starts at zero (a fill kernel, no memset packet)
	k_fill_u32<<<1, 64, 0, st>>>(queueHead, 1, 0u);
	const size_t shm = sizeof(float4) * 4 * size_t(ldsBodies);
	if (evStart)
		hipExtLaunchKernelGGL(k_solve_worlds, dim3(grid), dim3(SOLVE_WORLD_THREADS), shm, st, evStart, evStop, 0,
		                      nWorlds, worldOrder, queueHead, bodyStart, segStart, CB, C, rows, sb.vel, sb.p, sb.dq,
		                      sb.i0, sb.i1, posIters, velIters, stepDt, muS, muD, invStepDt, biasScale, scale, broken,
		                      ldsBodies);
	else
		k_solve_worlds<<<grid, SOLVE_WORLD_THREADS, shm, st>>>(nWorlds, worldOrder, queueHead, bodyStart, segStart, CB,
		                                                       C, rows, sb.vel, sb.p, sb.dq, sb.i0, sb.i1, posIters,
		                                                       velIters, stepDt, muS, muD, invStepDt, biasScale,
		                                                       scale, broken, ldsBodies);
}
void launch_prep_pgs(int C, const int* order, const int* conA, const int* conB, const int* conSlot,
                     const float4* ctN, const float4* ctP, const float4* pos, const float4* rot, SolverBodies sb,
                     float invDt, float muS, float muD, float restDist, float fot, Rows rows, float* rowsU,
                     PatchIO patches, unsigned long long* counters, hipStream_t st)
{
	(void)pos;
	(void)rot;
	(void)conA;
	(void)conB;
	(void)conSlot;
	PXH_LAUNCH(k_prep_pgs, C, C, order, ctN, ctP, sb.rec, invDt, muS, muD, restDist, fot, rows, rowsU, patches,
	           counters + 4);
	k_sum_counters<<<1, 256, 0, st>>>(nblk(C, 256), counters + 4, counters);
}
void launch_solve_pgs(int begin, int end, int C, Rows rows, const float* rowsU, SolverBodies sb, bool doFriction,
                      bool velocityPass, uint32_t* broken, hipStream_t st, hipEvent_t evStart, hipEvent_t evStop)
{
	if (end <= begin)
		return;
	const dim3 grid(nblk(end - begin, SOLVE_THREADS)), block(SOLVE_THREADS);
	if (evStart)
		hipExtLaunchKernelGGL(k_solve_pgs, grid, block, 0, st, evStart, evStop, 0, begin, end, C, rows, rowsU, sb.vel,
		                      doFriction ? 1 : 0, velocityPass ? 1 : 0, broken);
	else
		k_solve_pgs<<<grid, block, 0, st>>>(begin, end, C, rows, rowsU, sb.vel, doFriction ? 1 : 0,
		                                    velocityPass ? 1 : 0, broken);
}
void launch_pgs_save_motion(int n, SolverBodies sb, hipStream_t st)
{
	PXH_LAUNCH(k_pgs_save_motion, n, n, sb.vel, sb.p, sb.dq);
}
void launch_pgs_finalize(int n, float dt, float4* pos, float4* rot, float4* lin, float4* ang, const float4* invI,
                         SolverBodies sb, SleepState ss, hipStream_t st)
{
	PXH_LAUNCH(k_pgs_finalize, n, n, dt, pos, rot, lin, ang, invI, sb.vel, sb.p, sb.dq, sb.i0, sb.i1, ss);
}
void launch_integrate(int n, float dt, SolverBodies sb, hipStream_t st)
{
	PXH_LAUNCH(k_integrate, n, n, dt, sb.vel, sb.p, sb.dq, sb.i0, sb.i1);
}
void launch_finalize(int n, float dt, float4* pos, float4* rot, float4* lin, float4* ang, const float4* invI,
                     SolverBodies sb, SleepState ss, hipStream_t st)
{
	PXH_LAUNCH(k_finalize, n, n, dt, pos, rot, lin, ang, invI, sb.vel, sb.p, sb.dq, sb.i0, sb.i1, ss);
}
void launch_get_data(int n, const uint32_t* idx, int what, const float4* pos, const float4* rot, const float4* lin,
                     const float4* ang, const float4* prevLin, const float4* prevAng, float invDt, float* out,
                     hipStream_t st)
{
	PXH_LAUNCH(k_get_data, n, n, idx, what, pos, rot, lin, ang, prevLin, prevAng, invDt, out);
}
void launch_set_data(int n, const uint32_t* idx, int what, float4* pos, float4* rot, float4* lin, float4* ang,
                     const float4* invI, float4* extLin, float4* extAng, const float* in, SleepState ss,
                     hipStream_t st)
{
	PXH_LAUNCH(k_set_data, n, n, idx, what, pos, rot, lin, ang, invI, extLin, extAng, in, ss);
}
void launch_touch_events(int C, const int* prevRow, int R, uint32_t* found, uint32_t* matched, uint32_t* lost,
                         hipStream_t st)
{
	PXH_LAUNCH(k_fill_u32, int64_t(R) + 1, matched, int64_t(R) + 1, 0u);
	PXH_LAUNCH(k_found_flags, C, C, prevRow, found, matched);
	PXH_LAUNCH(k_lost_flags, R, R, matched, lost);
}
void launch_pair_keys(int P, const uint2* pairs, const uint32_t* ids, uint64_t* keys, uint32_t* vals, hipStream_t st)
{
	PXH_LAUNCH(k_pair_keys, P, P, pairs, ids, keys, vals);
}
void launch_set_diff(int na, const uint64_t* a, int nb, const uint64_t* b, uint32_t* flag, hipStream_t st)
{
	PXH_LAUNCH(k_set_diff_flags, na, na, a, nb, b, flag);
}
void launch_compact_keys(int na, const uint64_t* a, const uint32_t* flag, const uint32_t* scanned, uint2* out,
                         hipStream_t st)
{
	PXH_LAUNCH(k_compact_keys, na, na, a, flag, scanned, out);
}

void launch_np_immediate(int n, const int* type0, const float* geom0, const float* pose0, const int* type1,
                         const float* geom1, const float* pose1, float cd, float tol, float* out, int* counts,
                         hipStream_t st)
{
	PXH_LAUNCH(k_np_immediate, n, n, type0, geom0, pose0, type1, geom1, pose1, cd, tol, out, counts);
}

__global__ void k_bp_inflate(int n, const float* bounds, const float* dist, const uint32_t* groups, float4* bmin,
                             float4* bmax, int* scene, int* group, int* bodyStart)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i == 0)
	{
		bodyStart[0] = 0;
		bodyStart[1] = n;
	}
	if (i >= n)
		return;
	const float* b = bounds + 6 * int64_t(i);
	const float d = dist ? dist[i] : 0.0f;
	// geometry type slot unused by the standalone broadphase (all pairs class 2)
	bmin[i] = make_float4(b[0] - d, b[1] - d, b[2] - d, float(GEOM_BOX));
	bmax[i] = make_float4(b[3] + d, b[4] + d, b[5] + d, 0.0f);
	scene[i] = 0;
	group[i] = groups ? int(groups[i]) : i;
}

void launch_bp_inflate(int n, const float* bounds, const float* dist, const uint32_t* groups, float4* bmin,
                       float4* bmax, int* scene, int* group, int* bodyStart, hipStream_t st)
{
	k_bp_inflate<<<nblk(n > 0 ? n : 1, 256), 256, 0, st>>>(n, bounds, dist, groups, bmin, bmax, scene, group,
	                                                       bodyStart);
}

}  // namespace pxh
