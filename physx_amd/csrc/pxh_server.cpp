// Serving loop over one batched scene: step requests for independent scenes
// of different sizes arrive at any time, are admitted as world slots and HBM
// budget allow, are stepped together with every other admitted scene (one
// launch sequence per step for all of them) and leave the batch when their
// requested steps are done. A scene's bodies, friction patches and colours
// stay resident in HBM from admission to completion.
//
// The reference has no serving loop; this is the runtime around the plug-in
// boundary. Everything a scene computes is a function of its own content
// (world-local colouring keys, world-relative pair keys, per-world
// broadphase grids), so a scene's result is the same alone and in a mix.
#include "../../include/pxhip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

thread_local std::string gServerError;

int serverFail(int code, const char* msg)
{
	gServerError = msg;
	return code;
}

double nowSec()
{
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct PendingScene
{
	int64_t id;
	int n, m, steps;
	std::vector<int32_t> type;
	std::vector<float> dims, pose, lin, ang, mass, inertia, planes;
	double tSubmit;
};

struct ActiveScene
{
	int64_t id;
	int world;          // current world index in the batched scene
	int n, stepsLeft, stepsDone;
	int64_t bytes;      // admitted HBM estimate
	double tSubmit, tAdmit;
};

struct Result
{
	std::vector<float> pose, lin, ang;
	double tSubmit, tAdmit, tDone;
};

}  // namespace

struct pxh_server_t
{
	pxh_scene scene = nullptr;
	int64_t budget = 0, used = 0;
	int32_t maxWorlds = 0;
	int64_t bytesPerBody = 0;
	int64_t nextId = 1;
	std::deque<PendingScene> queue;
	std::vector<ActiveScene> active;    // in world order
	std::unordered_map<int64_t, Result> done;
	int64_t steps = 0, completed = 0, bodySteps = 0;
	int64_t deviceBytes = 0;   // the batched scene's device bytes after the last step
	// device staging for finished scenes' bodies (13 floats per body) and their indices
	float* devOut = nullptr;
	uint32_t* devIdx = nullptr;
	int devCap = 0;

	bool reserve(int count)
	{
		if (count <= devCap)
			return true;
		int cap = std::max(count, 2 * devCap);
		(void)hipFree(devOut);
		(void)hipFree(devIdx);
		devOut = nullptr;
		devIdx = nullptr;
		devCap = 0;
		if (hipMalloc(reinterpret_cast<void**>(&devOut), sizeof(float) * 13 * size_t(cap)) != hipSuccess ||
		    hipMalloc(reinterpret_cast<void**>(&devIdx), sizeof(uint32_t) * size_t(cap)) != hipSuccess)
			return false;
		devCap = cap;
		return true;
	}
};

extern "C" {

const char* pxh_server_last_error(void) { return gServerError.c_str(); }

int32_t pxh_server_create(const pxh_scene_desc* desc, int64_t hbm_budget_bytes, int32_t max_worlds,
                          int64_t bytes_per_body, pxh_server* out)
{
	if (!desc || !out || hbm_budget_bytes <= 0 || max_worlds <= 0 || bytes_per_body <= 0)
		return serverFail(PXH_ERR_INVALID, "pxh_server_create: bad arguments");
	pxh_server_t* sv = new pxh_server_t();
	const int32_t r = pxh_scene_create(desc, &sv->scene);
	if (r)
	{
		delete sv;
		return serverFail(r, pxh_last_error());
	}
	sv->budget = hbm_budget_bytes;
	sv->maxWorlds = max_worlds;
	sv->bytesPerBody = bytes_per_body;
	*out = sv;
	return PXH_OK;
}

int32_t pxh_server_release(pxh_server sv)
{
	if (!sv)
		return PXH_OK;
	pxh_scene_release(sv->scene);
	(void)hipFree(sv->devOut);
	(void)hipFree(sv->devIdx);
	delete sv;
	return PXH_OK;
}

int32_t pxh_server_submit(pxh_server sv, int32_t n, const int32_t* type, const float* dims, const float* pose,
                          const float* lin_vel, const float* ang_vel, const float* mass, const float* inertia,
                          int32_t m, const float* planes, int32_t steps, int64_t* out_id)
{
	if (!sv || n <= 0 || m < 0 || steps <= 0 || !type || !dims || !pose || !mass || !inertia ||
	    (m > 0 && !planes) || !out_id)
		return serverFail(PXH_ERR_INVALID, "pxh_server_submit: bad arguments");
	if (int64_t(n) * sv->bytesPerBody > sv->budget)
		return serverFail(PXH_ERR_CAPACITY, "pxh_server_submit: scene larger than the HBM budget");
	PendingScene p;
	p.id = sv->nextId++;
	p.n = n;
	p.m = m;
	p.steps = steps;
	p.type.assign(type, type + n);
	p.dims.assign(dims, dims + 4 * size_t(n));
	p.pose.assign(pose, pose + 7 * size_t(n));
	p.lin.assign(size_t(3) * n, 0.0f);
	p.ang.assign(size_t(3) * n, 0.0f);
	if (lin_vel)
		std::memcpy(p.lin.data(), lin_vel, sizeof(float) * 3 * size_t(n));
	if (ang_vel)
		std::memcpy(p.ang.data(), ang_vel, sizeof(float) * 3 * size_t(n));
	p.mass.assign(mass, mass + n);
	p.inertia.assign(inertia, inertia + 3 * size_t(n));
	p.planes.assign(planes, planes + 4 * size_t(m));
	p.tSubmit = nowSec();
	sv->queue.push_back(std::move(p));
	*out_id = sv->queue.back().id;
	return PXH_OK;
}

// One serving step: admit queued scenes (in arrival order, while the world
// slots and the HBM budget allow), step every admitted scene once, copy the
// finished scenes' bodies to host results and take their worlds out.
int32_t pxh_server_step(pxh_server sv, float dt, int32_t* out_admitted, int32_t* out_completed)
{
	if (!sv || !(dt > 0.0f))
		return serverFail(PXH_ERR_INVALID, "pxh_server_step: bad arguments");
	// admission: queued scenes in arrival order while slots and budget allow,
	// appended as new worlds with one add call for all of them. Each body is
	// charged bytes_per_body; besides, while the batched scene's measured
	// device bytes exceed the budget, nothing is admitted until the admitted
	// scenes have finished (an idle server always admits the next scene, so
	// the queue keeps moving). The batch is staged first; server state
	// changes only once the scene holds it.
	int admitted = 0;
	if (sv->active.empty() || sv->deviceBytes <= sv->budget)
	{
		const int world0 = pxh_scene_world_count(sv->scene);
		if (world0 < 0)
			return serverFail(world0, pxh_last_error());
		int64_t used = sv->used;
		std::vector<int32_t> type, wb, wp;
		std::vector<float> dims, pose, lin, ang, mass, inertia, planes;
		std::vector<ActiveScene> staged;
		const double tAdmit = nowSec();
		for (const PendingScene& p : sv->queue)
		{
			const int64_t bytes = int64_t(p.n) * sv->bytesPerBody;
			if (int(sv->active.size() + staged.size()) >= sv->maxWorlds || used + bytes > sv->budget)
				break;
			const int world = world0 + int(staged.size());
			type.insert(type.end(), p.type.begin(), p.type.end());
			dims.insert(dims.end(), p.dims.begin(), p.dims.end());
			pose.insert(pose.end(), p.pose.begin(), p.pose.end());
			lin.insert(lin.end(), p.lin.begin(), p.lin.end());
			ang.insert(ang.end(), p.ang.begin(), p.ang.end());
			mass.insert(mass.end(), p.mass.begin(), p.mass.end());
			inertia.insert(inertia.end(), p.inertia.begin(), p.inertia.end());
			planes.insert(planes.end(), p.planes.begin(), p.planes.end());
			wb.insert(wb.end(), size_t(p.n), world);
			wp.insert(wp.end(), size_t(p.m), world);
			staged.push_back(ActiveScene{p.id, world, p.n, p.steps, 0, bytes, p.tSubmit, tAdmit});
			used += bytes;
		}
		if (!type.empty())
		{
			const int32_t r = pxh_scene_add_bodies(sv->scene, int32_t(type.size()), type.data(), dims.data(),
			                                       pose.data(), lin.data(), ang.data(), mass.data(), inertia.data(),
			                                       wb.data());
			if (r < 0)
				return serverFail(r, pxh_last_error());
		}
		if (!wp.empty())
		{
			const int32_t r = pxh_scene_add_planes(sv->scene, int32_t(wp.size()), planes.data(), wp.data());
			if (r)
			{
				// take the staged worlds' bodies out again: the scene is as before
				std::string why = pxh_last_error();
				std::vector<int32_t> ws;
				for (int w = world0; w < pxh_scene_world_count(sv->scene); w++)
					ws.push_back(w);
				(void)pxh_scene_remove_worlds(sv->scene, ws.data(), int32_t(ws.size()));
				return serverFail(r, why.c_str());
			}
		}
		for (ActiveScene& a : staged)
		{
			sv->used += a.bytes;
			sv->active.push_back(a);
			sv->queue.pop_front();
			admitted++;
		}
	}
	int completed = 0;
	if (!sv->active.empty())
	{
		int32_t r = pxh_scene_simulate(sv->scene, dt);
		if (r)
			return serverFail(r, pxh_last_error());
		r = pxh_scene_fetch_results(sv->scene, 1);
		if (r)
			return serverFail(r, pxh_last_error());
		sv->steps++;
		{
			int64_t st[12] = {};
			if (pxh_scene_get_stats(sv->scene, st, 12) == 12)
				sv->deviceBytes = st[11];
		}
		// finished scenes: their bodies in one gather per quantity, one copy out
		std::vector<int32_t> finished;
		std::vector<uint32_t> idx;
		std::vector<size_t> at;
		for (ActiveScene& a : sv->active)
		{
			a.stepsLeft--;
			a.stepsDone++;
			sv->bodySteps += a.n;
			if (a.stepsLeft > 0)
				continue;
			int32_t first = 0, count = 0;
			r = pxh_scene_world_range(sv->scene, a.world, &first, &count);
			if (r)
				return serverFail(r, pxh_last_error());
			at.push_back(idx.size());
			for (int i = 0; i < count; i++)
				idx.push_back(uint32_t(first + i));
			finished.push_back(a.world);
		}
		if (!finished.empty())
		{
			const int total = int(idx.size());
			if (!sv->reserve(total))
				return serverFail(PXH_ERR_DEVICE, "pxh_server_step: result buffers");
			hipStream_t st = static_cast<hipStream_t>(pxh_scene_get_stream(sv->scene));
			float* dev = sv->devOut;
			std::vector<float> host(size_t(13) * total);
			if (total > 0)
			{
				if (hipMemcpyAsync(sv->devIdx, idx.data(), sizeof(uint32_t) * size_t(total), hipMemcpyHostToDevice,
				                   st) != hipSuccess)
					return serverFail(PXH_ERR_DEVICE, "pxh_server_step: index copy");
				r = pxh_scene_get_rigid_dynamic_data(sv->scene, dev, sv->devIdx, PXH_DATA_GLOBAL_POSE, total, st);
				if (!r)
					r = pxh_scene_get_rigid_dynamic_data(sv->scene, dev + 7 * size_t(total), sv->devIdx,
					                                     PXH_DATA_LINEAR_VELOCITY, total, st);
				if (!r)
					r = pxh_scene_get_rigid_dynamic_data(sv->scene, dev + 10 * size_t(total), sv->devIdx,
					                                     PXH_DATA_ANGULAR_VELOCITY, total, st);
				if (r)
					return serverFail(r, pxh_last_error());
				if (hipMemcpyAsync(host.data(), dev, host.size() * sizeof(float), hipMemcpyDeviceToHost, st) !=
				        hipSuccess ||
				    hipStreamSynchronize(st) != hipSuccess)
					return serverFail(PXH_ERR_DEVICE, "pxh_server_step: reading finished scenes");
			}
			const double tDone = nowSec();
			size_t k = 0;
			for (const ActiveScene& a : sv->active)
			{
				if (a.stepsLeft > 0)
					continue;
				const size_t o = at[k++];
				Result res;
				res.tSubmit = a.tSubmit;
				res.tAdmit = a.tAdmit;
				res.tDone = tDone;
				res.pose.resize(size_t(7) * a.n);
				res.lin.resize(size_t(3) * a.n);
				res.ang.resize(size_t(3) * a.n);
				for (int i = 0; i < a.n; i++)
				{
					// PxTransform (q, p) -> p xyz, q xyzw
					const float* t = host.data() + 7 * (o + i);
					float* q = res.pose.data() + 7 * size_t(i);
					q[0] = t[4]; q[1] = t[5]; q[2] = t[6]; q[3] = t[0]; q[4] = t[1]; q[5] = t[2]; q[6] = t[3];
					for (int c = 0; c < 3; c++)
					{
						res.lin[3 * size_t(i) + c] = host[7 * size_t(total) + 3 * (o + i) + c];
						res.ang[3 * size_t(i) + c] = host[10 * size_t(total) + 3 * (o + i) + c];
					}
				}
				sv->done.emplace(a.id, std::move(res));
				sv->used -= a.bytes;
				completed++;
			}
			r = pxh_scene_remove_worlds(sv->scene, finished.data(), int32_t(finished.size()));
			if (r)
				return serverFail(r, pxh_last_error());
			std::vector<ActiveScene> keep;
			for (const ActiveScene& a : sv->active)
				if (a.stepsLeft > 0)
					keep.push_back(a);
			for (size_t i = 0; i < keep.size(); i++)
				keep[i].world = int(i);  // worlds stay in admission order after removal
			sv->active.swap(keep);
			sv->completed += completed;
		}
	}
	if (out_admitted)
		*out_admitted = admitted;
	if (out_completed)
		*out_completed = completed;
	return PXH_OK;
}

// Final bodies of a finished scene (pose p xyz + q xyzw, velocities) and its
// submit/admit/finish times in seconds (steady clock); the result is released.
int32_t pxh_server_result(pxh_server sv, int64_t id, float* pose, float* lin_vel, float* ang_vel, double* times)
{
	if (!sv)
		return serverFail(PXH_ERR_INVALID, "pxh_server_result: null server");
	auto it = sv->done.find(id);
	if (it == sv->done.end())
		return serverFail(PXH_ERR_STATE, "pxh_server_result: scene not finished (or already read)");
	const Result& r = it->second;
	if (pose)
		std::memcpy(pose, r.pose.data(), r.pose.size() * sizeof(float));
	if (lin_vel)
		std::memcpy(lin_vel, r.lin.data(), r.lin.size() * sizeof(float));
	if (ang_vel)
		std::memcpy(ang_vel, r.ang.data(), r.ang.size() * sizeof(float));
	if (times)
	{
		times[0] = r.tSubmit;
		times[1] = r.tAdmit;
		times[2] = r.tDone;
	}
	sv->done.erase(it);
	return PXH_OK;
}

// Ids of the finished scenes whose results are not read yet, ascending.
int32_t pxh_server_finished(pxh_server sv, int64_t* ids, int32_t max_ids)
{
	if (!sv || max_ids < 0 || (max_ids > 0 && !ids))
		return serverFail(PXH_ERR_INVALID, "pxh_server_finished: bad arguments");
	std::vector<int64_t> all;
	all.reserve(sv->done.size());
	for (const auto& kv : sv->done)
		all.push_back(kv.first);
	std::sort(all.begin(), all.end());
	const int32_t k = int32_t(std::min<size_t>(all.size(), size_t(max_ids)));
	std::copy(all.begin(), all.begin() + k, ids);
	return k;
}

// out[0..7]: queued scenes, admitted scenes, admitted bodies, HBM bytes
// admitted (at the admission price), finished scenes, serving steps,
// body-steps, device bytes the batched scene held after the last step
int32_t pxh_server_stats(pxh_server sv, int64_t* out, int32_t n)
{
	if (!sv || !out || n < 0)
		return serverFail(PXH_ERR_INVALID, "pxh_server_stats: bad arguments");
	int64_t bodies = 0;
	for (const ActiveScene& a : sv->active)
		bodies += a.n;
	const int64_t v[8] = {int64_t(sv->queue.size()), int64_t(sv->active.size()), bodies, sv->used, sv->completed,
	                      sv->steps, sv->bodySteps, sv->deviceBytes};
	for (int i = 0; i < n && i < 8; i++)
		out[i] = v[i];
	return std::min(n, 8);
}

}  // extern "C"
