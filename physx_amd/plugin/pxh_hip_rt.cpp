// HIP runtime side of the GPU plug-in (see pxh_hip_rt.h).
#include "pxh_hip_rt.h"

#include <hip/hip_runtime_api.h>

#include <cstring>

namespace pxhrt {

namespace {
int code(hipError_t e)
{
	switch (e)
	{
	case hipSuccess: return kOk;
	case hipErrorInvalidValue: return kInvalid;
	case hipErrorOutOfMemory: return kOutOfMemory;
	case hipErrorNotReady: return kNotReady;
	default: return kUnknown;
	}
}
hipMemcpyKind kindOf(int k)
{
	switch (k)
	{
	case kHostToDevice: return hipMemcpyHostToDevice;
	case kDeviceToHost: return hipMemcpyDeviceToHost;
	case kDeviceToDevice: return hipMemcpyDeviceToDevice;
	default: return hipMemcpyDefault;
	}
}
hipStream_t st(void* s) { return static_cast<hipStream_t>(s); }
hipEvent_t ev(void* e) { return static_cast<hipEvent_t>(e); }
}  // namespace

int deviceCount(int* count) { return code(hipGetDeviceCount(count)); }
int setDevice(int device) { return code(hipSetDevice(device)); }
int malloc(void** p, size_t bytes) { return code(hipMalloc(p, bytes)); }
int free(void* p) { return code(hipFree(p)); }
int hostMalloc(void** p, size_t bytes)
{
	return code(hipHostMalloc(p, bytes, hipHostMallocMapped | hipHostMallocPortable));
}
int hostFree(void* p) { return code(hipHostFree(p)); }
int hostDevicePointer(void** d, void* host, unsigned flags) { return code(hipHostGetDevicePointer(d, host, flags)); }
int streamCreate(void** s, bool nonBlocking)
{
	hipStream_t h = nullptr;
	const int r = code(hipStreamCreateWithFlags(&h, nonBlocking ? hipStreamNonBlocking : hipStreamDefault));
	*s = h;
	return r;
}
int streamDestroy(void* s) { return code(hipStreamDestroy(st(s))); }
int streamSynchronize(void* s) { return code(hipStreamSynchronize(st(s))); }
int streamWaitEvent(void* s, void* e, unsigned flags) { return code(hipStreamWaitEvent(st(s), ev(e), flags)); }
int eventCreate(void** e, unsigned flags)
{
	hipEvent_t h = nullptr;
	const int r = code(hipEventCreateWithFlags(&h, flags));
	*e = h;
	return r;
}
int eventRecord(void* e, void* s) { return code(hipEventRecord(ev(e), st(s))); }
int eventQuery(void* e) { return code(hipEventQuery(ev(e))); }
int eventSynchronize(void* e) { return code(hipEventSynchronize(ev(e))); }
int eventDestroy(void* e) { return code(hipEventDestroy(ev(e))); }
int memcpy(void* dst, const void* src, size_t bytes, int kind) { return code(hipMemcpy(dst, src, bytes, kindOf(kind))); }
int memcpyAsync(void* dst, const void* src, size_t bytes, int kind, void* s)
{
	return code(hipMemcpyAsync(dst, src, bytes, kindOf(kind), st(s)));
}
int memsetD32(void* dst, unsigned v, size_t n, void* s)
{
	return code(s ? hipMemsetD32Async(hipDeviceptr_t(dst), int(v), n, st(s)) : hipMemsetD32(hipDeviceptr_t(dst), int(v), n));
}
int memsetD16(void* dst, unsigned short v, size_t n) { return code(hipMemsetD16(hipDeviceptr_t(dst), v, n)); }
int memsetD8(void* dst, unsigned char v, size_t n, void* s)
{
	return code(s ? hipMemsetD8Async(hipDeviceptr_t(dst), v, n, st(s)) : hipMemsetD8(hipDeviceptr_t(dst), v, n));
}
int lastError() { return code(hipGetLastError()); }
int memInfo(size_t* free, size_t* total) { return code(hipMemGetInfo(free, total)); }
int driverVersion(int* v) { return code(hipDriverGetVersion(v)); }

int deviceInfo(int device, DeviceInfo* out)
{
	hipDeviceProp_t p{};
	const int r = code(hipGetDeviceProperties(&p, device));
	if (r != kOk)
		return r;
	std::memset(out, 0, sizeof(*out));
	std::strncpy(out->name, p.name, sizeof(out->name) - 1);
	out->totalMem = p.totalGlobalMem;
	out->sharedPerBlock = p.sharedMemPerBlock;
	out->sharedPerMP = p.maxSharedMemoryPerMultiProcessor;
	out->multiProcessors = p.multiProcessorCount;
	out->clockKHz = p.clockRate;
	out->maxThreadsPerBlock = p.maxThreadsPerBlock;
	out->integrated = p.integrated;
	out->canMapHostMemory = p.canMapHostMemory;
	return kOk;
}

}  // namespace pxhrt
