// The HIP runtime calls of the GPU plug-in, behind plain C types so the
// plug-in's PhysX translation unit needs no HIP header (PhysX declares vector
// types of its own that clash with HIP's). Every function returns a
// CUresult-compatible code: 0 success, 1 invalid value, 2 out of memory, 600
// not ready, 999 any other failure.
#pragma once
#include <stddef.h>

namespace pxhrt {

enum { kOk = 0, kInvalid = 1, kOutOfMemory = 2, kNotReady = 600, kUnknown = 999 };
enum { kHostToDevice = 1, kDeviceToHost = 2, kDeviceToDevice = 3, kDefault = 4 };

int deviceCount(int* count);
int setDevice(int device);
int malloc(void** p, size_t bytes);
int free(void* p);
int hostMalloc(void** p, size_t bytes);   // pinned, mapped, portable
int hostFree(void* p);
int hostDevicePointer(void** d, void* host, unsigned flags);
int streamCreate(void** s, bool nonBlocking);
int streamDestroy(void* s);
int streamSynchronize(void* s);
int streamWaitEvent(void* s, void* e, unsigned flags);
int eventCreate(void** e, unsigned flags);
int eventRecord(void* e, void* s);
int eventQuery(void* e);
int eventSynchronize(void* e);
int eventDestroy(void* e);
int memcpy(void* dst, const void* src, size_t bytes, int kind);
int memcpyAsync(void* dst, const void* src, size_t bytes, int kind, void* s);
int memsetD32(void* dst, unsigned v, size_t n, void* s);   // s null: synchronous
int memsetD16(void* dst, unsigned short v, size_t n);
int memsetD8(void* dst, unsigned char v, size_t n, void* s);
int lastError();
int memInfo(size_t* free, size_t* total);
int driverVersion(int* v);

struct DeviceInfo
{
	char name[256];
	size_t totalMem, sharedPerBlock, sharedPerMP;
	int multiProcessors, clockKHz, maxThreadsPerBlock, integrated, canMapHostMemory;
};
int deviceInfo(int device, DeviceInfo* out);

}  // namespace pxhrt
