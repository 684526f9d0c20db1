/* A stand-in HIP runtime for the CPU failure-handling tests: exports the
 * entry points the backend resolves, and fails on purpose according to
 * FAKE_HIP_MODE:
 *   nodevice  hipGetDeviceCount reports 0 devices
 *   noload    one device; loading the code object fails
 *   nomalloc  one device, code object loads; hipMalloc fails
 *   nolaunch  everything succeeds (allocations are host memory) except
 *             kernel launches, which fail: a device failure mid-count
 * Nothing here touches a GPU. */
#include <stdlib.h>
#include <string.h>
#include <stddef.h>

typedef int hipError_t;
enum { ok = 0, errInvalid = 1, errOOM = 2, errNoBinary = 209 };

static const char* mode(void) { const char* m = getenv("FAKE_HIP_MODE"); return m ? m : "nodevice"; }

hipError_t hipGetDeviceCount(int* n) { *n = strcmp(mode(), "nodevice") ? 1 : 0; return ok; }
hipError_t hipDeviceGetAttribute(int* v, int attr, int dev) { (void) attr; (void) dev; *v = 64 * 1024; return ok; }
hipError_t hipDeviceGetName(char* name, int len, int dev) { (void) dev; strncpy(name, "Fake GPU", (size_t) len); return ok; }
hipError_t hipDeviceTotalMem(size_t* b, int dev) { (void) dev; *b = (size_t) 1 << 30; return ok; }
hipError_t hipSetDevice(int dev) { (void) dev; return ok; }
hipError_t hipDriverGetVersion(int* v) { *v = 1; return ok; }
hipError_t hipRuntimeGetVersion(int* v) { *v = 1; return ok; }
const char* hipGetErrorString(hipError_t e) { return e == errOOM ? "fake out of memory" : "fake HIP error"; }
hipError_t hipModuleLoadData(void** m, const void* image) { (void) image; *m = (void*) 1; return strcmp(mode(), "noload") ? ok : errNoBinary; }
hipError_t hipModuleUnload(void* m) { (void) m; return ok; }
hipError_t hipModuleGetFunction(void** f, void* m, const char* n) { (void) m; (void) n; *f = (void*) 1; return ok; }
hipError_t hipModuleGetGlobal(void** p, size_t* b, void* m, const char* n)
{
  static void* arr = NULL;                      /* the code object's static arrays: host memory here */
  (void) m; (void) n;
  if (!arr) arr = calloc(1, (size_t) 64 << 20);
  *p = arr; *b = (size_t) 64 << 20;
  return arr ? ok : errOOM;
}
hipError_t hipModuleLaunchKernel() { return errInvalid; }
hipError_t hipModuleOccupancyMaxActiveBlocksPerMultiprocessor() { return errInvalid; }
hipError_t hipFuncGetAttribute(int* v, int attr, void* f) { (void) attr; (void) f; *v = 0; return ok; }
static int nolaunch(void) { return strcmp(mode(), "nolaunch") == 0; }
hipError_t hipMalloc(void** p, size_t n) { if (!nolaunch()) { *p = NULL; return errOOM; } *p = malloc(n ? n : 1); return ok; }
hipError_t hipFree(void* p) { if (nolaunch()) free(p); return ok; }
hipError_t hipHostMalloc(void** p, size_t n, unsigned f) { (void) f; *p = malloc(n); return ok; }
hipError_t hipHostFree(void* p) { free(p); return ok; }
hipError_t hipMemcpy(void* d, const void* s, size_t n, int kind) { (void) kind; if (!nolaunch()) return errInvalid; memcpy(d, s, n); return ok; }
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, int kind, void* st) { (void) st; return hipMemcpy(d, s, n, kind); }
hipError_t hipMemsetAsync(void* d, int v, size_t n, void* st) { (void) st; if (!nolaunch()) return errInvalid; memset(d, v, n); return ok; }
hipError_t hipStreamCreateWithFlags(void** s, unsigned f) { (void) f; *s = (void*) 1; return ok; }
hipError_t hipStreamDestroy(void* s) { (void) s; return ok; }
hipError_t hipStreamSynchronize(void* s) { (void) s; return ok; }
hipError_t hipStreamWaitEvent(void* s, void* e, unsigned f) { (void) s; (void) e; (void) f; return ok; }
hipError_t hipEventCreate(void** e) { *e = (void*) 1; return ok; }
hipError_t hipEventRecord() { return ok; }
hipError_t hipEventSynchronize() { return ok; }
hipError_t hipEventElapsedTime(float* ms, void* a, void* b) { (void) a; (void) b; *ms = 0; return ok; }
hipError_t hipEventDestroy(void* e) { (void) e; return ok; }
hipError_t hipStreamBeginCapture() { return errInvalid; }
hipError_t hipStreamEndCapture() { return errInvalid; }
hipError_t hipGraphInstantiate() { return errInvalid; }
hipError_t hipGraphLaunch() { return errInvalid; }
hipError_t hipGraphExecDestroy() { return ok; }
hipError_t hipGraphDestroy() { return ok; }
