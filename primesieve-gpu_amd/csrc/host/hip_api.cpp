///
/// @file  hip_api.cpp
/// @brief dlopen + dlsym resolution of the HIP runtime (see hip_api.hpp).
///

#include "hip_api.hpp"

#include <primesieve/primesieve_error.hpp>

#include <dlfcn.h>
#include <cstdlib>
#include <mutex>
#include <string>

namespace psg {
namespace {

std::once_flag g_once;
HipApi g_api;
bool g_ok = false;
std::string g_error;

template <typename F>
bool resolve(void* lib, const char* name, F& out, std::string& missing)
{
  void* sym = dlsym(lib, name);
  if (!sym)
  {
    if (!missing.empty()) missing += ", ";
    missing += name;
    return false;
  }
  out = reinterpret_cast<F>(sym);
  return true;
}

void load()
{
  const char* env = std::getenv("PSG_HIP_LIBRARY");
  const char* candidates[] = { env, "libamdhip64.so.7", "libamdhip64.so",
                               "/opt/rocm/lib/libamdhip64.so.7" };
  void* lib = nullptr;
  for (const char* c : candidates)
  {
    if (!c || !*c) continue;
    lib = dlopen(c, RTLD_NOW | RTLD_LOCAL);
    if (lib || c == env) break;
  }
  if (!lib)
  {
    const char* why = dlerror();
    g_error = std::string("HIP runtime not found (libamdhip64.so.7)") +
              (why ? std::string(": ") + why : std::string());
    return;
  }

  std::string missing;
  HipApi& a = g_api;
  bool ok = true;
  ok &= resolve(lib, "hipGetDeviceCount", a.GetDeviceCount, missing);
  ok &= resolve(lib, "hipDeviceGetAttribute", a.DeviceGetAttribute, missing);
  ok &= resolve(lib, "hipDeviceGetName", a.DeviceGetName, missing);
  ok &= resolve(lib, "hipDeviceTotalMem", a.DeviceTotalMem, missing);
  ok &= resolve(lib, "hipSetDevice", a.SetDevice, missing);
  ok &= resolve(lib, "hipDriverGetVersion", a.DriverGetVersion, missing);
  ok &= resolve(lib, "hipRuntimeGetVersion", a.RuntimeGetVersion, missing);
  ok &= resolve(lib, "hipGetErrorString", a.GetErrorString, missing);
  ok &= resolve(lib, "hipModuleLoadData", a.ModuleLoadData, missing);
  ok &= resolve(lib, "hipModuleUnload", a.ModuleUnload, missing);
  ok &= resolve(lib, "hipModuleGetFunction", a.ModuleGetFunction, missing);
  ok &= resolve(lib, "hipModuleGetGlobal", a.ModuleGetGlobal, missing);
  ok &= resolve(lib, "hipModuleLaunchKernel", a.ModuleLaunchKernel, missing);
  ok &= resolve(lib, "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor", a.ModuleOccupancy, missing);
  ok &= resolve(lib, "hipFuncGetAttribute", a.FuncGetAttribute, missing);
  ok &= resolve(lib, "hipMalloc", a.Malloc, missing);
  ok &= resolve(lib, "hipFree", a.Free, missing);
  ok &= resolve(lib, "hipHostMalloc", a.HostMalloc, missing);
  ok &= resolve(lib, "hipHostFree", a.HostFree, missing);
  ok &= resolve(lib, "hipMemcpy", a.Memcpy, missing);
  ok &= resolve(lib, "hipMemcpyAsync", a.MemcpyAsync, missing);
  ok &= resolve(lib, "hipMemsetAsync", a.MemsetAsync, missing);
  ok &= resolve(lib, "hipStreamCreateWithFlags", a.StreamCreateWithFlags, missing);
  ok &= resolve(lib, "hipStreamDestroy", a.StreamDestroy, missing);
  ok &= resolve(lib, "hipStreamSynchronize", a.StreamSynchronize, missing);
  ok &= resolve(lib, "hipStreamWaitEvent", a.StreamWaitEvent, missing);
  ok &= resolve(lib, "hipEventCreate", a.EventCreate, missing);
  ok &= resolve(lib, "hipEventRecord", a.EventRecord, missing);
  ok &= resolve(lib, "hipEventSynchronize", a.EventSynchronize, missing);
  ok &= resolve(lib, "hipEventElapsedTime", a.EventElapsedTime, missing);
  ok &= resolve(lib, "hipEventDestroy", a.EventDestroy, missing);
  ok &= resolve(lib, "hipStreamBeginCapture", a.StreamBeginCapture, missing);
  ok &= resolve(lib, "hipStreamEndCapture", a.StreamEndCapture, missing);
  ok &= resolve(lib, "hipGraphInstantiate", a.GraphInstantiate, missing);
  ok &= resolve(lib, "hipGraphLaunch", a.GraphLaunch, missing);
  ok &= resolve(lib, "hipGraphExecDestroy", a.GraphExecDestroy, missing);
  ok &= resolve(lib, "hipGraphDestroy", a.GraphDestroy, missing);

  if (!ok)
  {
    g_error = "HIP runtime is missing symbols: " + missing;
    return;
  }
  // the library handle stays open for the life of the process
  g_ok = true;
}

} // namespace

const HipApi* hipApi(std::string& error)
{
  std::call_once(g_once, load);
  if (!g_ok)
  {
    error = g_error;
    return nullptr;
  }
  return &g_api;
}

void hipCheck(const HipApi* api, hipError_t e, const char* what)
{
  if (e == hipSuccess)
    return;
  const char* s = api->GetErrorString ? api->GetErrorString(e) : nullptr;
  throw primesieve::primesieve_error(std::string(what) + ": " +
                                     (s ? s : "HIP error " + std::to_string((int) e)));
}

} // namespace psg
