///
/// @file  hip_api.hpp
/// @brief Runtime-resolved HIP entry points.  The host library never links
///        libamdhip64: it is dlopen'ed on first use and the functions below
///        are looked up by name, once per process (std::call_once), exactly
///        like the reference resolves the CUDA driver
///        (src/gpu/cuda/cuda_loader.hpp:18-54, cuda_loader.cpp:192-224).
///        A machine without ROCm therefore still loads the library and simply
///        reports "no GPU device found".
///
///        PSG_HIP_LIBRARY overrides the library path; the CPU tests use it to
///        substitute a fake runtime that fails on purpose.
///
#ifndef PSG_HIP_API_HPP
#define PSG_HIP_API_HPP

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#include <string>

namespace psg {

struct HipApi
{
  decltype(&hipGetDeviceCount) GetDeviceCount = nullptr;
  decltype(&hipDeviceGetAttribute) DeviceGetAttribute = nullptr;
  decltype(&hipDeviceGetName) DeviceGetName = nullptr;
  decltype(&hipDeviceTotalMem) DeviceTotalMem = nullptr;
  decltype(&hipSetDevice) SetDevice = nullptr;
  decltype(&hipDriverGetVersion) DriverGetVersion = nullptr;
  decltype(&hipRuntimeGetVersion) RuntimeGetVersion = nullptr;
  decltype(&hipGetErrorString) GetErrorString = nullptr;
  decltype(&hipModuleLoadData) ModuleLoadData = nullptr;
  decltype(&hipModuleUnload) ModuleUnload = nullptr;
  decltype(&hipModuleGetFunction) ModuleGetFunction = nullptr;
  decltype(&hipModuleGetGlobal) ModuleGetGlobal = nullptr;
  decltype(&hipModuleLaunchKernel) ModuleLaunchKernel = nullptr;
  decltype(&hipModuleOccupancyMaxActiveBlocksPerMultiprocessor) ModuleOccupancy = nullptr;
  decltype(&hipFuncGetAttribute) FuncGetAttribute = nullptr;
  hipError_t (*Malloc)(void**, size_t) = nullptr;
  decltype(&hipFree) Free = nullptr;
  hipError_t (*HostMalloc)(void**, size_t, unsigned int) = nullptr;
  decltype(&hipHostFree) HostFree = nullptr;
  decltype(&hipMemcpy) Memcpy = nullptr;
  decltype(&hipMemcpyAsync) MemcpyAsync = nullptr;
  decltype(&hipMemsetAsync) MemsetAsync = nullptr;
  decltype(&hipStreamCreateWithFlags) StreamCreateWithFlags = nullptr;
  decltype(&hipStreamDestroy) StreamDestroy = nullptr;
  decltype(&hipStreamSynchronize) StreamSynchronize = nullptr;
  decltype(&hipStreamWaitEvent) StreamWaitEvent = nullptr;
  decltype(&hipEventCreate) EventCreate = nullptr;
  decltype(&hipEventRecord) EventRecord = nullptr;
  decltype(&hipEventSynchronize) EventSynchronize = nullptr;
  decltype(&hipEventElapsedTime) EventElapsedTime = nullptr;
  decltype(&hipEventDestroy) EventDestroy = nullptr;
  decltype(&hipStreamBeginCapture) StreamBeginCapture = nullptr;
  decltype(&hipStreamEndCapture) StreamEndCapture = nullptr;
  decltype(&hipGraphInstantiate) GraphInstantiate = nullptr;
  decltype(&hipGraphLaunch) GraphLaunch = nullptr;
  decltype(&hipGraphExecDestroy) GraphExecDestroy = nullptr;
  decltype(&hipGraphDestroy) GraphDestroy = nullptr;
};

/// The resolved table, or nullptr when the runtime cannot be loaded; then
/// `error` says why.  Thread-safe, resolved once per process.
const HipApi* hipApi(std::string& error);

/// Throws primesieve_error("<what>: <hip error string>") when e != hipSuccess.
void hipCheck(const HipApi* api, hipError_t e, const char* what);

} // namespace psg

#endif
