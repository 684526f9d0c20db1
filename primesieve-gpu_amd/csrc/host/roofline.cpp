///
/// @file  roofline.cpp
/// @brief Measured ceilings of the device the backend runs on: LDS ds_and_b32
///        throughput (conflict-free and random addresses), LDS read streaming
///        and HBM copy bandwidth.  No reference counterpart; the sieve's
///        roofline fractions are stated against these numbers.
///

#include "GpuSieve.hpp"
#include "HipBackend.hpp"

#include <primesieve/primesieve_error.hpp>
#include <primesieve_gpu.h>

#include <cstring>
#include <mutex>
#include <string>

extern "C" const unsigned char psg_bench_hsaco_start[];

namespace {

struct Module
{
  const psg::HipApi* api;
  hipModule_t mod = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  explicit Module(const psg::HipApi* a) : api(a) {}
  ~Module()
  {
    if (e0) (void) api->EventDestroy(e0);
    if (e1) (void) api->EventDestroy(e1);
    if (stream) (void) api->StreamDestroy(stream);
    if (mod) (void) api->ModuleUnload(mod);
  }
};

double timeLaunch(Module& m, hipFunction_t f, unsigned grid, unsigned block, unsigned lds,
                  void** params)
{
  const psg::HipApi* api = m.api;
  // one untimed launch, then the timed one
  for (int rep = 0; rep < 2; rep++)
  {
    psg::hipCheck(api, api->EventRecord(m.e0, m.stream), "hipEventRecord");
    psg::hipCheck(api, api->ModuleLaunchKernel(f, grid, 1, 1, block, 1, 1, lds, m.stream,
                                               params, nullptr), "roofline launch");
    psg::hipCheck(api, api->EventRecord(m.e1, m.stream), "hipEventRecord");
    psg::hipCheck(api, api->StreamSynchronize(m.stream), "roofline kernel");
  }
  float ms = 0;
  psg::hipCheck(api, api->EventElapsedTime(&ms, m.e0, m.e1), "hipEventElapsedTime");
  return ms / 1e3;
}

} // namespace

extern "C" int psg_measure_roofline(psg_roofline* out)
{
  if (!out)
    return -1;
  std::memset(out, 0, sizeof *out);
  std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
  try
  {
    psg::HipBackend* b = primesieve::gpu::activeHipBackend();
    if (!b)
      throw primesieve::primesieve_error("GPU is not available: " + primesieve::gpu::lastErrorRef());
    std::string err;
    const psg::HipApi* api = psg::hipApi(err);
    psg::hipCheck(api, api->SetDevice(b->device().deviceIndex), "hipSetDevice");
    Module m(api);
    psg::hipCheck(api, api->ModuleLoadData(&m.mod, psg_bench_hsaco_start), "loading bench code object");
    psg::hipCheck(api, api->StreamCreateWithFlags(&m.stream, hipStreamNonBlocking), "hipStreamCreate");
    psg::hipCheck(api, api->EventCreate(&m.e0), "hipEventCreate");
    psg::hipCheck(api, api->EventCreate(&m.e1), "hipEventCreate");
    hipFunction_t fAnd, fRead, fCopy;
    psg::hipCheck(api, api->ModuleGetFunction(&fAnd, m.mod, "psg_lds_and"), "psg_lds_and");
    psg::hipCheck(api, api->ModuleGetFunction(&fRead, m.mod, "psg_lds_read"), "psg_lds_read");
    psg::hipCheck(api, api->ModuleGetFunction(&fCopy, m.mod, "psg_hbm_copy"), "psg_hbm_copy");

    const uint32_t cus = b->device().computeUnits ? b->device().computeUnits : 256;
    out->compute_units = cus;
    const unsigned grid = cus * 2, block = 1024, ldsBytes = 64 * 1024;
    psg::DeviceBuffer sink;
    sink.allocate(api, grid * 4);

    uint32_t words = ldsBytes / 4, iters = 8192, mode = 0;
    uint32_t* sinkp = (uint32_t*) sink.get();
    void* pAnd[] = { &words, &iters, &mode, &sinkp };
    const double atoms = (double) grid * block * iters;
    double t = timeLaunch(m, fAnd, grid, block, ldsBytes, pAnd);
    out->and_conflict_free_gops = atoms / t / 1e9;
    out->and_conflict_free_gbs = 8.0 * atoms / t / 1e9;
    mode = 1;
    t = timeLaunch(m, fAnd, grid, block, ldsBytes, pAnd);
    out->and_random_gops = atoms / t / 1e9;
    out->and_random_gbs = 8.0 * atoms / t / 1e9;
    mode = 2;
    t = timeLaunch(m, fAnd, grid, block, ldsBytes, pAnd);
    out->or_conflict_free_gbs = 8.0 * atoms / t / 1e9;
    mode = 3;
    t = timeLaunch(m, fAnd, grid, block, ldsBytes, pAnd);
    out->or_random_gbs = 8.0 * atoms / t / 1e9;

    void* pRead[] = { &words, &iters, &sinkp };
    t = timeLaunch(m, fRead, grid, block, ldsBytes, pRead);
    out->lds_read_gbs = 8.0 * atoms / t / 1e9;

    const uint64_t bytes = 1ull << 30;
    psg::DeviceBuffer src, dst;
    src.allocate(api, bytes);
    dst.allocate(api, bytes);
    void* ps = src.get();
    void* pd = dst.get();
    uint64_t n16 = bytes / 16;
    void* pCopy[] = { &ps, &pd, &n16 };
    t = timeLaunch(m, fCopy, cus * 16, 256, 0, pCopy);
    out->hbm_copy_gbs = 2.0 * (double) bytes / t / 1e9;
    return 0;
  }
  catch (const std::exception& e)
  {
    primesieve::gpu::lastErrorRef() = e.what();
    return -1;
  }
}
