///
/// @file  GpuSieve.cpp
/// @brief Device-independent half of the GPU count path: enumeration,
///        backend choice, the process-wide lazily created backend, the primes
///        2/3/5 that the wheel-30 bit array cannot represent, the 1e8
///        worthwhile threshold and the CPU-fallback policy.  Same semantics as
///        the reference (src/gpu/GpuSieve.cpp:27-365); the extra calls
///        (gpu_count_all, gpu_count_batch) use the same lock and errors.
///

#include "GpuSieve.hpp"
#include "GpuBackend.hpp"
#include "HipBackend.hpp"

#include <primesieve/gpu.hpp>
#include <primesieve/primesieve_error.hpp>

#include <algorithm>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace {

std::mutex g_mutex;
int g_deviceIndex = -1;
bool g_enabled = false;
primesieve::gpu::GpuConfig g_config;
std::unique_ptr<primesieve::gpu::GpuBackend> g_backend;
std::string g_lastError;

/// Below this range length a kernel launch plus the sieving-prime set-up
/// does not pay (reference src/gpu/GpuSieve.cpp:34-36).
const uint64_t kMinGpuRange = 100000000ull;

} // namespace

namespace primesieve {
namespace gpu {

std::string DeviceInfo::describe() const
{
  char buf[512];
  std::snprintf(buf, sizeof buf,
                "[%d] %s: %s (%s)\n"
                "      %u compute units, %u MHz, %llu MiB global, %llu KiB local, max work-group %llu",
                index, backend.c_str(), name.c_str(), version.c_str(),
                computeUnits, clockMHz,
                (unsigned long long) (globalMemBytes >> 20),
                (unsigned long long) (localMemBytes >> 10),
                (unsigned long long) maxWorkGroupSize);
  return std::string(buf);
}

std::vector<DeviceInfo> listDevices()
{
  std::vector<DeviceInfo> out;
  listHipDevices(out);
  for (std::size_t i = 0; i < out.size(); i++)
    out[i].index = (int) i;
  return out;
}

std::unique_ptr<GpuBackend> createBackend(int deviceIndex, const GpuConfig& config,
                                          std::string& error)
{
  std::vector<DeviceInfo> devices = listDevices();
  if (devices.empty())
  {
    std::string why;
    if (!psg::hipApi(why))
      error = "no GPU device found (" + why + ")";
    else
      error = "no GPU device found";
    return std::unique_ptr<GpuBackend>();
  }
  if (deviceIndex < 0)
  {
    std::size_t best = 0;
    for (std::size_t i = 1; i < devices.size(); i++)
      if (devices[i].computeUnits > devices[best].computeUnits)
        best = i;
    deviceIndex = (int) best;
  }
  if (deviceIndex >= (int) devices.size())
  {
    error = "GPU device index out of range";
    return std::unique_ptr<GpuBackend>();
  }
  return createHipBackend(devices[(std::size_t) deviceIndex], config, error);
}

/// The backend of the selected device, created on first use; nullptr when
/// no GPU is usable (g_lastError says why).  Caller holds g_mutex.
GpuBackend* activeBackend()
{
  if (!g_backend)
  {
    g_backend = createBackend(g_deviceIndex, g_config, g_lastError);
    if (!g_backend)
      return nullptr;
  }
  return g_backend.get();
}

psg::HipBackend* activeHipBackend()
{
  return dynamic_cast<psg::HipBackend*>(activeBackend());
}

psg::HipBackend* existingHipBackend()
{
  return dynamic_cast<psg::HipBackend*>(g_backend.get());
}

namespace {

/// 2, 3, 5 and the k-tuplets containing them, counted when start <= 5
/// (reference src/gpu/GpuSieve.cpp:144-169, src/PrimeSieveClass.cpp:38-48).
void addSmallPrimes(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  struct SmallTuple { uint64_t first; uint64_t last; int index; };
  static const SmallTuple small[8] = {
    { 2, 2, 0 }, { 3, 3, 0 }, { 5, 5, 0 }, { 3, 5, 1 },
    { 5, 7, 1 }, { 5, 11, 2 }, { 5, 13, 3 }, { 5, 17, 4 }
  };
  if (start > 5)
    return;
  for (const SmallTuple& t : small)
    if (t.first >= start && t.last <= stop)
      counts[t.index]++;
}

} // namespace

bool countAllOnGpu(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  std::fill(counts, counts + 6, 0);
  if (start > stop)
    return activeBackend() != nullptr;
  psg::HipBackend* backend = activeHipBackend();
  if (!backend)
    return false;
  addSmallPrimes(start, stop, counts);
  if (stop >= 7)
  {
    uint64_t c[6];
    backend->countAll(std::max<uint64_t>(start, 7), stop, c);
    for (int i = 0; i < 6; i++)
      counts[i] += c[i];
  }
  return true;
}

/// Counts of a batch of queries: out[i] = count of type k_i - 1 in
/// [start_i, stop_i].  mode 0 packs every query into the persistent engine;
/// mode 1 is the plain path, one count call (launch series and
/// synchronisation) per query, kept as the baseline the batch is measured
/// against.  No reference counterpart.
bool countBatchOnGpu(const GpuQuery* q, size_t n, uint64_t* out, int mode)
{
  for (size_t i = 0; i < n; i++)
    if (q[i].k < 1 || q[i].k > 6)
      throw primesieve_error("batch query k must be 1..6");
  psg::HipBackend* backend = activeHipBackend();
  if (!backend)
    return false;
  std::vector<psg::Interval> iv;
  iv.reserve(n);
  for (size_t i = 0; i < n; i++)
  {
    out[i] = 0;
    if (q[i].start > q[i].stop)
      continue;
    uint64_t small[6] = { 0, 0, 0, 0, 0, 0 };
    addSmallPrimes(q[i].start, q[i].stop, small);
    out[i] = small[q[i].k - 1];
    if (q[i].stop >= 7)
      iv.push_back(psg::Interval{ std::max<uint64_t>(q[i].start, 7), q[i].stop, (uint32_t) i });
  }
  if (mode == 1)
  {
    for (const psg::Interval& v : iv)
    {
      uint64_t c[6];
      backend->countAll(v.lo, v.hi, c);
      out[v.out] += c[q[v.out].k - 1];
    }
    return true;
  }
  std::vector<uint64_t> six(n * 6);
  backend->countBatch(iv, (uint32_t) n, six.data());
  for (const psg::Interval& v : iv)
    out[v.out] += six[(size_t) v.out * 6 + q[v.out].k - 1];
  return true;
}

/// Reference countOnGpu (src/gpu/GpuSieve.cpp:131-176).
uint64_t countOnGpu(uint64_t start, uint64_t stop, int countType, bool& usedGpu)
{
  usedGpu = false;
  if (start > stop)
    return 0;
  GpuBackend* backend = activeBackend();
  if (!backend)
    return 0;
  uint64_t small[6] = { 0, 0, 0, 0, 0, 0 };
  addSmallPrimes(start, stop, small);
  uint64_t total = small[countType];
  if (stop >= 7)
    total += backend->count(std::max<uint64_t>(start, 7), stop, countType);
  usedGpu = true;
  return total;
}

void setConfig(const GpuConfig& config)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_config = config;
  g_backend.reset();
}

std::mutex& gpuMutex() { return g_mutex; }
std::string& lastErrorRef() { return g_lastError; }
void resetBackend() { g_backend.reset(); }

} // namespace gpu

// ---------------------------------------------------------------------------
// Public API (include/primesieve/gpu.hpp)
// ---------------------------------------------------------------------------

bool gpu_available()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  return !gpu::listDevices().empty();
}

std::vector<std::string> gpu_devices()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  std::vector<gpu::DeviceInfo> devices = gpu::listDevices();
  std::vector<std::string> out;
  for (const gpu::DeviceInfo& d : devices)
    out.push_back(d.describe());
  return out;
}

void set_gpu_device(int deviceIndex)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (deviceIndex != g_deviceIndex)
  {
    g_deviceIndex = deviceIndex;
    g_backend.reset();
  }
}

int get_gpu_device()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_deviceIndex;
}

void set_gpu_enabled(bool enable)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_enabled = enable;
}

bool get_gpu_enabled()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_enabled;
}

void set_gpu_sieve_size(int kilobytes)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_config.sieveBytes = (kilobytes > 0) ? (uint32_t) kilobytes * 1024 : 0;
  g_backend.reset();
}

void set_gpu_work_group_size(int size)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_config.workGroupSize = (size > 0) ? (uint32_t) size : 0;
  g_backend.reset();
}

void set_gpu_segments_per_launch(int segments)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_config.segmentsPerLaunch = (segments > 0) ? (uint32_t) segments : 0;
  g_backend.reset();
}

void set_gpu_presieve(bool enable)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  g_config.preSieve = enable;
  g_backend.reset();
}

std::string gpu_last_error()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  return g_lastError;
}

std::string gpu_active_device()
{
  std::lock_guard<std::mutex> lock(g_mutex);
  gpu::GpuBackend* b = gpu::activeBackend();
  if (!b)
    return std::string();
  return b->device().describe();
}

bool gpu_worthwhile(uint64_t start, uint64_t stop)
{
  if (start > stop)
    return false;
  return (stop - start) >= kMinGpuRange;
}

namespace {

uint64_t gpuCountType(uint64_t start, uint64_t stop, int type)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  bool used = false;
  uint64_t n = gpu::countOnGpu(start, stop, type, used);
  if (!used)
    throw primesieve_error("GPU is not available: " + g_lastError);
  return n;
}

} // namespace

uint64_t gpu_count_primes(uint64_t start, uint64_t stop) { return gpuCountType(start, stop, 0); }
uint64_t gpu_count_twins(uint64_t start, uint64_t stop) { return gpuCountType(start, stop, 1); }
uint64_t gpu_count_triplets(uint64_t start, uint64_t stop) { return gpuCountType(start, stop, 2); }
uint64_t gpu_count_quadruplets(uint64_t start, uint64_t stop) { return gpuCountType(start, stop, 3); }
uint64_t gpu_count_quintuplets(uint64_t start, uint64_t stop) { return gpuCountType(start, stop, 4); }
uint64_t gpu_count_sextuplets(uint64_t start, uint64_t stop) { return gpuCountType(start, stop, 5); }

bool gpu_try_count(uint64_t start, uint64_t stop, int countType, uint64_t& result)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_enabled || !gpu_worthwhile(start, stop) || countType < 0 || countType > 5)
    return false;
  try
  {
    bool used = false;
    uint64_t n = gpu::countOnGpu(start, stop, countType, used);
    if (!used)
      return false;
    result = n;
    return true;
  }
  catch (const std::exception& e)
  {
    // a GPU failure must never break a count: remember why, drop the
    // backend, let the caller sieve on the CPU
    g_lastError = e.what();
    g_backend.reset();
    return false;
  }
}

void gpu_count_batch(const std::vector<GpuQuery>& queries, std::vector<uint64_t>& out)
{
  std::lock_guard<std::mutex> lock(g_mutex);
  out.assign(queries.size(), 0);
  if (!gpu::countBatchOnGpu(queries.data(), queries.size(), out.data(), 0))
    throw primesieve_error("GPU is not available: " + g_lastError);
}

void gpu_count_all(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!gpu::countAllOnGpu(start, stop, counts))
    throw primesieve_error("GPU is not available: " + g_lastError);
}

} // namespace primesieve
