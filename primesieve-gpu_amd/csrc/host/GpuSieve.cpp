///
/// @file  GpuSieve.cpp
/// @brief The device-independent half of the GPU count path for builds that
///        use this library on its own: the process-wide GPU state (device
///        choice, enable flag, configuration, the lazily created backend, the
///        last error), device listing, the primes 2, 3, 5 that the wheel-30
///        bit array cannot hold, the 1e8 worthwhile threshold and the
///        CPU-fallback policy.
///
///        Observable behaviour follows the reference's src/gpu/GpuSieve.cpp
///        (:27-365): the DeviceInfo::describe() text, the "no GPU device
///        found" / "GPU is not available: " errors, which setters drop the
///        backend, one lock around every entry.  A program that keeps the
///        reference's own GpuSieve.cpp links HipBackend.cpp without this file;
///        the calls below that the reference does not have (six counts in one
///        pass, batches, tuning) exist only in this build.
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

namespace primesieve {
namespace gpu {

namespace {

/// Everything the GPU API keeps between calls.
struct GpuState
{
  std::mutex mutex;
  int deviceIndex = -1;               // -1: the device with the most compute units
  bool enabled = false;               // gpu_try_count is off until enabled
  GpuConfig config;
  std::unique_ptr<GpuBackend> backend;
  std::string lastError;
};

GpuState& state()
{
  static GpuState s;
  return s;
}

/// Ranges shorter than this go to the CPU: a launch series plus the
/// sieving-prime set-up does not pay below it (reference
/// src/gpu/GpuSieve.cpp:34-36).
constexpr uint64_t kMinGpuRange = 100000000ull;

/// Lock the state and change its configuration; the backend is rebuilt on
/// next use.
template <typename Change>
void reconfigure(Change change)
{
  GpuState& s = state();
  std::lock_guard<std::mutex> lock(s.mutex);
  change(s.config);
  s.backend.reset();
}

/// Add 2, 3, 5 and the k-tuplets that contain them when start <= 5: the bit
/// array starts at 7 (reference src/gpu/GpuSieve.cpp:144-169,
/// src/PrimeSieveClass.cpp:38-48).  counts[t]: t = 0 primes .. 5 sextuplets.
void addSmallPrimes(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  if (start > 5)
    return;
  // (first member, last member, count type)
  static const uint64_t tuples[8][3] = {
    { 2, 2, 0 }, { 3, 3, 0 }, { 5, 5, 0 },   // primes
    { 3, 5, 1 }, { 5, 7, 1 },                // twins
    { 5, 11, 2 },                            // triplet (5, 7, 11)
    { 5, 13, 3 },                            // quadruplet (5, 7, 11, 13)
    { 5, 17, 4 } };                          // quintuplet (5, 7, 11, 13, 17)
  for (const auto& t : tuples)
    if (t[0] >= start && t[1] <= stop)
      counts[t[2]]++;
}

} // namespace

std::string DeviceInfo::describe() const
{
  // the text of the reference's --gpu-info (src/gpu/GpuSieve.cpp:43-55)
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
  std::vector<DeviceInfo> devices;
  listHipDevices(devices);
  int i = 0;
  for (DeviceInfo& d : devices)
    d.index = i++;
  return devices;
}

std::unique_ptr<GpuBackend> createBackend(int deviceIndex, const GpuConfig& config,
                                          std::string& error)
{
  const std::vector<DeviceInfo> devices = listDevices();
  if (devices.empty())
  {
    std::string why;
    error = psg::hipApi(why) ? "no GPU device found" : "no GPU device found (" + why + ")";
    return nullptr;
  }
  size_t pick = (size_t) deviceIndex;
  if (deviceIndex < 0)
  {
    // the most compute units (the discrete GPU on a machine with two)
    pick = (size_t) (std::max_element(devices.begin(), devices.end(),
                                      [](const DeviceInfo& a, const DeviceInfo& b) {
                                        return a.computeUnits < b.computeUnits;
                                      }) - devices.begin());
  }
  if (pick >= devices.size())
  {
    error = "GPU device index out of range";
    return nullptr;
  }
  return createHipBackend(devices[pick], config, error);
}

GpuBackend* activeBackend()
{
  GpuState& s = state();
  if (!s.backend)
    s.backend = createBackend(s.deviceIndex, s.config, s.lastError);
  return s.backend.get();
}

psg::HipBackend* activeHipBackend()
{
  return dynamic_cast<psg::HipBackend*>(activeBackend());
}

psg::HipBackend* existingHipBackend()
{
  return dynamic_cast<psg::HipBackend*>(state().backend.get());
}

std::mutex& gpuMutex() { return state().mutex; }
std::string& lastErrorRef() { return state().lastError; }
void resetBackend() { state().backend.reset(); }

void setTuning(const GpuConfig& config, const psg::HipTuning& tuning)
{
  psg::setHipTuning(tuning);
  reconfigure([&](GpuConfig& c) { c = config; });
}

/// The reference's countOnGpu (src/gpu/GpuSieve.cpp:131-176): usedGpu is
/// false when no backend can be created.
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

bool countAllOnGpu(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  std::fill(counts, counts + 6, 0);
  psg::HipBackend* backend = activeHipBackend();
  if (!backend)
    return false;
  if (start > stop)
    return true;
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

/// countAllOnGpu without device allocation: the backend must exist already
/// (creating it allocates), and the call's device memory is the code
/// object's scratch array.
bool countAllNoAllocOnGpu(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  std::fill(counts, counts + 6, 0);
  psg::HipBackend* backend = existingHipBackend();
  if (!backend)
    return false;
  if (start > stop)
    return true;
  addSmallPrimes(start, stop, counts);
  if (stop >= 7)
  {
    uint64_t c[6];
    backend->countAllNoAlloc(std::max<uint64_t>(start, 7), stop, c);
    for (int i = 0; i < 6; i++)
      counts[i] += c[i];
  }
  return true;
}

/// out[i] = count of type k_i - 1 in [start_i, stop_i].  mode 0 packs every
/// query into the persistent engine, mode 2 does the same through a captured
/// launch series replayed for every plan of the same shape, mode 1 is the
/// plain path (one count call, i.e. launch series and synchronisation, per
/// query), the baseline the batch is measured against.  No reference
/// counterpart.
bool countBatchOnGpu(const GpuQuery* q, size_t n, uint64_t* out, int mode)
{
  for (size_t i = 0; i < n; i++)
    if (q[i].k < 1 || q[i].k > 6)
      throw primesieve_error("batch query k must be 1..6");
  if (mode < 0 || mode > 2)
    throw primesieve_error("batch mode must be 0 (persistent), 1 (plain) or 2 (graph)");
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
  backend->countBatch(iv, (uint32_t) n, six.data(), mode == 2);
  for (const psg::Interval& v : iv)
    out[v.out] += six[(size_t) v.out * 6 + q[v.out].k - 1];
  return true;
}

} // namespace gpu

// ---------------------------------------------------------------------------
// Public API (include/primesieve/gpu.hpp)
// ---------------------------------------------------------------------------

bool gpu_available()
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  return !gpu::listDevices().empty();
}

std::vector<std::string> gpu_devices()
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  std::vector<std::string> out;
  for (const gpu::DeviceInfo& d : gpu::listDevices())
    out.push_back(d.describe());
  return out;
}

void set_gpu_device(int deviceIndex)
{
  gpu::GpuState& s = gpu::state();
  std::lock_guard<std::mutex> lock(s.mutex);
  if (deviceIndex == s.deviceIndex)
    return;                          // same device: keep the backend and its caches
  s.deviceIndex = deviceIndex;
  s.backend.reset();
}

int get_gpu_device()
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  return gpu::state().deviceIndex;
}

void set_gpu_enabled(bool enable)
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  gpu::state().enabled = enable;
}

bool get_gpu_enabled()
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  return gpu::state().enabled;
}

// The four tunables of the reference (each change rebuilds the backend;
// a value <= 0 restores the automatic choice).
void set_gpu_sieve_size(int kilobytes)
{
  gpu::reconfigure([&](gpu::GpuConfig& c) { c.sieveBytes = kilobytes > 0 ? (uint32_t) kilobytes * 1024u : 0u; });
}

void set_gpu_work_group_size(int size)
{
  gpu::reconfigure([&](gpu::GpuConfig& c) { c.workGroupSize = size > 0 ? (uint32_t) size : 0u; });
}

void set_gpu_segments_per_launch(int segments)
{
  gpu::reconfigure([&](gpu::GpuConfig& c) { c.segmentsPerLaunch = segments > 0 ? (uint32_t) segments : 0u; });
}

void set_gpu_presieve(bool enable)
{
  gpu::reconfigure([&](gpu::GpuConfig& c) { c.preSieve = enable; });
}

std::string gpu_last_error()
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  return gpu::lastErrorRef();
}

std::string gpu_active_device()
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  const gpu::GpuBackend* b = gpu::activeBackend();
  return b ? b->device().describe() : std::string();
}

bool gpu_worthwhile(uint64_t start, uint64_t stop)
{
  return start <= stop && stop - start >= gpu::kMinGpuRange;
}

namespace {

/// gpu_count_<type>: a GPU that cannot be used is an error.
uint64_t countOrThrow(uint64_t start, uint64_t stop, int type)
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  bool used = false;
  const uint64_t n = gpu::countOnGpu(start, stop, type, used);
  if (!used)
    throw primesieve_error("GPU is not available: " + gpu::lastErrorRef());
  return n;
}

} // namespace

uint64_t gpu_count_primes(uint64_t start, uint64_t stop) { return countOrThrow(start, stop, 0); }
uint64_t gpu_count_twins(uint64_t start, uint64_t stop) { return countOrThrow(start, stop, 1); }
uint64_t gpu_count_triplets(uint64_t start, uint64_t stop) { return countOrThrow(start, stop, 2); }
uint64_t gpu_count_quadruplets(uint64_t start, uint64_t stop) { return countOrThrow(start, stop, 3); }
uint64_t gpu_count_quintuplets(uint64_t start, uint64_t stop) { return countOrThrow(start, stop, 4); }
uint64_t gpu_count_sextuplets(uint64_t start, uint64_t stop) { return countOrThrow(start, stop, 5); }

bool gpu_try_count(uint64_t start, uint64_t stop, int countType, uint64_t& result)
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  if (!gpu::state().enabled || !gpu_worthwhile(start, stop) || countType < 0 || countType > 5)
    return false;
  try
  {
    bool used = false;
    const uint64_t n = gpu::countOnGpu(start, stop, countType, used);
    if (used)
      result = n;
    return used;
  }
  catch (const std::exception& e)
  {
    // a device failure never fails a count: keep the reason, drop the
    // backend and let the caller sieve on the CPU
    gpu::lastErrorRef() = e.what();
    gpu::resetBackend();
    return false;
  }
}

void gpu_count_batch(const std::vector<GpuQuery>& queries, std::vector<uint64_t>& out)
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  out.assign(queries.size(), 0);
  if (!gpu::countBatchOnGpu(queries.data(), queries.size(), out.data(), 0))
    throw primesieve_error("GPU is not available: " + gpu::lastErrorRef());
}

void gpu_count_all(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  std::lock_guard<std::mutex> lock(gpu::gpuMutex());
  if (!gpu::countAllOnGpu(start, stop, counts))
    throw primesieve_error("GPU is not available: " + gpu::lastErrorRef());
}

} // namespace primesieve
