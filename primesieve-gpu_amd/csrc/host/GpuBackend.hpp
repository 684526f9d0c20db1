///
/// @file  GpuBackend.hpp
/// @brief The backend boundary, kept as the reference defines it
///        (src/gpu/GpuBackend.hpp:26-98): DeviceInfo, GpuConfig, the abstract
///        GpuBackend with count(lo >= 7, hi, countType), listDevices() and
///        createBackend().  There is exactly one implementation, the gfx950
///        HipBackend; no per-vendor #if dispatch.
///
#ifndef PSG_GPU_BACKEND_HPP
#define PSG_GPU_BACKEND_HPP

#include <stdint.h>
#include <memory>
#include <string>
#include <vector>

namespace primesieve {
namespace gpu {

struct DeviceInfo
{
  std::string backend;       // "HIP"
  std::string name;
  std::string vendor;
  std::string version;
  int      index = -1;       // index in listDevices()
  int      platformIndex = 0;
  int      deviceIndex = 0;  // HIP device ordinal
  uint32_t computeUnits = 0;
  uint32_t clockMHz = 0;
  uint64_t globalMemBytes = 0;
  uint64_t localMemBytes = 0;
  uint64_t maxAllocBytes = 0;
  uint64_t maxWorkGroupSize = 0;

  std::string describe() const;
};

/// Tunables; 0 means "derive a default from the device".
struct GpuConfig
{
  uint32_t sieveBytes = 0;         // LDS bytes per segment (multiple of 256)
  uint32_t workGroupSize = 0;      // threads per workgroup, multiple of 64
  uint32_t segmentsPerLaunch = 0;  // segments per kernel launch
  bool preSieve = true;            // pre-sieve tiles for 7..47
  uint32_t slabSegments = 0;       // segments per big-prime bucket slab
  uint32_t bucketFactor = 0;       // primes > bucketFactor * sieveBytes use the
                                   // bucket tier; 0 = auto (7), ~0u = never
  uint32_t bucketMargin = 0;       // bucket list room beyond the expected count,
                                   // in standard deviations; 0 = auto (6)
};

class GpuBackend
{
public:
  virtual ~GpuBackend() {}
  virtual const DeviceInfo& device() const = 0;

  /// Primes (countType 0) or k-tuplets (1..5) in [lo, hi]; lo >= 7, the
  /// caller handles 2, 3 and 5 (reference src/gpu/GpuBackend.hpp:66-69).
  virtual uint64_t count(uint64_t lo, uint64_t hi, int countType) = 0;

  virtual uint32_t sieveBytes() const = 0;
  virtual uint32_t workGroupSize() const = 0;
};

std::vector<DeviceInfo> listDevices();

/// deviceIndex < 0 picks the device with the most compute units.
std::unique_ptr<GpuBackend> createBackend(int deviceIndex,
                                          const GpuConfig& config,
                                          std::string& error);

/// The gfx950 HIP implementation (HipBackend.cpp).
void listHipDevices(std::vector<DeviceInfo>& out);
std::unique_ptr<GpuBackend> createHipBackend(const DeviceInfo& dev,
                                             const GpuConfig& config,
                                             std::string& error);

} // namespace gpu
} // namespace primesieve

#endif
