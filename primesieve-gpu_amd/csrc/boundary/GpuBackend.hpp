///
/// @file  GpuBackend.hpp
/// @brief The backend boundary of the count path, kept member for member as
///        the reference defines it (src/gpu/GpuBackend.hpp:26-98): DeviceInfo,
///        GpuConfig (its four tunables and nothing else), the abstract
///        GpuBackend with count(lo >= 7, hi, countType), listDevices() and
///        createBackend().  The three class definitions are layout- and
///        member-identical to the reference's, so the HIP backend compiles
///        against either header and one program may hold both.  The
///        gfx950 HipBackend is the only implementation; no per-vendor #if
///        dispatch.  Its own extra knobs (bucket tier, slab size) are not
///        part of this boundary: see psg::HipTuning in HipBackend.hpp.
///
///        This header sits in its own include directory (not beside
///        HipBackend.hpp) so that a build can put the reference's
///        src/gpu/GpuBackend.hpp in its place.
///
#ifndef PRIMESIEVE_GPU_BACKEND_HPP
#define PRIMESIEVE_GPU_BACKEND_HPP

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
  /// LDS bytes per segment (rounded down to a multiple of 256).
  uint32_t sieveBytes = 0;
  /// Threads per workgroup, rounded down to a power of two (64 .. 1024).
  uint32_t workGroupSize = 0;
  /// Segments per persistent launch (one slab); 0 = 256 per compute unit.
  uint32_t segmentsPerLaunch = 0;
  /// Pre-sieve tiles for the primes 7..47.
  bool preSieve = true;
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
