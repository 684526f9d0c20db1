///
/// @file  HipBackend.hpp
/// @brief gfx950 implementation of GpuBackend.  Replaces the reference's
///        CudaBackend (src/gpu/cuda/CudaBackend.cpp:222-491): device set-up,
///        code-object load, sieving-prime cache, the launch loop and the host
///        reduction -- here the six counts are reduced on the device and one
///        48-byte copy ends a call.
///
#ifndef PSG_HIP_BACKEND_HPP
#define PSG_HIP_BACKEND_HPP

#include "GpuBackend.hpp"
#include "hip_api.hpp"
#include "tables.hpp"
#include "../kernel_abi.h"

#include <primesieve_gpu.h>

#include <stdint.h>
#include <string>
#include <vector>

namespace psg {

/// Device allocation owned by the backend.
class DeviceBuffer
{
public:
  DeviceBuffer() {}
  ~DeviceBuffer() { release(); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;

  void allocate(const HipApi* api, size_t bytes);   // keeps the old block if large enough
  void release();
  void* get() const { return ptr_; }
  uint64_t addr() const { return (uint64_t) (uintptr_t) ptr_; }
  size_t bytes() const { return bytes_; }

private:
  const HipApi* api_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

/// One query of a batch after host-side planning.
struct BatchItem
{
  uint64_t start;
  uint64_t stop;
  uint32_t k;        // 1..6
};

class HipBackend : public primesieve::gpu::GpuBackend
{
public:
  HipBackend(const HipApi* api, const primesieve::gpu::DeviceInfo& dev,
             const primesieve::gpu::GpuConfig& config);
  ~HipBackend() override;

  const primesieve::gpu::DeviceInfo& device() const override { return dev_; }
  uint64_t count(uint64_t lo, uint64_t hi, int countType) override;
  uint32_t sieveBytes() const override { return sieveBytes_; }
  uint32_t workGroupSize() const override { return wgSize_; }

  /// All six counts of [lo, hi] (lo >= 7 after clamping) in one pass.
  void countAll(uint64_t lo, uint64_t hi, uint64_t out[PSG_NCOUNTS]);

  const psg_stats& lastStats() const { return stats_; }

  /// The geometry actually in use.
  psg_tuning tuning() const
  {
    psg_tuning t;
    t.sieve_bytes = sieveBytes_;
    t.work_group_size = wgSize_;
    t.segments_per_launch = segmentsPerLaunch_;
    t.slab_segments = slabSegments_;
    t.bucket_factor = bucketFactor_ ? bucketFactor_ : 0xffffffffu;
    t.presieve = config_.preSieve ? 1u : 0u;
    return t;
  }

private:
  void init();
  void ensureSievingPrimes(uint32_t limit);
  void launchSieve(const PsgSieveArgs& args, uint32_t nseg);
  void recordPhase(size_t i);

  const HipApi* api_;
  primesieve::gpu::DeviceInfo dev_;
  primesieve::gpu::GpuConfig config_;
  uint32_t sieveBytes_ = 0;
  uint32_t wgSize_ = 0;
  uint32_t segmentsPerLaunch_ = 0;
  uint32_t bucketFactor_ = 1;     // primes > bucketFactor_ * S use the bucket tier (0 = never)
  uint32_t slabSegments_ = 0;     // segments per bucket slab
  uint32_t bucketChunks_ = 0;     // target workgroups of the bucket count / fill kernels
  uint32_t windowSegs_ = 64;      // segments per bucket window

  hipModule_t module_ = nullptr;
  hipFunction_t fnSieve_ = nullptr;
  hipFunction_t fnBucketCount_ = nullptr;
  hipFunction_t fnScanChunks_ = nullptr;
  hipFunction_t fnScanLists_ = nullptr;
  hipFunction_t fnBucketFill_ = nullptr;
  hipStream_t stream_ = nullptr;
  hipEvent_t evStart_ = nullptr;
  hipEvent_t evStop_ = nullptr;

  DeviceBuffer primes_;      // uint32 sieving primes >= first cross prime
  DeviceBuffer magics_;      // uint64 division magics
  DeviceBuffer tile0_;
  DeviceBuffer tile1_;
  DeviceBuffer counts_;      // uint64[6]
  uint64_t* hostCounts_ = nullptr;   // pinned
  DeviceBuffer bucketCounts_;        // uint32 [chunks][lists]
  DeviceBuffer listOff_;             // uint32 [lists + 1]
  DeviceBuffer entries_;             // bucket entries
  DeviceBuffer overflow_;            // uint32 flag
  uint32_t* hostOverflow_ = nullptr; // pinned
  std::vector<hipEvent_t> phaseEvents_;

  std::vector<uint32_t> hostPrimes_; // the primes held in primes_
  std::vector<double> invPrefix_;    // invPrefix_[i] = sum_{j<i} 1/p_j
  uint32_t primesLimit_ = 0;         // every prime <= primesLimit_ is held
  uint32_t firstCrossPrime_ = 29;

  psg_stats stats_{};
};

} // namespace psg

#endif
