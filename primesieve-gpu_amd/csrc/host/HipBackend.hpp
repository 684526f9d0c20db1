///
/// @file  HipBackend.hpp
/// @brief gfx950 implementation of GpuBackend.  Replaces the reference's
///        CudaBackend (src/gpu/cuda/CudaBackend.cpp:222-491): device set-up,
///        code-object load, sieving-prime cache, the launch loop and the
///        reduction -- here the counts are reduced on the device and one small
///        copy ends a call.
///
///        One execution engine serves both the reference's single-interval
///        count and the batch API: intervals are cut into runs of segments,
///        runs are packed into slabs, and each slab is a big-prime bucket
///        build plus one persistent, work-stealing sieve launch.
///
#ifndef PSG_HIP_BACKEND_HPP
#define PSG_HIP_BACKEND_HPP

#include "GpuBackend.hpp"
#include "hip_api.hpp"
#include "tables.hpp"
#include "../kernel_abi.h"

#include <primesieve_gpu.h>

#include <stdint.h>
#include <functional>
#include <string>
#include <utility>
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
  void swap(DeviceBuffer& o)
  {
    std::swap(api_, o.api_);
    std::swap(ptr_, o.ptr_);
    std::swap(bytes_, o.bytes_);
  }

private:
  const HipApi* api_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

/// Pinned host memory (hipHostMalloc), grown on demand; contents are not kept
/// when it grows.  Copies from and to it are asynchronous DMA, while a copy
/// from pageable memory is staged by the runtime before hipMemcpyAsync
/// returns (65 us for a call's descriptors).
class PinnedBuffer
{
public:
  PinnedBuffer() {}
  ~PinnedBuffer() { release(); }
  PinnedBuffer(const PinnedBuffer&) = delete;
  PinnedBuffer& operator=(const PinnedBuffer&) = delete;

  void* ensure(const HipApi* api, size_t bytes);
  void release();

private:
  const HipApi* api_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

/// An interval [lo, hi] whose six counts are added to output slot `out`.
struct Interval
{
  uint64_t lo;
  uint64_t hi;
  uint32_t out;
};

/// Knobs of the HIP backend beyond the reference's GpuConfig (which keeps
/// exactly the reference's four members); 0 means auto.  A backend reads
/// them when it is created (createHipBackend), so a change applies to the
/// next backend, like a GpuConfig change.
struct HipTuning
{
  uint32_t slabSegments = 0;   // segments per big-prime bucket slab (auto: 4 GiB of sieve bytes)
  uint32_t bucketFactor = 0;   // primes > bucketFactor * sieve bytes use the bucket tier;
                               // auto = 7, 0xffffffff = never
  uint32_t bucketMargin = 0;   // bucket list room beyond the expected count, in
                               // standard deviations; auto = 6
};

/// The process-wide HipTuning that new backends take (thread-safe).
void setHipTuning(const HipTuning& t);
HipTuning hipTuning();

/// How the engine cuts work (all in segments).
struct ExecOptions
{
  uint32_t runSegs = 1024;        // longest run
  uint32_t slabSegs = 1024;       // segments per slab (one persistent launch)
  uint32_t windowSegs = 64;       // segments per bucket window
  uint32_t primesPerChunk = 1024; // big primes per fill workgroup ...
  uint32_t maxChunks = 256;       // ... but at most maxChunks workgroups per window
  double entryBudget = 3.0e9;     // expected bucket entries per slab (12 GB)
  bool pipeline = false;          // fill slab k+1 on a second stream while slab k is sieved
  bool graph = false;             // replay a captured launch series (captured once per shape)
  bool noAlloc = false;           // per-call device memory from psg_scratch only
  uint32_t bucketFactor = 0;      // bucket threshold in segments for this call (0: the backend's)
  uint64_t listOut = 0;           // list mode: uint32_t* for the primes (one segment per interval)
  uint64_t listOffsets = 0;       // list mode: const uint32_t[outs], first output index per interval
};

class HipBackend : public primesieve::gpu::GpuBackend
{
public:
  HipBackend(const HipApi* api, const primesieve::gpu::DeviceInfo& dev,
             const primesieve::gpu::GpuConfig& config, const HipTuning& tuning);
  ~HipBackend() override;

  const primesieve::gpu::DeviceInfo& device() const override { return dev_; }
  uint64_t count(uint64_t lo, uint64_t hi, int countType) override;
  uint32_t sieveBytes() const override { return sieveBytes_; }
  uint32_t workGroupSize() const override { return wgSize_; }

  /// All six counts of [lo, hi] (lo >= 7 after clamping) in one pass.
  void countAll(uint64_t lo, uint64_t hi, uint64_t out[PSG_NCOUNTS]);

  /// Six counts per output slot for a batch of intervals (lo >= 7 after
  /// clamping), all in one launch series; with `graph` the series is a
  /// captured one, replayed for every batch of the same shape.
  void countBatch(const std::vector<Interval>& iv, uint32_t outs, uint64_t* out6, bool graph = false);

  /// countAll without any device allocation: the call's descriptors,
  /// counters and bucket lists live in the code object's psg_scratch array
  /// (PSG_SCRATCH_BYTES), and the sieving primes must already be cached.
  /// Throws when the interval needs more scratch than there is.
  void countAllNoAlloc(uint64_t lo, uint64_t hi, uint64_t out[PSG_NCOUNTS]);

  const psg_stats& lastStats() const { return stats_; }
  /// Make the device cache hold every sieving prime up to `limit` and copy
  /// the primes <= limit (up to cap; ascending) and their division magics
  /// (read back from the device) to the caller's arrays (either may be null).
  /// Returns the number of cached primes <= limit.
  uint64_t sievingPrimes(uint32_t limit, uint32_t* primes, uint64_t* magics, uint64_t cap);

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
    t.bucket_margin = bucketMargin_;
    return t;
  }

private:
  void init();
  void ensureSievingPrimes(uint32_t limit);
  void execute(const std::vector<Interval>& iv, uint32_t outs, uint64_t* out6,
               const ExecOptions& opt);
  bool launchPlan(uint32_t outs, const std::vector<uint32_t>& room, const ExecOptions& opt);
  void installHostPrimes(uint32_t want);
  void generatePrimesOnDevice(uint32_t want);
  uint32_t primesUpTo(uint64_t x) const;
  uint32_t persistentGrid(uint32_t nSegs);
  uint32_t fillGrid(uint32_t nItems);
  hipGraphExec_t graphFor(const std::vector<uint64_t>& key, bool& captured,
                          const std::function<void()>& record);
  // LDS of the segment kernel: the segment, its tail and (when they fit) the
  // small pre-sieve tiles
  // The small pre-sieve tiles go into LDS when they fit beside the segment and
  // do not lower the number of workgroups a CU holds (a 64 KiB segment of 8
  // waves runs two workgroups per CU only with its tiles in global memory).
  bool tilesFitLds() const;
  uint32_t sieveLds() const
  {
    return sieveBytes_ + PSG_SIEVE_LDS_TAIL
         + ((config_.preSieve && tilesFitLds()) ? 4u * PSG_SMALL_TILE_WORDS : 0u);
  }
  hipEvent_t phaseEvent(size_t i);

  const HipApi* api_;
  primesieve::gpu::DeviceInfo dev_;
  primesieve::gpu::GpuConfig config_;
  HipTuning hipTuning_;
  uint32_t sieveBytes_ = 0;
  uint32_t ldsBytes_ = 0;         // LDS per workgroup the device offers (capped)
  uint32_t wgSize_ = 0;
  uint32_t segmentsPerLaunch_ = 0;
  uint32_t bucketFactor_ = 8;     // primes > bucketFactor_ * S use the bucket tier (0 = never)
  bool bucketAuto_ = true;        // bucketFactor_ is the default, not set by a tuning
  uint32_t slabSegments_ = 0;     // segments per slab of a single interval
  uint32_t batchSlabSegs_ = 32768; // segments per slab of a batch (8192: 9601, 16384: 9720,
                                   // 32768: 9757 intervals/s on the 8192-interval batch)
  uint32_t windowSegs_ = 128;     // segments per bucket window of a single interval
  uint32_t persistentWGs_ = 0;    // resident workgroups of the sieve kernel
  mutable int tilesLds_ = -1;     // tilesFitLds(), once the sieve kernel is loaded
  uint32_t fillWGs_ = 0;          // resident workgroups of the bucket fill
  uint32_t bucketMargin_ = 6;     // bucket list room: expected + margin * sqrt(expected) + 64

  hipModule_t module_ = nullptr;
  hipFunction_t fnSieve_ = nullptr;
  hipFunction_t fnBucketFill_ = nullptr;
  hipFunction_t fnMagics_ = nullptr;
  hipFunction_t fnClear_ = nullptr;
  static constexpr uint32_t kClearWGs = 256;   // grid of psg_clear (grid-stride)
  uint64_t scratchAddr_ = 0;      // device address of psg_scratch on this device
  uint64_t scratchBytes_ = 0;
  uint64_t listCountAddr_ = 0;    // bucket list counters of the current call

  // captured launch series, keyed by shape (descriptor address, slab count,
  // grids, which slabs fill buckets)
  struct CachedGraph
  {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
  };
  static constexpr size_t kMaxGraphs = 16;
  std::vector<CachedGraph> graphs_;
  hipStream_t stream_ = nullptr;
  hipStream_t fillStream_ = nullptr;   // bucket fills of a pipelined call
  hipEvent_t evSetup_ = nullptr;
  hipEvent_t evStart_ = nullptr;
  hipEvent_t evStop_ = nullptr;

  DeviceBuffer primes_;        // uint32 sieving primes >= first cross prime
  DeviceBuffer magics_;        // uint64 division magics
  DeviceBuffer tiles_;        // pre-sieve tiles
  DeviceBuffer counts_;        // uint64 [outputs][6]
  DeviceBuffer descriptors_;   // runs and windows of every slab of a call
  DeviceBuffer workCounters_;  // uint32 per slab
  DeviceBuffer listCount_;     // uint32 per segment of the call: bucket entries appended
  DeviceBuffer entries_[2];    // bucket entries of a slab; two alternate when pipelined
  DeviceBuffer flags_;         // uint32 overflow flag, pad, uint64 entries appended
  uint64_t* hostFlags_ = nullptr;      // pinned copy of flags_
  std::vector<uint64_t> hostCounts_;
  std::vector<unsigned char> staging_;  // descriptor blob of calls that must not allocate
  PinnedBuffer pinnedDesc_;            // descriptor blob of the other calls
  PinnedBuffer pinnedCounts_;          // their counts, copied back
  std::vector<hipEvent_t> phaseEvents_;

  // the plan of the current call (built by execute, launched by launchPlan)
  struct Slab
  {
    size_t run0, run1, win0, win1;
    uint32_t nSegs, nWGs, segBase;
    double est;
  };
  std::vector<PsgRun> runs_;
  std::vector<PsgWindow> wins_;
  std::vector<Slab> slabs_;
  std::vector<double> segExpect_;    // expected bucket entries per segment (0: none)
  uint32_t bigBegin_ = 0;

  std::vector<uint32_t> hostPrimes_; // the primes held in primes_
  std::vector<double> invPrefix_;    // invPrefix_[i] = sum_{j<i} 1/p_j
  uint32_t primesLimit_ = 0;         // every prime <= primesLimit_ is held
  uint32_t firstCrossPrime_ = 29;

  psg_stats stats_{};
};

} // namespace psg

#endif
