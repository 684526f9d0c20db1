///
/// @file  HipBackend.cpp
/// @brief The gfx950 backend: device enumeration, code-object load, the
///        sieving-prime cache and the segment launch loop.
///
///        Reference counterparts: listCudaDevices / createCudaBackend
///        (src/gpu/cuda/CudaBackend.cpp:498-552), CudaBackend::init (:222-344),
///        ensureSievingPrimes (:346-388), runLaunch (:401-441) and count
///        (:443-491).
///

#include "HipBackend.hpp"

#include <primesieve/primesieve_error.hpp>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

// gfx950 code object of kernels/psg_sieve.hip, linked in by embed.S
extern "C" const unsigned char psg_sieve_hsaco_start[];
extern "C" const unsigned char psg_sieve_hsaco_end[];

using primesieve::primesieve_error;
using primesieve::gpu::DeviceInfo;
using primesieve::gpu::GpuBackend;
using primesieve::gpu::GpuConfig;

namespace psg {

void DeviceBuffer::allocate(const HipApi* api, size_t bytes)
{
  if (ptr_ && bytes_ >= bytes)
    return;
  release();
  api_ = api;
  void* p = nullptr;
  hipCheck(api, api->Malloc(&p, bytes ? bytes : 1), "hipMalloc");
  ptr_ = p;
  bytes_ = bytes;
}

void DeviceBuffer::release()
{
  if (ptr_ && api_)
    (void) api_->Free(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

HipBackend::HipBackend(const HipApi* api, const DeviceInfo& dev, const GpuConfig& config)
  : api_(api), dev_(dev), config_(config)
{
  init();
}

HipBackend::~HipBackend()
{
  if (!api_)
    return;
  (void) api_->SetDevice(dev_.deviceIndex);
  primes_.release();
  magics_.release();
  tile0_.release();
  tile1_.release();
  counts_.release();
  bucketCounts_.release();
  listOff_.release();
  entries_.release();
  overflow_.release();
  if (hostCounts_) (void) api_->HostFree(hostCounts_);
  if (hostOverflow_) (void) api_->HostFree(hostOverflow_);
  for (hipEvent_t e : phaseEvents_) (void) api_->EventDestroy(e);
  if (evStart_) (void) api_->EventDestroy(evStart_);
  if (evStop_) (void) api_->EventDestroy(evStop_);
  if (stream_) (void) api_->StreamDestroy(stream_);
  if (module_) (void) api_->ModuleUnload(module_);
}

void HipBackend::init()
{
  hipCheck(api_, api_->SetDevice(dev_.deviceIndex), "hipSetDevice");

  // geometry: LDS segment, workgroup, launch size
  const uint64_t ldsMax = std::min<uint64_t>(dev_.localMemBytes, PSG_MAX_SEG_BYTES + 1024);
  uint32_t sb = config_.sieveBytes ? config_.sieveBytes : 128u * 1024u;
  sb = (sb / 256u) * 256u;
  if (sb < 256u) sb = 256u;
  if (sb > PSG_MAX_SEG_BYTES) sb = PSG_MAX_SEG_BYTES;
  while (sb > 256u && (uint64_t) sb + 1024u > ldsMax) sb -= 256u;
  sieveBytes_ = sb;

  uint32_t wg = config_.workGroupSize ? config_.workGroupSize : 1024u;
  wg = (wg / 64u) * 64u;
  if (wg < 64u) wg = 64u;
  if (wg > 1024u) wg = 1024u;
  if (dev_.maxWorkGroupSize && wg > dev_.maxWorkGroupSize)
    wg = (uint32_t) (dev_.maxWorkGroupSize / 64u) * 64u;
  wgSize_ = wg;

  segmentsPerLaunch_ = config_.segmentsPerLaunch ? config_.segmentsPerLaunch : 65536u;
  bucketFactor_ = (config_.bucketFactor == 0) ? 8u
                : (config_.bucketFactor == 0xffffffffu) ? 0u : config_.bucketFactor;
  // a slab spans ~128 MiB of sieve bytes, at most 2048 segments (the count
  // and fill kernels keep one LDS counter per (segment, class): 64 KiB)
  uint32_t slab = config_.slabSegments ? config_.slabSegments
                                       : (uint32_t) ((128ull << 20) / sieveBytes_);
  slabSegments_ = std::max<uint32_t>(1, std::min<uint32_t>(slab, 2048));
  bucketChunks_ = std::max<uint32_t>(64, 4 * (dev_.computeUnits ? dev_.computeUnits : 256));
  windowSegs_ = 64;
  firstCrossPrime_ = config_.preSieve ? PSG_FIRST_CROSS_PRIME : 7u;

  const size_t coBytes = (size_t) (psg_sieve_hsaco_end - psg_sieve_hsaco_start);
  hipCheck(api_, api_->ModuleLoadData(&module_, psg_sieve_hsaco_start),
           ("loading the gfx950 code object (" + std::to_string(coBytes) + " bytes)").c_str());
  const struct { hipFunction_t* f; const char* name; } fns[] = {
    { &fnSieve_, "psg_sieve_count" }, { &fnBucketCount_, "psg_bucket_count" },
    { &fnScanChunks_, "psg_bucket_scan_chunks" }, { &fnScanLists_, "psg_bucket_scan_lists" },
    { &fnBucketFill_, "psg_bucket_fill" } };
  for (const auto& f : fns)
    hipCheck(api_, api_->ModuleGetFunction(f.f, module_, f.name),
             (std::string("hipModuleGetFunction(") + f.name + ")").c_str());
  hipCheck(api_, api_->StreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  hipCheck(api_, api_->EventCreate(&evStart_), "hipEventCreate");
  hipCheck(api_, api_->EventCreate(&evStop_), "hipEventCreate");

  if (config_.preSieve)
  {
    const std::vector<uint32_t> t0 = tileWindows({ 7, 11, 13, 17 });
    const std::vector<uint32_t> t1 = tileWindows({ 19, 23 });
    if (t0.size() != PSG_TILE0_PERIOD || t1.size() != PSG_TILE1_PERIOD)
      throw primesieve_error("pre-sieve tile period mismatch");
    tile0_.allocate(api_, t0.size() * 4);
    tile1_.allocate(api_, t1.size() * 4);
    hipCheck(api_, api_->Memcpy(tile0_.get(), t0.data(), t0.size() * 4, hipMemcpyHostToDevice), "hipMemcpy tile0");
    hipCheck(api_, api_->Memcpy(tile1_.get(), t1.data(), t1.size() * 4, hipMemcpyHostToDevice), "hipMemcpy tile1");
  }
  counts_.allocate(api_, PSG_NCOUNTS * sizeof(uint64_t));
  void* hc = nullptr;
  hipCheck(api_, api_->HostMalloc(&hc, PSG_NCOUNTS * sizeof(uint64_t), 0), "hipHostMalloc");
  hostCounts_ = (uint64_t*) hc;
  hc = nullptr;
  hipCheck(api_, api_->HostMalloc(&hc, 4, 0), "hipHostMalloc");
  hostOverflow_ = (uint32_t*) hc;
  *hostOverflow_ = 0;
}

void HipBackend::ensureSievingPrimes(uint32_t limit)
{
  if (limit <= primesLimit_ && !hostPrimes_.empty())
    return;
  // grow geometrically so that a slowly rising stop does not regenerate
  uint64_t want = std::max<uint64_t>(limit, (uint64_t) primesLimit_ * 2);
  want = std::max<uint64_t>(want, 1u << 16);
  want = std::min<uint64_t>(want, 0xffffffffull);
  std::vector<uint32_t> pr = primesBetween(firstCrossPrime_, (uint32_t) want);
  std::vector<uint64_t> mg(pr.size());
  for (size_t i = 0; i < pr.size(); i++)
    mg[i] = divisionMagic(pr[i]);
  primes_.allocate(api_, pr.size() * 4 + 4);
  magics_.allocate(api_, mg.size() * 8 + 8);
  if (!pr.empty())
  {
    hipCheck(api_, api_->Memcpy(primes_.get(), pr.data(), pr.size() * 4, hipMemcpyHostToDevice), "hipMemcpy primes");
    hipCheck(api_, api_->Memcpy(magics_.get(), mg.data(), mg.size() * 8, hipMemcpyHostToDevice), "hipMemcpy magics");
  }
  invPrefix_.assign(pr.size() + 1, 0.0);
  for (size_t i = 0; i < pr.size(); i++)
    invPrefix_[i + 1] = invPrefix_[i] + 1.0 / pr[i];
  hostPrimes_.swap(pr);
  primesLimit_ = (uint32_t) want;
}

void HipBackend::launchSieve(const PsgSieveArgs& args, uint32_t nseg)
{
  PsgSieveArgs a = args;
  void* params[] = { &a };
  hipCheck(api_, api_->ModuleLaunchKernel(fnSieve_, nseg, 1, 1, wgSize_, 1, 1,
                                          sieveBytes_, stream_, params, nullptr),
           "hipModuleLaunchKernel(psg_sieve_count)");
}

void HipBackend::countAll(uint64_t lo, uint64_t hi, uint64_t out[PSG_NCOUNTS])
{
  const auto t0 = std::chrono::steady_clock::now();
  stats_ = psg_stats{};
  stats_.sieve_bytes = sieveBytes_;
  stats_.work_group_size = wgSize_;
  std::fill(out, out + PSG_NCOUNTS, 0);
  if (lo < 7) lo = 7;
  if (hi < 7 || lo > hi)
    return;

  hipCheck(api_, api_->SetDevice(dev_.deviceIndex), "hipSetDevice");
  const ByteRange br = byteRange(lo, hi);
  const uint32_t sq = (uint32_t) isqrt64(hi);
  ensureSievingPrimes(sq);

  const uint32_t S = sieveBytes_;
  const auto upto = [&](uint64_t x) {
    return (uint32_t) (std::upper_bound(hostPrimes_.begin(), hostPrimes_.end(),
                                        (uint32_t) std::min<uint64_t>(x, 0xffffffffull)) -
                       hostPrimes_.begin());
  };
  const uint32_t nprimes = upto(sq);

  PsgSieveArgs args;
  std::memset(&args, 0, sizeof args);
  args.range.gLo = br.gLo;
  args.range.gHi = br.gHi;
  args.range.loMask = br.loMask;
  args.range.hiMask = br.hiMask;
  args.range.nprimes = nprimes;
  args.primes = primes_.addr();
  args.magics = magics_.addr();
  args.tile0 = tile0_.addr();
  args.tile1 = tile1_.addr();
  args.counts = counts_.addr();
  args.segBytes = S;
  args.tier0End = std::min(upto(std::min<uint32_t>(256, S / 8)), nprimes);
  args.tier1End = std::min(upto(S / 8), nprimes);
  args.tier2End = std::min(upto(S), nprimes);
  args.tier3End = nprimes;

  const uint64_t nseg = (br.gHi - br.gLo) / S + 1;

  // big primes (p > bucket threshold) go through the bucket tier when the
  // segment size is a power of two; otherwise tier 3 handles every p > S
  const bool pow2 = (S & (S - 1)) == 0;
  const uint64_t threshold = (uint64_t) S * bucketFactor_;
  const uint32_t bigBegin = (pow2 && bucketFactor_) ? std::min(upto(threshold), nprimes) : nprimes;
  const bool useBuckets = bigBegin < nprimes;
  args.tier3End = bigBegin;

  hipCheck(api_, api_->MemsetAsync(counts_.get(), 0, PSG_NCOUNTS * sizeof(uint64_t), stream_),
           "hipMemsetAsync");
  hipCheck(api_, api_->EventRecord(evStart_, stream_), "hipEventRecord");
  uint32_t launches = 0;
  size_t nev = 0;
  if (!useBuckets)
  {
    for (uint64_t s = 0; s < nseg; s += segmentsPerLaunch_)
    {
      const uint32_t n = (uint32_t) std::min<uint64_t>(segmentsPerLaunch_, nseg - s);
      args.range.B0 = br.gLo + s * S;
      recordPhase(nev++);
      launchSieve(args, n);
      launches++;
    }
    recordPhase(nev++);
  }
  else
  {
    uint32_t shift = 0;
    while ((1u << shift) < S) shift++;
    const uint32_t L = slabSegments_;
    const uint32_t nBig = nprimes - bigBegin;
    const uint32_t entryBytes = (S <= 65536u) ? 2u : 4u;
    // Each class of each prime hits a slab of B bytes at most B/p + 1 times,
    // so this capacity is a strict upper bound (the overflow flag is a guard).
    const double invSum = invPrefix_[nprimes] - invPrefix_[bigBegin];
    const uint64_t cap64 = (uint64_t) (8.0 * (double) L * S * invSum) + 8ull * nBig + 4096;
    if (cap64 > 0x7fffffffull)
      throw primesieve_error("bucket slab too large; lower the slab size");
    const uint32_t capacity = (uint32_t) cap64;
    // workgroups = windows x prime chunks, about bucketChunks_ of them
    const uint32_t R = std::min(windowSegs_, L);
    const uint32_t nWindowsMax = (L + R - 1) / R;
    const uint32_t nChunks = std::max<uint32_t>(1, std::min<uint32_t>(
        (bucketChunks_ + nWindowsMax - 1) / nWindowsMax, (nBig + 63) / 64));
    const uint32_t nListsMax = 8u * L;
    bucketCounts_.allocate(api_, (size_t) nChunks * nListsMax * 4);
    listOff_.allocate(api_, ((size_t) nListsMax + 1) * 4);
    entries_.allocate(api_, (size_t) capacity * entryBytes);
    overflow_.allocate(api_, 4);
    hipCheck(api_, api_->MemsetAsync(overflow_.get(), 0, 4, stream_), "hipMemsetAsync");

    PsgBucketArgs b;
    std::memset(&b, 0, sizeof b);
    b.primes = primes_.addr();
    b.magics = magics_.addr();
    b.counts = bucketCounts_.addr();
    b.listOff = listOff_.addr();
    b.entries = entries_.addr();
    b.overflow = overflow_.addr();
    b.segShift = shift;
    b.bigBegin = bigBegin;
    b.bigEnd = nprimes;
    b.nChunks = nChunks;
    b.capacity = capacity;
    b.entryBytes = entryBytes;
    b.windowSegs = R;

    args.buckets = entries_.addr();
    args.bucketOffs = listOff_.addr();
    args.entryBytes = entryBytes;
    args.capacity = capacity;
    args.segBegin = 0;
    args.slabSegBegin = 0;

    for (uint64_t s = 0; s < nseg; s += L)
    {
      const uint32_t n = (uint32_t) std::min<uint64_t>(L, nseg - s);
      b.slabB = br.gLo + s * S;
      b.slabBytes = n * S;
      b.nLists = 8u * n;
      b.nWindows = (n + R - 1) / R;
      const uint32_t grid = b.nWindows * nChunks;
      const uint32_t ldsBytes = 8u * R * 4u;
      void* pb[] = { &b };
      recordPhase(nev++);
      hipCheck(api_, api_->ModuleLaunchKernel(fnBucketCount_, grid, 1, 1, 512, 1, 1,
                                              ldsBytes, stream_, pb, nullptr),
               "hipModuleLaunchKernel(psg_bucket_count)");
      hipCheck(api_, api_->ModuleLaunchKernel(fnScanChunks_, (b.nLists + 255) / 256, 1, 1, 256, 1, 1,
                                              0, stream_, pb, nullptr),
               "hipModuleLaunchKernel(psg_bucket_scan_chunks)");
      hipCheck(api_, api_->ModuleLaunchKernel(fnScanLists_, 1, 1, 1, 1024, 1, 1, 0, stream_, pb, nullptr),
               "hipModuleLaunchKernel(psg_bucket_scan_lists)");
      hipCheck(api_, api_->ModuleLaunchKernel(fnBucketFill_, grid, 1, 1, 512, 1, 1,
                                              ldsBytes, stream_, pb, nullptr),
               "hipModuleLaunchKernel(psg_bucket_fill)");
      recordPhase(nev++);
      args.range.B0 = b.slabB;
      launchSieve(args, n);
      launches++;
    }
    recordPhase(nev++);
    hipCheck(api_, api_->MemcpyAsync(hostOverflow_, overflow_.get(), 4, hipMemcpyDeviceToHost, stream_),
             "hipMemcpyAsync overflow");
    stats_.bucket_entries = 0;
  }
  hipCheck(api_, api_->EventRecord(evStop_, stream_), "hipEventRecord");
  hipCheck(api_, api_->MemcpyAsync(hostCounts_, counts_.get(), PSG_NCOUNTS * sizeof(uint64_t),
                                   hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync counts");
  hipCheck(api_, api_->StreamSynchronize(stream_), "sieve kernel");
  if (useBuckets && *hostOverflow_)
    throw primesieve_error("big-prime bucket capacity exceeded");
  float ms = 0.f;
  hipCheck(api_, api_->EventElapsedTime(&ms, evStart_, evStop_), "hipEventElapsedTime");

  // phase timings: with buckets, events alternate fill | sieve | fill | ...
  double fillMs = 0, sieveMs = 0;
  for (size_t e = 0; e + 1 < nev; e++)
  {
    float t = 0.f;
    hipCheck(api_, api_->EventElapsedTime(&t, phaseEvents_[e], phaseEvents_[e + 1]), "hipEventElapsedTime");
    if (useBuckets && e % 2 == 0) fillMs += t;
    else sieveMs += t;
  }

  std::copy(hostCounts_, hostCounts_ + PSG_NCOUNTS, out);
  stats_.kernel_ms = sieveMs;
  stats_.fill_ms = fillMs;
  stats_.segments = nseg;
  stats_.launches = launches;
  stats_.sieving_primes = nprimes;
  stats_.device_ms = ms;
  stats_.wall_ms = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - t0).count();
}

void HipBackend::recordPhase(size_t i)
{
  while (phaseEvents_.size() <= i)
  {
    hipEvent_t e = nullptr;
    hipCheck(api_, api_->EventCreate(&e), "hipEventCreate");
    phaseEvents_.push_back(e);
  }
  hipCheck(api_, api_->EventRecord(phaseEvents_[i], stream_), "hipEventRecord");
}

uint64_t HipBackend::count(uint64_t lo, uint64_t hi, int countType)
{
  if (countType < 0 || countType >= PSG_NCOUNTS)
    throw primesieve_error("invalid count type");
  uint64_t c[PSG_NCOUNTS];
  countAll(lo, hi, c);
  return c[countType];
}

} // namespace psg

// ---------------------------------------------------------------------------
// Enumeration and creation (reference src/gpu/cuda/CudaBackend.cpp:498-552)
// ---------------------------------------------------------------------------

namespace primesieve {
namespace gpu {

void listHipDevices(std::vector<DeviceInfo>& out)
{
  std::string err;
  const psg::HipApi* api = psg::hipApi(err);
  if (!api)
    return;
  int n = 0;
  if (api->GetDeviceCount(&n) != hipSuccess || n <= 0)
    return;
  int drv = 0, rt = 0;
  (void) api->DriverGetVersion(&drv);
  (void) api->RuntimeGetVersion(&rt);
  for (int i = 0; i < n; i++)
  {
    DeviceInfo d;
    d.backend = "HIP";
    d.vendor = "AMD";
    char name[256] = { 0 };
    if (api->DeviceGetName(name, sizeof name - 1, i) == hipSuccess)
      d.name = name;
    char ver[96];
    std::snprintf(ver, sizeof ver, "HIP runtime %d, driver %d", rt, drv);
    d.version = ver;
    d.deviceIndex = i;
    int v = 0;
    if (api->DeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, i) == hipSuccess)
      d.computeUnits = (uint32_t) v;
    if (api->DeviceGetAttribute(&v, hipDeviceAttributeClockRate, i) == hipSuccess)
      d.clockMHz = (uint32_t) (v / 1000);
    if (api->DeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, i) == hipSuccess)
      d.localMemBytes = (uint64_t) v;
    if (api->DeviceGetAttribute(&v, hipDeviceAttributeMaxThreadsPerBlock, i) == hipSuccess)
      d.maxWorkGroupSize = (uint64_t) v;
    size_t mem = 0;
    if (api->DeviceTotalMem(&mem, i) == hipSuccess)
      d.globalMemBytes = d.maxAllocBytes = mem;
    out.push_back(d);
  }
}

std::unique_ptr<GpuBackend> createHipBackend(const DeviceInfo& dev,
                                             const GpuConfig& config,
                                             std::string& error)
{
  std::string err;
  const psg::HipApi* api = psg::hipApi(err);
  if (!api)
  {
    error = err;
    return std::unique_ptr<GpuBackend>();
  }
  try
  {
    return std::unique_ptr<GpuBackend>(new psg::HipBackend(api, dev, config));
  }
  catch (const std::exception& e)
  {
    error = e.what();
    return std::unique_ptr<GpuBackend>();
  }
}

} // namespace gpu
} // namespace primesieve
