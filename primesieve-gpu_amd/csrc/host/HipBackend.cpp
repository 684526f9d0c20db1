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
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
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

void* PinnedBuffer::ensure(const HipApi* api, size_t bytes)
{
  if (ptr_ && bytes_ >= bytes)
    return ptr_;
  const size_t want = std::max(bytes, 2 * bytes_);   // grow geometrically
  release();
  api_ = api;
  void* p = nullptr;
  hipCheck(api, api->HostMalloc(&p, want ? want : 1, 0), "hipHostMalloc");
  ptr_ = p;
  bytes_ = want;
  return ptr_;
}

void PinnedBuffer::release()
{
  if (ptr_ && api_)
    (void) api_->HostFree(ptr_);
  ptr_ = nullptr;
  bytes_ = 0;
}

namespace {

std::mutex g_tuningMutex;
HipTuning g_tuning;

} // namespace

void setHipTuning(const HipTuning& t)
{
  std::lock_guard<std::mutex> lock(g_tuningMutex);
  g_tuning = t;
}

HipTuning hipTuning()
{
  std::lock_guard<std::mutex> lock(g_tuningMutex);
  return g_tuning;
}

HipBackend::HipBackend(const HipApi* api, const DeviceInfo& dev, const GpuConfig& config,
                       const HipTuning& tuning)
  : api_(api), dev_(dev), config_(config), hipTuning_(tuning)
{
  init();
}

HipBackend::~HipBackend()
{
  if (!api_)
    return;
  (void) api_->SetDevice(dev_.deviceIndex);
  for (const CachedGraph& g : graphs_)
    (void) api_->GraphExecDestroy(g.exec);
  primes_.release();
  magics_.release();
  tiles_.release();
  counts_.release();
  descriptors_.release();
  workCounters_.release();
  listCount_.release();
  entries_[0].release();
  entries_[1].release();
  flags_.release();
  if (hostFlags_) (void) api_->HostFree(hostFlags_);
  pinnedDesc_.release();
  pinnedCounts_.release();
  for (hipEvent_t e : phaseEvents_) (void) api_->EventDestroy(e);
  if (evSetup_) (void) api_->EventDestroy(evSetup_);
  if (evStart_) (void) api_->EventDestroy(evStart_);
  if (evStop_) (void) api_->EventDestroy(evStop_);
  if (stream_) (void) api_->StreamDestroy(stream_);
  if (fillStream_) (void) api_->StreamDestroy(fillStream_);
  if (module_) (void) api_->ModuleUnload(module_);
}

void HipBackend::init()
{
  hipCheck(api_, api_->SetDevice(dev_.deviceIndex), "hipSetDevice");

  // geometry: LDS segment, workgroup, launch size
  const uint64_t ldsMax = std::min<uint64_t>(dev_.localMemBytes, PSG_MAX_SEG_BYTES + PSG_SIEVE_LDS_TAIL);
  ldsBytes_ = (uint32_t) ldsMax;
  uint32_t sb = config_.sieveBytes ? config_.sieveBytes : 128u * 1024u;
  sb = (sb / 256u) * 256u;
  if (sb < 256u) sb = 256u;
  if (sb > PSG_MAX_SEG_BYTES) sb = PSG_MAX_SEG_BYTES;
  while (sb > 256u && (uint64_t) sb + PSG_SIEVE_LDS_TAIL > ldsMax) sb -= 256u;  // the segment and its tail
  sieveBytes_ = sb;

  // A power of two, as the reference requires of workGroupSize
  // (src/gpu/GpuBackend.hpp:49-50, rounded down in CudaBackend.cpp:260-263):
  // tier 0 splits a segment into equal per-wave slices of S / waves bytes,
  // which must be whole 16-byte multiples for every S that is a multiple of
  // 256, i.e. waves must divide 16.
  uint32_t wg = config_.workGroupSize ? config_.workGroupSize : 1024u;
  if (dev_.maxWorkGroupSize && wg > dev_.maxWorkGroupSize)
    wg = (uint32_t) std::min<uint64_t>(dev_.maxWorkGroupSize, 0xffffffffu);
  wg = std::max<uint32_t>(64u, std::min<uint32_t>(wg, 1024u));
  while (wg & (wg - 1u))
    wg &= wg - 1u;
  wgSize_ = wg;

  // segments per persistent launch (one slab); the reference's default is
  // 256 per compute unit (src/gpu/cuda/CudaBackend.cpp:315-319)
  segmentsPerLaunch_ = config_.segmentsPerLaunch
                     ? config_.segmentsPerLaunch
                     : std::max<uint32_t>(dev_.computeUnits, 1u) * 256u;
  bucketAuto_ = hipTuning_.bucketFactor == 0;
  bucketFactor_ = bucketAuto_ ? 7u
                : (hipTuning_.bucketFactor == 0xffffffffu) ? 0u : hipTuning_.bucketFactor;
  // a slab spans up to 4 GiB of sieve bytes (one launch of each kernel
  // for a 1e11 interval: no launch tails between slabs; its bucket entries,
  // 4 B per crossing of a big prime and at most the entry budget, fit in HBM)
  uint32_t slab = hipTuning_.slabSegments ? hipTuning_.slabSegments
                                          : (uint32_t) ((1ull << 32) / sieveBytes_);
  slabSegments_ = std::max<uint32_t>(1, std::min<uint32_t>(slab, 32768));
  windowSegs_ = 128;
  bucketMargin_ = hipTuning_.bucketMargin ? hipTuning_.bucketMargin : 6u;
  firstCrossPrime_ = config_.preSieve ? PSG_FIRST_CROSS_PRIME : 7u;

  const size_t coBytes = (size_t) (psg_sieve_hsaco_end - psg_sieve_hsaco_start);
  hipCheck(api_, api_->ModuleLoadData(&module_, psg_sieve_hsaco_start),
           ("loading the gfx950 code object (" + std::to_string(coBytes) + " bytes)").c_str());
  const struct { hipFunction_t* f; const char* name; } fns[] = {
    { &fnSieve_, "psg_sieve_count" }, { &fnBucketFill_, "psg_bucket_fill" },
    { &fnMagics_, "psg_prime_magics" }, { &fnClear_, "psg_clear" } };
  for (const auto& f : fns)
    hipCheck(api_, api_->ModuleGetFunction(f.f, module_, f.name),
             (std::string("hipModuleGetFunction(") + f.name + ")").c_str());
  // The code object's static scratch array (the device memory of the
  // no-allocation entry point): its address on this device, resolved once
  // here with the device current.  Host code uses only this address.
  {
    hipDeviceptr_t p = nullptr;
    size_t bytes = 0;
    hipCheck(api_, api_->ModuleGetGlobal(&p, &bytes, module_, "psg_scratch"),
             "hipModuleGetGlobal(psg_scratch)");
    if (!p || bytes < PSG_SCRATCH_BYTES)
      throw primesieve_error("psg_scratch: unexpected size " + std::to_string(bytes));
    scratchAddr_ = (uint64_t) (uintptr_t) p;
    scratchBytes_ = PSG_SCRATCH_BYTES;
  }
  // the segment kernel addresses its sieve words by LDS offset, which holds
  // only while its dynamic LDS starts at 0, i.e. it has no static LDS
  int staticLds = -1;
  hipCheck(api_, api_->FuncGetAttribute(&staticLds, HIP_FUNC_ATTRIBUTE_SHARED_SIZE_BYTES, fnSieve_),
           "hipFuncGetAttribute(psg_sieve_count)");
  if (staticLds != 0)
    throw primesieve_error("psg_sieve_count has " + std::to_string(staticLds) +
                                       " bytes of static LDS; its sieve must start at LDS address 0");
  hipCheck(api_, api_->StreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  hipCheck(api_, api_->StreamCreateWithFlags(&fillStream_, hipStreamNonBlocking), "hipStreamCreate");
  hipCheck(api_, api_->EventCreate(&evSetup_), "hipEventCreate");
  hipCheck(api_, api_->EventCreate(&evStart_), "hipEventCreate");
  hipCheck(api_, api_->EventCreate(&evStop_), "hipEventCreate");

  if (config_.preSieve)
  {
    const std::vector<std::vector<uint32_t>> groups = { { 7, 11, 13, 17 }, { 19, 23 }, { 29, 31 },
                                                        { 37, 41 }, { 43, 47 } };
    const uint32_t periods[PSG_NTILES] = PSG_TILE_PERIODS;
    const uint32_t rows[PSG_NTILES] = PSG_TILE_ROWS;
    const uint32_t offsets[PSG_NTILES] = PSG_TILE_OFFSETS;
    std::vector<uint32_t> all;
    for (size_t t = 0; t < groups.size(); t++)
    {
      const std::vector<uint32_t> w = tileWindows(groups[t]);
      const uint32_t Q = rows[t];
      if (w.size() != periods[t] || all.size() != offsets[t] || Q != (periods[t] + 3) / 4 + 3)
        throw primesieve_error("pre-sieve tile layout mismatch");
      std::vector<uint32_t> tile((size_t) 4 * Q, 0);
      for (uint32_t r = 0; r < 4; r++)
        for (uint32_t j = 0; j < Q; j++)
          tile[(size_t) r * Q + j] = ~w[(r + 4ull * j) % periods[t]];      // the kernel marks composites
      all.insert(all.end(), tile.begin(), tile.end());
    }
    if (all.size() != PSG_TILE_WORDS)
      throw primesieve_error("pre-sieve tile layout mismatch");
    tiles_.allocate(api_, all.size() * 4);
    hipCheck(api_, api_->Memcpy(tiles_.get(), all.data(), all.size() * 4, hipMemcpyHostToDevice), "hipMemcpy tiles");
  }
  void* hc = nullptr;
  hipCheck(api_, api_->HostMalloc(&hc, 16, 0), "hipHostMalloc");
  hostFlags_ = (uint64_t*) hc;
  hostFlags_[0] = hostFlags_[1] = 0;
}

// Sieving-prime caches that reach this limit are generated on the device
// (below it the host's sieve is about as fast as the device round trips).
constexpr uint32_t kDevicePrimesMin = 1u << 24;

void HipBackend::ensureSievingPrimes(uint32_t limit)
{
  if (limit <= primesLimit_ && !hostPrimes_.empty())
    return;
  // grow geometrically so that a slowly rising stop does not regenerate
  uint64_t want = std::max<uint64_t>(limit, (uint64_t) primesLimit_ * 2);
  want = std::max<uint64_t>(want, 1u << 16);
  want = std::min<uint64_t>(want, 0xffffffffull);
  // list mode gives every thread S/8/threads whole 64-bit words
  if (want < kDevicePrimesMin || (sieveBytes_ / 8u) % wgSize_ != 0)
    installHostPrimes((uint32_t) want);
  else
    generatePrimesOnDevice((uint32_t) want);
}

/// The sieving primes firstCrossPrime_..want from the host's segmented sieve,
/// with their magics, uploaded to the device cache.
void HipBackend::installHostPrimes(uint32_t want)
{
  std::vector<uint32_t> pr = primesBetween(firstCrossPrime_, want);
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
  primesLimit_ = want;
}

/// The sieving primes firstCrossPrime_..want generated on the device: the
/// host sieves the seed primes up to sqrt(want); the segment kernel counts
/// the primes of every segment of [firstCrossPrime_, want] (one interval per
/// segment), the host turns the counts into output offsets, and a second pass
/// of the same kernel in list mode writes the primes there; psg_prime_magics
/// adds their magics.  The host keeps a copy of the primes (read back) for
/// its tier and bucket-room tables.
void HipBackend::generatePrimesOnDevice(uint32_t want)
{
  const psg_stats keep = stats_;
  const uint32_t S = sieveBytes_;
  const uint32_t seed = (uint32_t) std::min<uint64_t>(isqrt64(want) + 1, 0xffffffffull);
  if (seed > primesLimit_ || hostPrimes_.empty())
    installHostPrimes(std::max<uint32_t>(seed, 1u << 16));

  // one interval per segment of the global byte grid from firstCrossPrime_ on
  const uint64_t g0 = (firstCrossPrime_ - 7u) / 30u;
  std::vector<Interval> iv;
  for (uint64_t g = g0; ; g += S)
  {
    const uint64_t lo = (g == g0) ? firstCrossPrime_ : 30 * g + 7;
    if (lo > want)
      break;
    const uint64_t hi = std::min<uint64_t>(30 * (g + S) + 6, want);
    iv.push_back(Interval{ lo, hi, (uint32_t) iv.size() });
  }
  const uint32_t n = (uint32_t) iv.size();
  ExecOptions opt;
  opt.runSegs = 1;
  opt.slabSegs = std::max<uint32_t>(n, 1);
  std::vector<uint64_t> six((size_t) n * PSG_NCOUNTS);
  execute(iv, n, six.data(), opt);                       // pass 1: counts

  std::vector<uint32_t> offsets(n);
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; i++)
  {
    offsets[i] = (uint32_t) total;
    total += six[(size_t) i * PSG_NCOUNTS];
  }
  DeviceBuffer offs, newPrimes, newMagics;
  offs.allocate(api_, (size_t) n * 4);
  newPrimes.allocate(api_, total * 4 + 4);
  newMagics.allocate(api_, total * 8 + 8);
  hipCheck(api_, api_->Memcpy(offs.get(), offsets.data(), (size_t) n * 4, hipMemcpyHostToDevice),
           "hipMemcpy prime offsets");
  opt.listOut = newPrimes.addr();
  opt.listOffsets = offs.addr();
  execute(iv, n, six.data(), opt);                       // pass 2: the primes

  if (total)
  {
    uint64_t pa = newPrimes.addr(), ma = newMagics.addr();
    uint32_t cnt = (uint32_t) total;
    void* pm[] = { &pa, &ma, &cnt };
    hipCheck(api_, api_->ModuleLaunchKernel(fnMagics_, (cnt + 255u) / 256u, 1, 1, 256, 1, 1, 0,
                                            stream_, pm, nullptr), "hipModuleLaunchKernel(psg_prime_magics)");
  }
  std::vector<uint32_t> pr(total);
  hipCheck(api_, api_->StreamSynchronize(stream_), "psg_prime_magics");
  if (total)
    hipCheck(api_, api_->Memcpy(pr.data(), newPrimes.get(), total * 4, hipMemcpyDeviceToHost),
             "hipMemcpy primes to host");
  primes_.swap(newPrimes);
  magics_.swap(newMagics);
  invPrefix_.assign(pr.size() + 1, 0.0);
  for (size_t i = 0; i < pr.size(); i++)
    invPrefix_[i + 1] = invPrefix_[i] + 1.0 / pr[i];
  hostPrimes_.swap(pr);
  primesLimit_ = want;
  stats_ = keep;
}

uint64_t HipBackend::sievingPrimes(uint32_t limit, uint32_t* primes, uint64_t* magics, uint64_t cap)
{
  hipCheck(api_, api_->SetDevice(dev_.deviceIndex), "hipSetDevice");
  ensureSievingPrimes(limit);
  const uint64_t n = primesUpTo(limit);
  const uint64_t m = std::min<uint64_t>(n, cap);
  if (primes && m)
    std::memcpy(primes, hostPrimes_.data(), m * 4);
  if (magics && m)
    hipCheck(api_, api_->Memcpy(magics, magics_.get(), m * 8, hipMemcpyDeviceToHost), "hipMemcpy magics");
  return n;
}

uint32_t HipBackend::primesUpTo(uint64_t x) const
{
  const uint32_t v = (uint32_t) std::min<uint64_t>(x, 0xffffffffull);
  return (uint32_t) (std::upper_bound(hostPrimes_.begin(), hostPrimes_.end(), v) - hostPrimes_.begin());
}

hipEvent_t HipBackend::phaseEvent(size_t i)
{
  while (phaseEvents_.size() <= i)
  {
    hipEvent_t e = nullptr;
    hipCheck(api_, api_->EventCreate(&e), "hipEventCreate");
    phaseEvents_.push_back(e);
  }
  return phaseEvents_[i];
}

bool HipBackend::tilesFitLds() const
{
  if (tilesLds_ >= 0)
    return tilesLds_ == 1;
  const uint32_t base = sieveBytes_ + PSG_SIEVE_LDS_TAIL;
  const uint32_t with = base + 4u * PSG_SMALL_TILE_WORDS;
  bool fit = with <= ldsBytes_;
  if (fit && fnSieve_)
  {
    int occWith = 0, occWithout = 0;
    if (api_->ModuleOccupancy(&occWith, fnSieve_, (int) wgSize_, with) == hipSuccess &&
        api_->ModuleOccupancy(&occWithout, fnSieve_, (int) wgSize_, base) == hipSuccess)
      fit = occWith >= occWithout;
  }
  if (fnSieve_)
    tilesLds_ = fit ? 1 : 0;
  return fit;
}

uint32_t HipBackend::persistentGrid(uint32_t nSegs)
{
  if (!persistentWGs_)
  {
    int perCU = 0;
    if (api_->ModuleOccupancy(&perCU, fnSieve_, (int) wgSize_, sieveLds()) != hipSuccess || perCU < 1)
      perCU = 1;
    persistentWGs_ = (uint32_t) perCU * (dev_.computeUnits ? dev_.computeUnits : 256);
  }
  return std::max<uint32_t>(1, std::min(nSegs, persistentWGs_));
}

/// Workgroups of a persistent bucket-fill launch for nItems work items: as
/// many as the device holds at once, fewer when there is less work.
uint32_t HipBackend::fillGrid(uint32_t nItems)
{
  if (!fillWGs_)
  {
    int perCU = 0;
    if (api_->ModuleOccupancy(&perCU, fnBucketFill_, (int) PSG_BUCKET_THREADS, PSG_BUCKET_LDS) != hipSuccess
        || perCU < 1)
      perCU = 1;
    fillWGs_ = (uint32_t) perCU * (dev_.computeUnits ? dev_.computeUnits : 256);
  }
  return std::max<uint32_t>(1, std::min(nItems, fillWGs_));
}

/// Sieve every interval; out6[slot * 6 + t] receives the six counts of the
/// intervals whose `out` is slot.  Intervals are cut into runs of at most
/// opt.runSegs segments, runs are packed into slabs, and each slab is one
/// bucket fill (when big primes exist) plus one persistent sieve launch.
///
/// Every segment's bucket list gets a room of E + m*sqrt(E) + 64 entries,
/// E = 8 * S * sum(1/p) over its big primes (the expected crossings).  The
/// fill counts what it appends; if any list outgrew its room, the counted
/// totals become the rooms and the call runs once more.
void HipBackend::execute(const std::vector<Interval>& iv, uint32_t outs, uint64_t* out6,
                         const ExecOptions& opt)
{
  const auto t0 = std::chrono::steady_clock::now();
  stats_ = psg_stats{};
  stats_.sieve_bytes = sieveBytes_;
  stats_.work_group_size = wgSize_;
  std::fill(out6, out6 + (size_t) outs * PSG_NCOUNTS, 0);
  hipCheck(api_, api_->SetDevice(dev_.deviceIndex), "hipSetDevice");

  uint64_t maxHi = 0;
  for (const Interval& v : iv)
    if (v.lo <= v.hi && v.hi >= 7) maxHi = std::max(maxHi, v.hi);
  if (!maxHi)
    return;
  if (opt.noAlloc)
  {
    // the sieving-prime cache is device memory: it must be there already
    if (hostPrimes_.empty() || isqrt64(maxHi) > primesLimit_)
      throw primesieve_error("the sieving primes up to " + std::to_string(isqrt64(maxHi)) +
                             " are not cached on the device; a call that may allocate "
                             "(psg_gpu_count_all or psg_gpu_sieving_primes) must cache them first");
  }
  else
    ensureSievingPrimes((uint32_t) isqrt64(maxHi));

  const uint32_t S = sieveBytes_;
  uint32_t shift = 0;
  while ((1u << shift) < S) shift++;
  const bool pow2 = (1u << shift) == S;
  // bucket primes: p > bucketFactor * S, and always p > 2^16 (the fill
  // kernel's remainders rely on it)
  const uint32_t bucketFactor = opt.bucketFactor ? opt.bucketFactor : bucketFactor_;
  bigBegin_ = (pow2 && bucketFactor)
            ? primesUpTo(std::max<uint64_t>((uint64_t) S * bucketFactor, 1u << 16))
            : (uint32_t) hostPrimes_.size();
  const uint32_t windowSegs = std::max<uint32_t>(1, std::min(opt.windowSegs, PSG_MAX_WINDOW_SEGS));
  const double hBig = invPrefix_[bigBegin_];

  // ---- runs -----------------------------------------------------------------
  // A run never expects more bucket entries than a slab may hold: slabs are
  // cut between runs, so a long interval with a heavy bucket load (high
  // starts: ~4.9e5 entries per segment at 1e19) becomes several runs, and
  // several slabs, instead of one slab past the 2^32-entry index range.
  runs_.clear();
  uint64_t segsTotal = 0;
  for (const Interval& v : iv)
  {
    const uint64_t lo = std::max<uint64_t>(v.lo, 7);
    if (v.hi < 7 || lo > v.hi)
      continue;
    const ByteRange br = byteRange(lo, v.hi);
    const uint32_t np = primesUpTo(isqrt64(v.hi));
    const uint64_t nseg = (br.gHi - br.gLo) / S + 1;
    const double perSeg = (np > bigBegin_) ? 8.0 * S * (invPrefix_[np] - hBig) : 0.0;
    uint32_t runSegs = std::max<uint32_t>(1, opt.runSegs);
    if (perSeg > 0 && perSeg * runSegs > opt.entryBudget)
      runSegs = (uint32_t) std::max(1.0, std::floor(opt.entryBudget / perSeg));
    for (uint64_t s = 0; s < nseg; s += runSegs)
    {
      PsgRun r;
      std::memset(&r, 0, sizeof r);
      r.B0 = br.gLo + s * S;
      r.gLo = br.gLo;
      r.gHi = br.gHi;
      r.loMask = br.loMask;
      r.hiMask = br.hiMask;
      r.nprimes = np;
      r.out = v.out;
      r.nseg = (uint32_t) std::min<uint64_t>(runSegs, nseg - s);
      runs_.push_back(r);
    }
    segsTotal += nseg;
  }
  if (runs_.empty())
    return;

  // ---- slabs, windows and the expected bucket entries per segment ------------
  slabs_.clear();
  wins_.clear();
  segExpect_.clear();
  {
    Slab cur{ 0, 0, 0, 0, 0, 0, 0, 0.0 };
    for (size_t i = 0; i < runs_.size(); i++)
    {
      PsgRun& r = runs_[i];
      const bool big = r.nprimes > bigBegin_;
      const double perSeg = big ? 8.0 * S * (invPrefix_[r.nprimes] - hBig) : 0.0;
      const double e = perSeg * r.nseg;
      if (cur.run1 > cur.run0 &&
          (cur.nSegs + r.nseg > opt.slabSegs || cur.est + e > opt.entryBudget))
      {
        cur.win1 = wins_.size();
        slabs_.push_back(cur);
        cur = Slab{ i, i, wins_.size(), wins_.size(), 0, 0, cur.segBase + cur.nSegs, 0.0 };
      }
      r.segStart = cur.nSegs;
      if (big)
      {
        const uint32_t nBig = r.nprimes - bigBegin_;
        const uint32_t chunks = std::max<uint32_t>(1, std::min<uint32_t>(
            opt.maxChunks, (nBig + opt.primesPerChunk - 1) / opt.primesPerChunk));
        for (uint32_t s = 0; s < r.nseg; s += windowSegs)
        {
          PsgWindow w;
          std::memset(&w, 0, sizeof w);
          w.segs = std::min(windowSegs, r.nseg - s);
          w.base = r.B0 + (uint64_t) s * S;
          w.firstSeg = cur.nSegs + s;
          w.bigEnd = r.nprimes;
          w.nChunks = chunks;
          w.wgBase = cur.nWGs;
          cur.nWGs += chunks;
          wins_.push_back(w);
        }
      }
      segExpect_.insert(segExpect_.end(), r.nseg, perSeg);
      cur.est += e;
      cur.nSegs += r.nseg;
      cur.run1 = i + 1;
    }
    cur.win1 = wins_.size();
    slabs_.push_back(cur);
  }

  // the segments of a run share their expectation: one room per run
  std::vector<uint32_t> room(segExpect_.size());
  double lastE = -1.0;
  uint32_t lastRoom = 0;
  for (size_t s = 0; s < room.size(); s++)
  {
    const double E = segExpect_[s];
    if (E != lastE)
    {
      lastE = E;
      lastRoom = E > 0 ? (uint32_t) std::min(std::ceil(E + bucketMargin_ * std::sqrt(E) + 64.0), 4.0e9) : 0u;
    }
    room[s] = lastRoom;
  }
  if (!launchPlan(outs, room, opt))
  {
    // a list outgrew its room: the fill counted every list exactly
    hipCheck(api_, api_->Memcpy(room.data(), (const void*) (uintptr_t) listCountAddr_, room.size() * 4,
                                hipMemcpyDeviceToHost),
             "hipMemcpy bucket list counts");
    stats_.bucket_retries = 1;
    if (!launchPlan(outs, room, opt))
      throw primesieve_error("big-prime bucket list outgrew its counted room");
  }
  std::copy(hostCounts_.begin(), hostCounts_.end(), out6);
  stats_.segments = segsTotal;
  stats_.launches = (uint32_t) slabs_.size();
  stats_.sieving_primes = primesUpTo(isqrt64(maxHi));
  stats_.wall_ms = std::chrono::duration<double, std::milli>(
                       std::chrono::steady_clock::now() - t0).count();
}

/// Launch the current plan with the given bucket list rooms (one per
/// segment of the call).  False when a list outgrew its room (the counts are
/// then incomplete and the list counters at listCountAddr_ hold the exact
/// list sizes).
///
/// Every argument the kernels take is in one host-built blob, uploaded with a
/// single copy (the PSG_DESC_* layout of kernel_abi.h): the clear block, one
/// record per slab, then the run, window and list-offset tables.  A kernel's
/// only parameter is the address of its record, so a launch series depends on
/// the plan only through its slab count and which slabs fill buckets.  Three
/// forms:
///   * eager: psg_clear, then per slab psg_bucket_fill (when it has big
///     primes) and psg_sieve_count, each timed with events; a batch with two
///     or more bucket slabs fills slab k+1 on a second stream, into the other
///     entry buffer, while slab k is sieved;
///   * graph (opt.graph): the same series on one stream with fixed persistent
///     grids, captured once per shape and replayed; no memset nodes (the
///     clear kernel zeroes counters and counts);
///   * no allocation (opt.noAlloc): every per-call buffer is carved out of the
///     code object's psg_scratch array (resolved at init); one stream, one
///     entry buffer reused by every slab.
bool HipBackend::launchPlan(uint32_t outs, const std::vector<uint32_t>& room, const ExecOptions& opt)
{
  const uint32_t S = sieveBytes_;
  uint32_t shift = 0;
  while ((1u << shift) < S) shift++;

  // per-slab list offsets, and the largest slab's entries
  std::vector<uint32_t> listOff;
  std::vector<size_t> offPos(slabs_.size());
  uint64_t maxEntries = 0;
  for (size_t si = 0; si < slabs_.size(); si++)
  {
    const Slab& sl = slabs_[si];
    offPos[si] = listOff.size();
    uint64_t acc = 0;
    for (uint32_t s = 0; s < sl.nSegs; s++)
    {
      listOff.push_back((uint32_t) acc);
      acc += room[sl.segBase + s];
    }
    if (acc > 0xffffffffull - (1u << 20))   // entry indices plus a few thread strides stay below 2^32
      throw primesieve_error("big-prime bucket slab exceeds 2^32 entries");
    listOff.push_back((uint32_t) acc);
    maxEntries = std::max(maxEntries, acc);
  }

  // ---- sizes -----------------------------------------------------------------
  const size_t nSlabs = slabs_.size();
  const size_t nSegsAll = room.size();
  const size_t runOff = PSG_DESC_HEADER + nSlabs * PSG_SLAB_RECORD;
  const size_t winOff = runOff + runs_.size() * sizeof(PsgRun);
  const size_t offOff = winOff + wins_.size() * sizeof(PsgWindow);
  const size_t descBytes = offOff + listOff.size() * 4;
  const size_t countBytes = (size_t) outs * PSG_NCOUNTS * sizeof(uint64_t);
  const size_t workBytes = nSlabs * 8;             // per slab: sieve and fill work counters
  const size_t listCountBytes = nSegsAll * 4 + 4;
  const size_t entryBytes = (size_t) maxEntries * 4;
  size_t bucketSlabs = 0;
  for (const Slab& sl : slabs_)
    bucketSlabs += (sl.win1 > sl.win0) ? 1 : 0;
  // Pipelined (eager batches): neither kernel's workgroup fits in LDS beside
  // the other's, so the fill of slab k+1 takes the CUs that the segment
  // kernel's workgroups leave at the end of slab k: the overlap is the
  // launch tails.
  const bool pipelined = opt.pipeline && !opt.graph && !opt.noAlloc && bucketSlabs >= 2;

  // ---- device memory -----------------------------------------------------------
  uint64_t desc, counts, work, listCount, flags, entries[2] = { 0, 0 };
  if (opt.noAlloc)
  {
    // the scratch array: flags | counts | work counters | list counters |
    // descriptors | bucket entries, each 256-byte aligned
    uint64_t at = 0;
    const auto carve = [&](size_t bytes) {
      const uint64_t a = scratchAddr_ + at;
      at += (bytes + 255) & ~(uint64_t) 255;
      return a;
    };
    flags = carve(16);
    counts = carve(countBytes);
    work = carve(workBytes);
    listCount = carve(listCountBytes);
    desc = carve(descBytes);
    entries[0] = carve(entryBytes);
    if (at > scratchBytes_)
      throw primesieve_error("count needs " + std::to_string(at) + " bytes of device scratch; the "
                             "no-allocation entry point has " + std::to_string(scratchBytes_) +
                             " (use psg_gpu_count_all for this interval)");
  }
  else
  {
    descriptors_.allocate(api_, descBytes);
    counts_.allocate(api_, countBytes);
    workCounters_.allocate(api_, workBytes);
    listCount_.allocate(api_, listCountBytes);
    flags_.allocate(api_, 16);
    if (maxEntries)
      entries_[0].allocate(api_, entryBytes);
    if (maxEntries && pipelined)
      entries_[1].allocate(api_, entryBytes);
    desc = descriptors_.addr();
    counts = counts_.addr();
    work = workCounters_.addr();
    listCount = listCount_.addr();
    flags = flags_.addr();
    entries[0] = entries_[0].addr();
    entries[1] = entries_[1].addr();
  }
  listCountAddr_ = listCount;

  // ---- the descriptor blob ----------------------------------------------------
  unsigned char* blob;
  if (opt.noAlloc)
  {
    staging_.assign(descBytes, 0);
    blob = staging_.data();
  }
  else
  {
    blob = (unsigned char*) pinnedDesc_.ensure(api_, descBytes);
    std::memset(blob, 0, descBytes);
  }
  PsgClearArgs clr;
  std::memset(&clr, 0, sizeof clr);
  clr.addr[0] = counts;    clr.words[0] = (uint32_t) (countBytes / 4);
  clr.addr[1] = work;      clr.words[1] = (uint32_t) (workBytes / 4);
  clr.addr[2] = listCount; clr.words[2] = (uint32_t) nSegsAll;
  clr.addr[3] = flags;     clr.words[3] = 4;
  static_assert(sizeof(PsgClearArgs) <= PSG_DESC_HEADER, "clear block");
  static_assert(sizeof(PsgSieveArgs) <= PSG_SLAB_FILL_OFFSET, "sieve record");
  static_assert(PSG_SLAB_FILL_OFFSET + sizeof(PsgBucketArgs) <= PSG_SLAB_RECORD, "fill record");
  std::memcpy(blob, &clr, sizeof clr);

  PsgSieveArgs args;
  std::memset(&args, 0, sizeof args);
  args.primes = primes_.addr();
  args.magics = magics_.addr();
  args.tiles = tiles_.addr();
  args.tilesInLds = (tiles_.addr() && tilesFitLds()) ? 1u : 0u;
  args.counts = counts;
  args.segBytes = S;
  args.tier0End = primesUpTo(std::min<uint32_t>(PSG_TIER0_MAX_PRIME, S / 8));
  args.tier1End = primesUpTo(std::min<uint32_t>(PSG_TIER1_MAX_PRIME, S));
  args.tier2End = primesUpTo(S);
  args.tier3End = bigBegin_;
  args.primesOut = opt.listOut;
  args.outOffsets = opt.listOffsets;

  PsgBucketArgs b;
  std::memset(&b, 0, sizeof b);
  b.primes = primes_.addr();
  b.overflow = flags;
  b.total = flags + 8;
  b.segShift = shift;
  b.bigBegin = bigBegin_;

  std::vector<uint32_t> parityOf(nSlabs, 0);
  uint32_t parity = 0;
  for (size_t si = 0; si < nSlabs; si++)
  {
    const Slab& sl = slabs_[si];
    const bool buckets = sl.win1 > sl.win0;
    const uint64_t slabOff = desc + offOff + offPos[si] * 4;
    const uint64_t slabCount = listCount + (uint64_t) sl.segBase * 4;
    parityOf[si] = parity;
    args.runs = desc + runOff + sl.run0 * sizeof(PsgRun);
    args.nRuns = (uint32_t) (sl.run1 - sl.run0);
    args.nSegs = sl.nSegs;
    args.buckets = buckets ? entries[parity] : 0;
    args.listOff = slabOff;
    args.listCount = slabCount;
    args.workCounter = work + si * 8;
    b.windows = desc + winOff + sl.win0 * sizeof(PsgWindow);
    b.nWindows = (uint32_t) (sl.win1 - sl.win0);
    b.nItems = buckets ? sl.nWGs : 0;
    b.listOff = slabOff;
    b.listCount = slabCount;
    b.entries = entries[parity];
    b.workCounter = work + si * 8 + 4;
    unsigned char* rec = blob + PSG_DESC_HEADER + si * PSG_SLAB_RECORD;
    std::memcpy(rec, &args, sizeof args);
    std::memcpy(rec + PSG_SLAB_FILL_OFFSET, &b, sizeof b);
    if (buckets && pipelined)
      parity ^= 1u;
  }
  std::memcpy(blob + runOff, runs_.data(), runs_.size() * sizeof(PsgRun));
  if (!wins_.empty())
    std::memcpy(blob + winOff, wins_.data(), wins_.size() * sizeof(PsgWindow));
  std::memcpy(blob + offOff, listOff.data(), listOff.size() * 4);
  hipCheck(api_, api_->MemcpyAsync((void*) (uintptr_t) desc, blob, descBytes, hipMemcpyHostToDevice, stream_),
           "hipMemcpyAsync descriptors");

  // ---- launches -----------------------------------------------------------------
  const uint32_t fillLds = PSG_BUCKET_LDS;
  const bool fixedGrids = opt.graph;
  const auto recordAddr = [&](size_t si) { return desc + PSG_DESC_HEADER + si * PSG_SLAB_RECORD; };
  const auto launchClear = [&](hipStream_t st) {
    uint64_t a = desc;
    void* pa[] = { &a };
    hipCheck(api_, api_->ModuleLaunchKernel(fnClear_, kClearWGs, 1, 1, 256, 1, 1, 0, st, pa, nullptr),
             "hipModuleLaunchKernel(psg_clear)");
  };
  const auto launchFill = [&](size_t si, hipStream_t st) {
    uint64_t a = recordAddr(si) + PSG_SLAB_FILL_OFFSET;
    void* pa[] = { &a };
    const uint32_t grid = fixedGrids ? fillGrid(~0u) : fillGrid(slabs_[si].nWGs);
    hipCheck(api_, api_->ModuleLaunchKernel(fnBucketFill_, grid, 1, 1, PSG_BUCKET_THREADS, 1, 1,
                                            fillLds, st, pa, nullptr),
             "hipModuleLaunchKernel(psg_bucket_fill)");
  };
  const auto launchSieve = [&](size_t si, hipStream_t st) {
    uint64_t a = recordAddr(si);
    void* pa[] = { &a };
    const uint32_t grid = fixedGrids ? persistentGrid(~0u) : persistentGrid(slabs_[si].nSegs);
    hipCheck(api_, api_->ModuleLaunchKernel(fnSieve_, grid, 1, 1, wgSize_, 1, 1, sieveLds(), st,
                                            pa, nullptr),
             "hipModuleLaunchKernel(psg_sieve_count)");
  };

  struct SlabEvents { hipEvent_t fb, fe, sb, se; };
  std::vector<SlabEvents> ev;
  stats_.graph = 0;
  hipCheck(api_, api_->EventRecord(evStart_, stream_), "hipEventRecord");
  if (opt.graph)
  {
    std::vector<uint64_t> key = { desc, nSlabs, persistentGrid(~0u), fillGrid(~0u), sieveLds() };
    for (const Slab& sl : slabs_)
      key.push_back(sl.win1 > sl.win0);
    bool captured = false;
    hipGraphExec_t exec = graphFor(key, captured, [&]() {
      launchClear(stream_);
      for (size_t si = 0; si < nSlabs; si++)
      {
        if (slabs_[si].win1 > slabs_[si].win0)
          launchFill(si, stream_);
        launchSieve(si, stream_);
      }
    });
    hipCheck(api_, api_->GraphLaunch(exec, stream_), "hipGraphLaunch");
    stats_.graph = captured ? 2u : 1u;
  }
  else if (opt.noAlloc)
  {
    launchClear(stream_);
    for (size_t si = 0; si < nSlabs; si++)
    {
      if (slabs_[si].win1 > slabs_[si].win0)
        launchFill(si, stream_);
      launchSieve(si, stream_);
    }
  }
  else
  {
    launchClear(stream_);
    hipStream_t fs = stream_;
    if (pipelined)
    {
      // the fills start after the upload and the clear
      fs = fillStream_;
      hipCheck(api_, api_->EventRecord(evSetup_, stream_), "hipEventRecord");
      hipCheck(api_, api_->StreamWaitEvent(fs, evSetup_, 0), "hipStreamWaitEvent");
    }
    // events: fill begin/end on fs, sieve begin/end on stream_, per slab
    ev.assign(nSlabs, SlabEvents{ nullptr, nullptr, nullptr, nullptr });
    size_t nev = 0;
    hipEvent_t lastUse[2] = { nullptr, nullptr };   // the last sieve reading each entry buffer
    for (size_t si = 0; si < nSlabs; si++)
    {
      const bool buckets = slabs_[si].win1 > slabs_[si].win0;
      const uint32_t par = parityOf[si];
      if (buckets)
      {
        if (pipelined && lastUse[par])
          hipCheck(api_, api_->StreamWaitEvent(fs, lastUse[par], 0), "hipStreamWaitEvent");
        ev[si].fb = phaseEvent(nev++);
        ev[si].fe = phaseEvent(nev++);
        hipCheck(api_, api_->EventRecord(ev[si].fb, fs), "hipEventRecord");
        launchFill(si, fs);
        hipCheck(api_, api_->EventRecord(ev[si].fe, fs), "hipEventRecord");
        if (pipelined)
          hipCheck(api_, api_->StreamWaitEvent(stream_, ev[si].fe, 0), "hipStreamWaitEvent");
      }
      ev[si].sb = phaseEvent(nev++);
      ev[si].se = phaseEvent(nev++);
      hipCheck(api_, api_->EventRecord(ev[si].sb, stream_), "hipEventRecord");
      launchSieve(si, stream_);
      hipCheck(api_, api_->EventRecord(ev[si].se, stream_), "hipEventRecord");
      if (buckets)
        lastUse[par] = ev[si].se;
    }
  }
  hipCheck(api_, api_->EventRecord(evStop_, stream_), "hipEventRecord");
  const size_t nCounts = (size_t) outs * PSG_NCOUNTS;
  hostCounts_.resize(nCounts);
  uint64_t* countsTo = opt.noAlloc ? hostCounts_.data()
                                   : (uint64_t*) pinnedCounts_.ensure(api_, nCounts * 8);
  hipCheck(api_, api_->MemcpyAsync(hostFlags_, (void*) (uintptr_t) flags, 16, hipMemcpyDeviceToHost, stream_),
           "hipMemcpyAsync flags");
  hipCheck(api_, api_->MemcpyAsync(countsTo, (void*) (uintptr_t) counts, nCounts * 8,
                                   hipMemcpyDeviceToHost, stream_), "hipMemcpyAsync counts");
  hipCheck(api_, api_->StreamSynchronize(stream_), "sieve kernels");
  if (countsTo != hostCounts_.data())
    std::copy(countsTo, countsTo + nCounts, hostCounts_.begin());

  float ms = 0.f;
  hipCheck(api_, api_->EventElapsedTime(&ms, evStart_, evStop_), "hipEventElapsedTime");
  double fillMs = 0, sieveMs = 0;
  for (const SlabEvents& e : ev)
  {
    float t = 0.f;
    if (e.fb)
    {
      hipCheck(api_, api_->EventElapsedTime(&t, e.fb, e.fe), "hipEventElapsedTime");
      fillMs += t;
    }
    hipCheck(api_, api_->EventElapsedTime(&t, e.sb, e.se), "hipEventElapsedTime");
    sieveMs += t;
  }
  stats_.kernel_ms += sieveMs;
  stats_.fill_ms += fillMs;
  stats_.device_ms += ms;
  stats_.bucket_entries = hostFlags_[1];
  return (hostFlags_[0] & 0xffffffffull) == 0;
}

/// The instantiated launch series for `key`, captured from `record` on the
/// backend's stream the first time the key is seen (captured = true).  At
/// most kMaxGraphs series are kept; the oldest is dropped first.
hipGraphExec_t HipBackend::graphFor(const std::vector<uint64_t>& key, bool& captured,
                                    const std::function<void()>& record)
{
  captured = false;
  for (const CachedGraph& g : graphs_)
    if (g.key == key)
      return g.exec;
  hipCheck(api_, api_->StreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal),
           "hipStreamBeginCapture");
  try
  {
    record();
  }
  catch (...)
  {
    hipGraph_t g = nullptr;
    if (api_->StreamEndCapture(stream_, &g) == hipSuccess && g)
      (void) api_->GraphDestroy(g);
    throw;
  }
  hipGraph_t graph = nullptr;
  hipCheck(api_, api_->StreamEndCapture(stream_, &graph), "hipStreamEndCapture");
  hipGraphExec_t exec = nullptr;
  const hipError_t e = api_->GraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void) api_->GraphDestroy(graph);
  hipCheck(api_, e, "hipGraphInstantiate");
  if (graphs_.size() >= kMaxGraphs)
  {
    (void) api_->GraphExecDestroy(graphs_.front().exec);
    graphs_.erase(graphs_.begin());
  }
  graphs_.push_back(CachedGraph{ key, exec });
  captured = true;
  return exec;
}

void HipBackend::countAll(uint64_t lo, uint64_t hi, uint64_t out[PSG_NCOUNTS])
{
  // a slab (one persistent launch of each kernel) holds at most
  // segmentsPerLaunch_ segments, the reference's bound on a launch
  // (src/gpu/cuda/CudaBackend.cpp:466)
  ExecOptions opt;
  opt.runSegs = std::min(slabSegments_, segmentsPerLaunch_);
  opt.slabSegs = opt.runSegs;
  opt.windowSegs = windowSegs_;
  execute({ Interval{ lo, hi, 0 } }, 1, out, opt);
}

void HipBackend::countBatch(const std::vector<Interval>& iv, uint32_t outs, uint64_t* out6, bool graph)
{
  ExecOptions opt;
  opt.graph = graph;
  opt.runSegs = std::min<uint32_t>({ 2048u, std::max<uint32_t>(slabSegments_, 1), segmentsPerLaunch_ });
  opt.windowSegs = PSG_MAX_WINDOW_SEGS;  // few prime sweeps per run
  opt.primesPerChunk = 4096;
  opt.maxChunks = 16384;
  opt.slabSegs = std::min(std::max<uint32_t>(opt.runSegs, batchSlabSegs_), segmentsPerLaunch_);
  opt.pipeline = true;
  // A batch of short intervals spends its fill time on each interval's sweep
  // over the big primes, one window visit each: by default its bucket tier
  // starts at 12 segments' bytes instead of 7 (the 8192-interval batch:
  // 9318 -> 9636 intervals/s; 10: 9578, 16: 8968).  A tuning's explicit
  // bucket_factor applies to batches as well.
  if (bucketAuto_ && bucketFactor_)
    opt.bucketFactor = 12;
  execute(iv, outs, out6, opt);
}

void HipBackend::countAllNoAlloc(uint64_t lo, uint64_t hi, uint64_t out[PSG_NCOUNTS])
{
  ExecOptions opt;
  opt.runSegs = std::min(slabSegments_, segmentsPerLaunch_);
  opt.slabSegs = opt.runSegs;
  opt.windowSegs = windowSegs_;
  opt.noAlloc = true;
  // one slab's expected bucket entries take at most 40% of the scratch, so
  // that its rooms (expected + margin) and the call's tables fit beside them
  opt.entryBudget = std::min(opt.entryBudget, 0.4 * (double) scratchBytes_ / 4.0);
  execute({ Interval{ lo, hi, 0 } }, 1, out, opt);
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
    return std::unique_ptr<GpuBackend>(new psg::HipBackend(api, dev, config, psg::hipTuning()));
  }
  catch (const std::exception& e)
  {
    error = e.what();
    return std::unique_ptr<GpuBackend>();
  }
}

} // namespace gpu
} // namespace primesieve
