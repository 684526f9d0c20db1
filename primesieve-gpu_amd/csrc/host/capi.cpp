///
/// @file  capi.cpp
/// @brief extern "C" layer (include/primesieve_gpu.h) over the C++ GPU API.
///        uint64_t-returning calls follow the reference C API's convention
///        (src/api-c.cpp:143-225): catch, print to stderr, errno = EDOM,
///        return PSG_ERROR.
///

#include "GpuBackend.hpp"
#include "GpuSieve.hpp"
#include "HipBackend.hpp"
#include "tables.hpp"

#include <primesieve/gpu.hpp>
#include <primesieve_gpu.h>
#include <primesieve/primesieve_error.hpp>

#include <cerrno>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

namespace {

int copyOut(const std::string& s, char* buf, size_t len)
{
  if (buf && len > 0)
  {
    size_t n = s.size() < len - 1 ? s.size() : len - 1;
    std::memcpy(buf, s.data(), n);
    buf[n] = '\0';
  }
  return (int) s.size();
}

template <typename F>
uint64_t guarded(const char* name, F f)
{
  try
  {
    return f();
  }
  catch (const std::exception& e)
  {
    std::cerr << name << ": " << e.what() << std::endl;
    errno = EDOM;
    return PSG_ERROR;
  }
}

void setLastError(const std::string& s)
{
  std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
  primesieve::gpu::lastErrorRef() = s;
}

} // namespace

extern "C" {

int psg_gpu_available(void) { return primesieve::gpu_available() ? 1 : 0; }

int psg_gpu_device_count(void) { return (int) primesieve::gpu_devices().size(); }

int psg_gpu_device_describe(int index, char* buf, size_t len)
{
  std::vector<std::string> d = primesieve::gpu_devices();
  if (index < 0 || index >= (int) d.size())
    return -1;
  return copyOut(d[(size_t) index], buf, len);
}

void psg_set_gpu_device(int device_index) { primesieve::set_gpu_device(device_index); }
int  psg_get_gpu_device(void) { return primesieve::get_gpu_device(); }
void psg_set_gpu_enabled(int enable) { primesieve::set_gpu_enabled(enable != 0); }
int  psg_get_gpu_enabled(void) { return primesieve::get_gpu_enabled() ? 1 : 0; }
void psg_set_gpu_sieve_size(int kilobytes) { primesieve::set_gpu_sieve_size(kilobytes); }
void psg_set_gpu_work_group_size(int size) { primesieve::set_gpu_work_group_size(size); }
void psg_set_gpu_segments_per_launch(int s) { primesieve::set_gpu_segments_per_launch(s); }
void psg_set_gpu_presieve(int enable) { primesieve::set_gpu_presieve(enable != 0); }

int psg_set_gpu_tuning(const psg_tuning* t)
{
  if (!t)
    return -1;
  primesieve::gpu::GpuConfig c;        // the reference's four tunables
  c.sieveBytes = t->sieve_bytes;
  c.workGroupSize = t->work_group_size;
  c.segmentsPerLaunch = t->segments_per_launch;
  c.preSieve = t->presieve != 0;
  psg::HipTuning h;                     // the HIP backend's own
  h.slabSegments = t->slab_segments;
  h.bucketFactor = t->bucket_factor;
  h.bucketMargin = t->bucket_margin;
  primesieve::gpu::setTuning(c, h);
  return 0;
}

int psg_get_gpu_tuning(psg_tuning* t)
{
  if (!t)
    return -1;
  std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
  psg::HipBackend* b = primesieve::gpu::activeHipBackend();
  if (!b)
    return -1;
  *t = b->tuning();
  return 0;
}

int psg_gpu_active_device(char* buf, size_t len)
{
  return copyOut(primesieve::gpu_active_device(), buf, len);
}

int psg_gpu_last_error(char* buf, size_t len)
{
  return copyOut(primesieve::gpu_last_error(), buf, len);
}

int psg_gpu_worthwhile(uint64_t start, uint64_t stop)
{
  return primesieve::gpu_worthwhile(start, stop) ? 1 : 0;
}

uint64_t psg_gpu_count_primes(uint64_t a, uint64_t b)
{ return guarded("psg_gpu_count_primes", [&] { return primesieve::gpu_count_primes(a, b); }); }
uint64_t psg_gpu_count_twins(uint64_t a, uint64_t b)
{ return guarded("psg_gpu_count_twins", [&] { return primesieve::gpu_count_twins(a, b); }); }
uint64_t psg_gpu_count_triplets(uint64_t a, uint64_t b)
{ return guarded("psg_gpu_count_triplets", [&] { return primesieve::gpu_count_triplets(a, b); }); }
uint64_t psg_gpu_count_quadruplets(uint64_t a, uint64_t b)
{ return guarded("psg_gpu_count_quadruplets", [&] { return primesieve::gpu_count_quadruplets(a, b); }); }
uint64_t psg_gpu_count_quintuplets(uint64_t a, uint64_t b)
{ return guarded("psg_gpu_count_quintuplets", [&] { return primesieve::gpu_count_quintuplets(a, b); }); }
uint64_t psg_gpu_count_sextuplets(uint64_t a, uint64_t b)
{ return guarded("psg_gpu_count_sextuplets", [&] { return primesieve::gpu_count_sextuplets(a, b); }); }

int psg_gpu_try_count(uint64_t start, uint64_t stop, int count_type, uint64_t* result)
{
  uint64_t r = 0;
  if (!result || !primesieve::gpu_try_count(start, stop, count_type, r))
    return 0;
  *result = r;
  return 1;
}

int psg_gpu_count_all(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  if (!counts)
    return -1;
  try
  {
    primesieve::gpu_count_all(start, stop, counts);
    return 0;
  }
  catch (const std::exception& e)
  {
    setLastError(e.what());
    return -1;
  }
}

int psg_gpu_count_all_noalloc(uint64_t start, uint64_t stop, uint64_t counts[6])
{
  if (!counts)
    return -1;
  std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
  try
  {
    if (!primesieve::gpu::countAllNoAllocOnGpu(start, stop, counts))
      throw primesieve::primesieve_error("no GPU backend yet: a call that may allocate (psg_gpu_count_all, "
                                         "psg_gpu_sieving_primes) creates it and its sieving-prime cache");
    return 0;
  }
  catch (const std::exception& e)
  {
    primesieve::gpu::lastErrorRef() = e.what();
    return -1;
  }
}

int psg_gpu_count_batch(const psg_query* queries, size_t n, uint64_t* out, int mode)
{
  if ((!queries || !out) && n)
    return -1;
  try
  {
    std::vector<primesieve::GpuQuery> q(n);
    for (size_t i = 0; i < n; i++)
      q[i] = primesieve::GpuQuery{ queries[i].start, queries[i].stop, queries[i].k };
    std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
    if (!primesieve::gpu::countBatchOnGpu(q.data(), n, out, mode))
      throw primesieve::primesieve_error("GPU is not available: " + primesieve::gpu::lastErrorRef());
    return 0;
  }
  catch (const std::exception& e)
  {
    std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
    primesieve::gpu::lastErrorRef() = e.what();
    return -1;
  }
}

uint64_t psg_gpu_sieving_primes(uint32_t limit, uint32_t* primes, uint64_t* magics, uint64_t cap)
{
  return guarded("psg_gpu_sieving_primes", [&] {
    std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
    psg::HipBackend* b = primesieve::gpu::activeHipBackend();
    if (!b)
      throw primesieve::primesieve_error("GPU is not available: " + primesieve::gpu::lastErrorRef());
    return b->sievingPrimes(limit, primes, magics, cap);
  });
}

int psg_gpu_last_stats(psg_stats* out)
{
  if (!out)
    return -1;
  std::lock_guard<std::mutex> lock(primesieve::gpu::gpuMutex());
  psg::HipBackend* b = primesieve::gpu::existingHipBackend();
  if (!b)
    return -1;
  *out = b->lastStats();
  return 0;
}

int psg_algorithmic_work(uint64_t start, uint64_t stop, uint32_t seg_bytes, psg_work* out)
{
  if (!out)
    return -1;
  std::memset(out, 0, sizeof *out);
  try
  {
    uint64_t lo = start < 7 ? 7 : start;
    if (stop < 7 || lo > stop)
      return 0;
    const psg::ByteRange br = psg::byteRange(lo, stop);
    const uint32_t sq = (uint32_t) psg::isqrt64(stop);
    const std::vector<uint32_t> pr = psg::primesBetween(PSG_FIRST_CROSS_PRIME, sq);
    const uint64_t bytes = br.gHi - br.gLo + 1;
    out->words32 = (bytes + 3) / 4;
    out->sieving_primes = pr.size();
    out->crossings = psg::crossings(pr, lo, stop, PSG_FIRST_CROSS_PRIME, sq);
    const uint64_t big = 8ull * seg_bytes;
    out->crossings_big = big < sq ? psg::crossings(pr, lo, stop, (uint32_t) (big + 1), sq) : 0;
    return 0;
  }
  catch (const std::exception& e)
  {
    setLastError(e.what());
    return -1;
  }
}

} // extern "C"
