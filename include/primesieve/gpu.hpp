///
/// @file  gpu.hpp
/// @brief GPU count API of the MI355X (gfx950) HIP backend.
///
///        The first block keeps the reference's public GPU API verbatim
///        (include/primesieve/gpu.hpp:32-81): same names, argument meaning
///        and error behaviour.  The gpu_count_*() functions throw
///        primesieve_error when no GPU is usable; gpu_try_count() never
///        throws and returns false so that the caller can sieve on the CPU.
///
///        The second block is new: a single-pass six-count call, the batched
///        interval API and the measurement hooks used by the benchmarks.
///
#ifndef PRIMESIEVE_GPU_HPP
#define PRIMESIEVE_GPU_HPP

#include <stdint.h>
#include <string>
#include <vector>

namespace primesieve {

// ---- reference API (include/primesieve/gpu.hpp:32-81) -------------------

bool gpu_available();
std::vector<std::string> gpu_devices();
void set_gpu_device(int deviceIndex);
int  get_gpu_device();
void set_gpu_enabled(bool enable);
bool get_gpu_enabled();
void set_gpu_sieve_size(int kilobytes);
void set_gpu_work_group_size(int size);
void set_gpu_segments_per_launch(int segments);
void set_gpu_presieve(bool enable);
std::string gpu_active_device();
std::string gpu_last_error();
bool gpu_worthwhile(uint64_t start, uint64_t stop);
uint64_t gpu_count_primes(uint64_t start, uint64_t stop);
uint64_t gpu_count_twins(uint64_t start, uint64_t stop);
uint64_t gpu_count_triplets(uint64_t start, uint64_t stop);
uint64_t gpu_count_quadruplets(uint64_t start, uint64_t stop);
uint64_t gpu_count_quintuplets(uint64_t start, uint64_t stop);
uint64_t gpu_count_sextuplets(uint64_t start, uint64_t stop);
bool gpu_try_count(uint64_t start, uint64_t stop, int countType, uint64_t& result);

// ---- additions of the HIP backend ---------------------------------------

/// All six counts (primes, twins, ..., sextuplets) of [start, stop] from one
/// sieve pass.  Throws like gpu_count_primes().
void gpu_count_all(uint64_t start, uint64_t stop, uint64_t counts[6]);

/// One query of a batch: count type k = 1..6 (1 = primes, 2 = twins, ...,
/// 6 = sextuplets) over [start, stop].
struct GpuQuery
{
  uint64_t start;
  uint64_t stop;
  uint32_t k;
};

/// Count every query of a batch in one persistent launch series.  out[i]
/// receives the count of queries[i].  Throws like gpu_count_primes().
void gpu_count_batch(const std::vector<GpuQuery>& queries, std::vector<uint64_t>& out);

} // namespace primesieve

#endif
