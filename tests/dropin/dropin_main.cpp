// Driver of the drop-in check (tests/test_dropin.py): a program built from
// the reference's own count API (src/api.cpp, src/api-c.cpp,
// src/gpu/GpuSieve.cpp) and this repository's HIP backend.  It enables the
// GPU, counts through the reference's C API -- which tries gpu_try_count
// first -- and reports what the GPU layer said, one "key value" per line.
#include <primesieve.h>
#include <primesieve/gpu.hpp>

#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

int main(int argc, char** argv)
{
  if (argc < 3)
    return 2;
  const uint64_t a = std::strtoull(argv[1], nullptr, 10);
  const uint64_t b = std::strtoull(argv[2], nullptr, 10);
  primesieve::set_gpu_enabled(true);
  std::printf("available %d\n", primesieve::gpu_available() ? 1 : 0);
  std::printf("devices %zu\n", primesieve::gpu_devices().size());
  std::printf("primes %llu\n", (unsigned long long) primesieve_count_primes(a, b));
  std::printf("twins %llu\n", (unsigned long long) primesieve_count_twins(a, b));
  std::string err = primesieve::gpu_last_error();
  for (char& c : err)
    if (c == '\n') c = ' ';
  std::printf("last_error %s\n", err.c_str());
  try
  {
    std::printf("gpu_count %llu\n", (unsigned long long) primesieve::gpu_count_primes(a, b));
  }
  catch (const std::exception& e)
  {
    std::printf("gpu_count_threw %s\n", e.what());
  }
  return 0;
}
