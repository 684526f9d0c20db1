/*
 * ref_count6.cpp -- test and benchmark infrastructure, built only into
 * oracle/_ref/ together with the reference CPU library (never into the
 * product).  One C entry point over the reference's multithreaded CPU sieve
 * that counts primes and all five k-tuplet types in ONE pass, as
 * `primesieve START --dist=N --count=123456 --threads=T` does: a
 * ParallelSieve (src/ParallelSieve.hpp:25-36) run with the six COUNT_*
 * flags (src/PrimeSieveClass.hpp:30-35), counts read with getCount(i).
 */
#include "ParallelSieve.hpp"

#include <stdint.h>

extern "C" void psg_ref_count6(uint64_t start, uint64_t stop, uint64_t out[6], int threads)
{
  primesieve::ParallelSieve ps;
  if (threads > 0)
    ps.setNumThreads(threads);
  ps.sieve(start, stop, primesieve::COUNT_PRIMES | primesieve::COUNT_TWINS |
                        primesieve::COUNT_TRIPLETS | primesieve::COUNT_QUADRUPLETS |
                        primesieve::COUNT_QUINTUPLETS | primesieve::COUNT_SEXTUPLETS);
  for (int i = 0; i < 6; i++)
    out[i] = ps.getCount(i);
}
