/*
 * kernel_abi.h -- argument blocks shared by the gfx950 code object and the
 * host library.  Compiled by both hipcc (device) and g++ (host), so only
 * fixed-width integers and no pointers: device addresses travel as uint64_t.
 *
 * Sieve layout (the reference's wheel-30 bit array, src/gpu/GpuWheel.hpp:125,
 * doc/GPU.md:84-88): bit b of global byte g stands for the integer
 * 30*g + {7,11,13,17,19,23,29,31}[b].  A segment is `segBytes` consecutive
 * global bytes held in LDS as 32-bit little-endian words.
 */
#ifndef PSG_KERNEL_ABI_H
#define PSG_KERNEL_ABI_H

#include <stdint.h>

#define PSG_NCOUNTS 6          /* primes, twins, triplets, ..., sextuplets */
#define PSG_MAX_SEG_BYTES (160u * 1024u)
#define PSG_TILE0_PERIOD 17017u  /* 7*11*13*17 bytes */
#define PSG_TILE1_PERIOD 437u    /* 19*23 bytes */
#define PSG_FIRST_CROSS_PRIME 29u

/* One contiguous byte range [gLo, gHi] counted by one launch series. */
typedef struct PsgRange {
  uint64_t B0;          /* global byte index of segment 0 */
  uint64_t gLo;         /* first counted global byte */
  uint64_t gHi;         /* last counted global byte */
  uint32_t loMask;      /* bits of byte gLo that are >= lo */
  uint32_t hiMask;      /* bits of byte gHi that are <= hi */
  uint32_t nprimes;     /* sieving primes (>= 29) with p*p <= hi */
  uint32_t pad0;
} PsgRange;

/* Arguments of psg_sieve_count: sieve segments B0 + s*segBytes for
 * s in [segBegin, segBegin + gridDim.x) and add the six counts. */
typedef struct PsgSieveArgs {
  PsgRange range;
  uint64_t primes;      /* const uint32_t*: sieving primes >= 29, ascending */
  uint64_t magics;      /* const uint64_t*: ceil(2^(64+L)/p), L = floor(log2 p) */
  uint64_t tile0;       /* const uint32_t*: sliding windows, PSG_TILE0_PERIOD entries */
  uint64_t tile1;       /* const uint32_t*: sliding windows, PSG_TILE1_PERIOD entries */
  uint64_t counts;      /* uint64_t[PSG_NCOUNTS], accumulated atomically */
  uint64_t buckets;     /* const uint32_t*: big-prime bucket entries (0 = none) */
  uint64_t bucketOffs;  /* const uint32_t*: per-segment entry offsets within the slab */
  uint32_t segBytes;    /* multiple of 256, <= PSG_MAX_SEG_BYTES */
  uint32_t segBegin;    /* first segment index of this launch */
  uint32_t tier1End;    /* primes[0, tier1End): p <= segBytes/8 */
  uint32_t tier2End;    /* primes[tier1End, tier2End): p <= segBytes */
  uint32_t tier3End;    /* primes[tier2End, tier3End): per-segment big primes */
  uint32_t slabSegBegin;/* first segment index covered by the bucket slab */
} PsgSieveArgs;

#endif
