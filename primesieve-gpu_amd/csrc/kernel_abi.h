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
  uint32_t tier1End;    /* primes[tier0End, tier1End): p <= segBytes/8 */
  uint32_t tier2End;    /* primes[tier1End, tier2End): p <= segBytes */
  uint32_t tier3End;    /* primes[tier2End, tier3End): per-segment big primes */
  uint32_t slabSegBegin;/* first segment index covered by the bucket slab */
  uint32_t entryBytes;  /* 2: uint16 byte offsets (segBytes <= 65536); 4: uint32 */
  uint32_t capacity;    /* bucket entries allocated */
  uint32_t tier0End;    /* primes[0, tier0End): p <= 256, struck per wave slice */
} PsgSieveArgs;

/* Big-prime bucket tier.  A slab is `slabBytes / segBytes` consecutive
 * segments starting at global byte slabB.  Every prime in
 * primes[bigBegin, bigEnd) is walked over the whole slab once per phase;
 * its hits go to list (segment * 8 + class), class = bit index.  Lists are
 * laid out by a counting sort with no global atomics:
 *   psg_bucket_count  per-workgroup LDS histogram -> counts[chunk][list]
 *   psg_bucket_scan_chunks  exclusive prefix over chunks per list, totals
 *   psg_bucket_scan_lists   exclusive prefix over lists -> listOff
 *   psg_bucket_fill   re-walk, LDS cursors, store byte offsets
 * The sieve kernel then strikes list (s*8 + j) of segment s with bit j. */
typedef struct PsgBucketArgs {
  uint64_t slabB;       /* global byte index of the slab's first byte */
  uint64_t primes;      /* const uint32_t* */
  uint64_t magics;      /* const uint64_t* */
  uint64_t counts;      /* uint32_t[nChunks][nLists] */
  uint64_t listOff;     /* uint32_t[nLists + 1] */
  uint64_t entries;     /* uint16_t* or uint32_t* (entryBytes) */
  uint64_t overflow;    /* uint32_t flag, set when capacity is exceeded */
  uint32_t slabBytes;   /* bytes covered, multiple of segBytes, < 2^31 */
  uint32_t segShift;    /* log2(segBytes) */
  uint32_t nLists;      /* 8 * segments in the slab */
  uint32_t bigBegin;
  uint32_t bigEnd;
  uint32_t nChunks;     /* prime chunks (primes interleaved over chunks) */
  uint32_t capacity;    /* entries allocated */
  uint32_t entryBytes;  /* 2 or 4 */
  uint32_t windowSegs;  /* segments per window: workgroup (w, c) walks the
                           primes of chunk c over window w only, so that its
                           sub-lists are contiguous and its LDS counters few */
  uint32_t nWindows;    /* ceil(segments in slab / windowSegs) */
} PsgBucketArgs;

#endif
