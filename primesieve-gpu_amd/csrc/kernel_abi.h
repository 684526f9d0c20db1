/*
 * kernel_abi.h -- argument blocks shared by the gfx950 code object and the
 * host library.  Compiled by both hipcc (device) and g++ (host), so only
 * fixed-width integers and no pointers: device addresses travel as uint64_t.
 *
 * Sieve layout (the reference's wheel-30 bit array, src/gpu/GpuWheel.hpp:125,
 * doc/GPU.md:84-88): bit b of global byte g stands for the integer
 * 30*g + {7,11,13,17,19,23,29,31}[b].  A segment is `segBytes` consecutive
 * global bytes held in LDS as 32-bit little-endian words.
 *
 * Work is described by runs: a run is a stretch of consecutive segments of
 * one interval (a whole query, or a piece of a long interval).  A slab is a
 * list of runs whose segments are numbered 0 .. nSegs-1 in run order; one
 * slab is sieved by one persistent launch of psg_sieve_count, after its
 * big-prime bucket lists have been built.
 */
#ifndef PSG_KERNEL_ABI_H
#define PSG_KERNEL_ABI_H

#include <stdint.h>

#define PSG_NCOUNTS 6          /* primes, twins, triplets, ..., sextuplets */
#define PSG_MAX_SEG_BYTES (160u * 1024u)
#define PSG_SIEVE_LDS_TAIL 3800u /* LDS after the segment: class table, counters, wave totals,
                                    tier-0 primes and magics, class-row table,
                                    cofactor-ordered class table */
/* Pre-sieve tiles: the bit pattern of each group of small primes repeats
 * every P = product-of-the-group bytes.  A tile holds the P sliding 32-bit
 * windows of its pattern (window i = pattern bytes i..i+3, wrapping; a set
 * bit marks a multiple of a tile prime, as in the kernel's sieve) in
 * residue-major order: 4 rows of L = ceil(P/4) + 3 words, word j of row r
 * holding window (r + 4j) mod P.  Window i < P sits at row i % 4, word i / 4,
 * and the 4 windows i, i+4, i+8, i+12 that 4 consecutive sieve words need
 * follow it in the same row (each row runs 3 windows past the period).
 * Tiles are stored one after another; tiles 1.. (PSG_SMALL_TILE_WORDS
 * words) are copied into LDS when the segment leaves room. */
#define PSG_NTILES 5
#define PSG_TILE_PERIODS { 17017u, 437u, 899u, 1517u, 2021u }   /* 7*11*13*17, 19*23, 29*31, 37*41, 43*47 */
#define PSG_TILE_ROWS { 4258u, 113u, 228u, 383u, 509u }         /* L = ceil(P / 4) + 3 */
#define PSG_TILE_OFFSETS { 0u, 17032u, 17484u, 18396u, 19928u } /* first word of each tile */
#define PSG_TILE_WORDS 21964u
#define PSG_SMALL_TILE_WORDS (PSG_TILE_WORDS - 17032u)
#define PSG_FIRST_CROSS_PRIME 53u
#define PSG_TIER0_MAX_PRIME 256u
#define PSG_TIER1_MAX_PRIME 2048u
#define PSG_TIER0_SLOTS 64u      /* LDS slots for the tier-0 primes (>= the 51 primes 7..251) */

/* One run (48 bytes). */
typedef struct PsgRun {
  uint64_t B0;          /* global byte of the run's first segment */
  uint64_t gLo;         /* first counted global byte of the interval */
  uint64_t gHi;         /* last counted global byte of the interval */
  uint32_t loMask;      /* bits of byte gLo whose values are >= lo */
  uint32_t hiMask;      /* bits of byte gHi whose values are <= hi */
  uint32_t nprimes;     /* sieving primes (from primes[0]) with p*p <= hi */
  uint32_t out;         /* output slot: counts go to out * PSG_NCOUNTS + type */
  uint32_t segStart;    /* first slab segment of this run */
  uint32_t nseg;        /* segments in this run */
} PsgRun;

/* Arguments of psg_sieve_count (persistent: workgroups take slab segments
 * from *workCounter until nSegs is reached).  Like every argument block
 * here it lives in device memory, uploaded with the call's run and window
 * tables; the kernel's only parameter is its address.  A launch series thus
 * depends on the plan only through the number of slabs, and one captured
 * series replays for every plan of that shape. */
typedef struct PsgSieveArgs {
  uint64_t runs;        /* const PsgRun*, sorted by segStart */
  uint64_t primes;      /* const uint32_t*: sieving primes, ascending */
  uint64_t magics;      /* const uint64_t*: ceil(2^(64+L)/p), L = floor(log2 p) */
  uint64_t tiles;       /* const uint32_t[PSG_TILE_WORDS]: pre-sieve tiles (0 = no pre-sieve) */
  uint64_t tilesInLds;  /* 1: the kernel keeps tiles 1.. in LDS after the segment and its tail */
  uint64_t counts;      /* uint64_t[outs][PSG_NCOUNTS], accumulated atomically */
  uint64_t buckets;     /* const uint32_t*: bucket entries of the slab (0 = none) */
  uint64_t listOff;     /* const uint32_t[nSegs + 1]: segment s owns entries [listOff[s], listOff[s+1]) */
  uint64_t listCount;   /* const uint32_t[nSegs]: entries appended to segment s (may exceed its room) */
  uint64_t workCounter; /* uint32_t*, zero at launch */
  uint64_t primesOut;   /* list mode (0 = count mode): uint32_t*; every run is one segment,
                           and its primes are written ascending from primesOut[outOffsets[run.out]] */
  uint64_t outOffsets;  /* const uint32_t[outs], list mode */
  uint32_t nRuns;
  uint32_t nSegs;       /* segments in the slab */
  uint32_t segBytes;    /* multiple of 256, <= PSG_MAX_SEG_BYTES */
  uint32_t tier0End;    /* primes[0, tier0End): p <= 256, struck per wave slice */
  uint32_t tier1End;    /* primes[tier0End, tier1End): p <= min(PSG_TIER1_MAX_PRIME, segBytes) */
  uint32_t tier2End;    /* primes[tier1End, tier2End): p <= segBytes */
  uint32_t tier3End;    /* primes[tier2End, tier3End): per-segment big primes;
                           larger primes come from the bucket lists */
  uint32_t pad;
} PsgSieveArgs;

/* Big-prime bucket tier.  A bucket entry is the bit index of one crossing
 * inside its segment, byte * 8 + bit (word = entry >> 5, bit = entry & 31),
 * appended to the list of that segment.  Lists have a fixed room each
 * (listOff, sized from the expected number of crossings plus a margin) and
 * a counter (listCount) that psg_bucket_fill advances atomically; a counter
 * that ends above its room raises *overflow and the host repeats the call
 * with rooms equal to the counted totals. */
#define PSG_BUCKET_THREADS 256u      /* threads of a psg_bucket_fill workgroup */
#define PSG_BUCKET_HITS 32u          /* hits a fill thread stages per stage */
#define PSG_BUCKET_STAGE (PSG_BUCKET_THREADS * PSG_BUCKET_HITS)
#define PSG_MAX_WINDOW_SEGS PSG_BUCKET_THREADS   /* segments per bucket window */
/* The fill's per-segment histogram has several copies, interleaved per
 * segment (word seg * C + copy, copy = thread mod C), so that lanes counting
 * hits of one segment rarely hit the same LDS word: C = 4 for windows of up
 * to 128 segments, 2 up to 256. */
#define PSG_BUCKET_HIST_WORDS (2u * PSG_MAX_WINDOW_SEGS)
/* dynamic LDS of psg_bucket_fill: stage, histogram copies, per-segment placement (u64) */
#define PSG_BUCKET_LDS ((PSG_BUCKET_STAGE + PSG_BUCKET_HIST_WORDS + 2u * PSG_MAX_WINDOW_SEGS) * 4u)

/* One bucket window (32 bytes): consecutive segments of one run walked by
 * nChunks workgroups, each over an interleaved share of the big primes
 * primes[bigBegin, bigEnd). */
typedef struct PsgWindow {
  uint64_t base;        /* global byte of the window's first byte */
  uint32_t firstSeg;    /* slab segment of the window's first segment */
  uint32_t segs;        /* segments in the window, <= PSG_MAX_WINDOW_SEGS */
  uint32_t bigEnd;      /* exclusive prime index bound of its interval */
  uint32_t nChunks;     /* workgroups walking this window */
  uint32_t wgBase;      /* first workgroup of the window in the fill grid */
  uint32_t pad;
} PsgWindow;

/* Arguments of psg_bucket_fill: the work items are the (window, chunk)
 * pairs, numbered 0 .. nItems-1 (window w owns items wgBase .. wgBase +
 * nChunks - 1); the workgroups of a persistent grid take them from
 * *workCounter. */
typedef struct PsgBucketArgs {
  uint64_t windows;     /* const PsgWindow*, sorted by wgBase */
  uint64_t primes;      /* const uint32_t*; every bucket prime is >= 2^16 */
  uint64_t listOff;     /* const uint32_t[nSegs + 1] */
  uint64_t listCount;   /* uint32_t[nSegs], zero at launch */
  uint64_t entries;     /* uint32_t* */
  uint64_t overflow;    /* uint32_t flag, set when a list outgrows its room */
  uint64_t total;       /* unsigned long long: entries appended (statistics) */
  uint64_t workCounter; /* uint32_t*, zero at launch */
  uint32_t nWindows;
  uint32_t nItems;      /* (window, chunk) work items; 0 = nothing to fill */
  uint32_t segShift;    /* log2(segBytes) */
  uint32_t bigBegin;
} PsgBucketArgs;

/* Device memory that psg_clear zeroes before a slab series: up to four
 * ranges of 32-bit words (counts, work counters, list counts, flags). */
#define PSG_CLEAR_RANGES 4
typedef struct PsgClearArgs {
  uint64_t addr[PSG_CLEAR_RANGES];
  uint32_t words[PSG_CLEAR_RANGES];
} PsgClearArgs;

/* Layout of a call's descriptor upload: the clear block, then one 256-byte
 * record per slab (its PsgSieveArgs, then its PsgBucketArgs), then the run,
 * window and list-offset tables the records point into. */
#define PSG_DESC_HEADER 64u
#define PSG_SLAB_RECORD 256u
#define PSG_SLAB_FILL_OFFSET 128u

/* The static scratch array of the code object (psg_scratch): the device
 * memory of count calls that must not allocate (descriptors, counters,
 * bucket lists). */
#define PSG_SCRATCH_BYTES (64u << 20)

#endif
