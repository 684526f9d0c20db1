/*
 * psg_oracle.c -- CPU oracle for the wheel-30 prime / k-tuplet count path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only tests/, the smoke check and the CPU-baseline leg of
 * bench.py use it, and only as the checker.  It is a plain, deliberately
 * simple restatement of the reference CPU sieve's semantics:
 *
 *   - byte layout: bit b of byte g stands for 30*g + {7,11,13,17,19,23,29,31}[b]
 *     (src/gpu/GpuWheel.hpp:125, src/Erat.cpp:30-47);
 *   - values == 1 (mod 30) live in the previous byte: byteRemainder(n) =
 *     (n - 7) % 30 + 7 (src/Erat.cpp:231-237);
 *   - bits below start / above stop are cleared in the first / last byte
 *     (src/Erat.cpp:255-291, tables unsetSmaller/unsetLarger :30-47);
 *   - sieving primes p >= 7 strike p*q for q >= p, gcd(q, 30) = 1
 *     (src/Wheel.hpp:50-89);
 *   - primes: popcount; k-tuplets: number of the reference bitmasks contained
 *     in each byte (src/CountPrintPrimes.cpp:33-41, 114-184);
 *   - 2, 3, 5 and the tuplets containing them come from a fixed table applied
 *     when start <= 5 (src/PrimeSieveClass.cpp:38-48, 253-265);
 *   - multi-threaded chunks end at n == 1 (mod 30) so that no byte (and hence
 *     no k-tuplet) is split (src/ParallelSieve.cpp:122-133).
 *
 * It also restates divisionMagic (src/gpu/GpuTables.cpp:170-200), the per-byte
 * k-tuplet table (src/gpu/GpuTables.cpp:132-168) and gives an independent
 * brute-force counter (trial division + explicit k-tuplet patterns, the role
 * tools/ref_sieve.cpp:97-116 plays in the reference) and a closed-form
 * crossing counter used for the algorithmic-work figures.
 *
 * Pinning (each function is checked by tests/test_oracle.py):
 *   psg_oracle_count6      golden KATs from the reference test suite, the
 *                          reference CPU library (oracle/_ref) and brute force
 *   psg_oracle_brute_count π(x) closed values, the explicit small tuplets
 *   psg_oracle_ktable      brute-force enumeration of k-tuplet patterns
 *   psg_oracle_division_magic  128-bit floor division
 *   psg_oracle_crossings   brute-force enumeration of crossings
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

static const uint32_t kOff[8] = {7, 11, 13, 17, 19, 23, 29, 31};

/* bit index of value v with (v - 30*g) == off, off in 7..31; -1 if not coprime */
static int bit_of_offset(uint32_t off)
{
  for (int b = 0; b < 8; b++)
    if (kOff[b] == off)
      return b;
  return -1;
}

uint64_t psg_oracle_isqrt(uint64_t n)
{
  uint64_t r = 0;
  /* bitwise integer square root: exact for all 64-bit n */
  uint64_t bit = (uint64_t) 1 << 62;
  uint64_t x = n;
  while (bit > x) bit >>= 2;
  while (bit) {
    if (x >= r + bit) { x -= r + bit; r = (r >> 1) + bit; }
    else r >>= 1;
    bit >>= 2;
  }
  return r;
}

/* All primes in [2, limit] (simple sieve); returns count, fills out if non-null. */
size_t psg_oracle_small_primes(uint32_t limit, uint32_t* out)
{
  if (limit < 2) return 0;
  uint8_t* comp = (uint8_t*) calloc((size_t) limit + 1, 1);
  size_t n = 0;
  for (uint64_t i = 2; i * i <= limit; i++)
    if (!comp[i])
      for (uint64_t j = i * i; j <= limit; j += i)
        comp[j] = 1;
  for (uint64_t i = 2; i <= limit; i++)
    if (!comp[i]) { if (out) out[n] = (uint32_t) i; n++; }
  free(comp);
  return n;
}

/* Reference bitmasks (src/CountPrintPrimes.cpp:33-41). */
static const uint8_t kMasks[6][4] = {
  {0},
  {0x06, 0x18, 0xc0, 0},
  {0x07, 0x0e, 0x1c, 0x38},
  {0x1e, 0, 0, 0},
  {0x1f, 0x3e, 0, 0},
  {0x3f, 0, 0, 0},
};

void psg_oracle_ktable(int type, uint8_t out[256])
{
  for (int j = 0; j < 256; j++) {
    int c = 0;
    if (type == 0) {
      for (int b = 0; b < 8; b++) c += (j >> b) & 1;
    } else {
      for (int m = 0; m < 4; m++)
        if (kMasks[type][m] && (j & kMasks[type][m]) == kMasks[type][m]) c++;
    }
    out[j] = (uint8_t) c;
  }
}

/* M = ceil(2^(64+L) / p), L = floor(log2 p)  (src/gpu/GpuTables.cpp:170-200) */
uint64_t psg_oracle_division_magic(uint32_t p)
{
  uint32_t L = 31 - (uint32_t) __builtin_clz(p);
  unsigned __int128 num = (unsigned __int128) 1 << (64 + L);
  unsigned __int128 q = num / p;
  if (q * p != num) q++;
  return (uint64_t) q;
}

/* ---------------------------------------------------------------------- */
/* Segmented sieve of one chunk [lo, hi], lo >= 7, byte-aligned chunk ends. */

typedef struct {
  const uint32_t* primes;   /* primes >= 7 up to isqrt(hi) */
  size_t nprimes;
  uint64_t lo, hi;
  uint64_t counts[6];
} chunk_job;

static uint64_t byte_remainder(uint64_t n) { return (n - 7) % 30 + 7; }

static void sieve_chunk(chunk_job* job)
{
  const uint64_t lo = job->lo, hi = job->hi;
  uint8_t ktab[6][256];
  for (int t = 0; t < 6; t++) psg_oracle_ktable(t, ktab[t]);
  memset(job->counts, 0, sizeof job->counts);
  if (lo > hi) return;

  const uint64_t gFirst = (lo - byte_remainder(lo)) / 30;
  const uint64_t gLast = (hi - byte_remainder(hi)) / 30;
  const uint64_t segBytes = 1 << 16;
  uint8_t* seg = (uint8_t*) malloc(segBytes);

  for (uint64_t g0 = gFirst; g0 <= gLast; g0 += segBytes) {
    uint64_t n = gLast - g0 + 1;
    if (n > segBytes) n = segBytes;
    memset(seg, 0xff, n);
    /* value range of this segment: [30*g0 + 7, 30*(g0+n-1) + 31] */
    const uint64_t vmin = 30 * g0 + 7;
    const uint64_t top = 30 * (g0 + n - 1);
    const uint64_t vmax = (top > UINT64_MAX - 31) ? UINT64_MAX : top + 31;

    for (size_t i = 0; i < job->nprimes; i++) {
      const uint64_t p = job->primes[i];
      if (p * p > vmax) break;
      uint64_t q = vmin / p + (vmin % p != 0);
      if (q < p) q = p;
      const uint64_t qmax = vmax / p;
      /* walk q upwards over the residues coprime to 30 */
      for (; q <= qmax; q++) {
        if (q % 2 == 0 || q % 3 == 0 || q % 5 == 0) continue;
        const uint64_t v = p * q;
        const uint64_t g = (v - 7) / 30;
        const int b = bit_of_offset((uint32_t) (v - 30 * g));
        seg[g - g0] &= (uint8_t) ~(1u << b);
      }
    }

    /* clear bits outside [lo, hi] */
    for (uint64_t i = 0; i < n; i++) {
      if (i > 0 && i + 1 < n) continue;
      for (int b = 0; b < 8; b++) {
        const uint64_t v = 30 * (g0 + i) + kOff[b];
        if (v < lo || v > hi) seg[i] &= (uint8_t) ~(1u << b);
      }
    }

    for (uint64_t i = 0; i < n; i++)
      for (int t = 0; t < 6; t++)
        job->counts[t] += ktab[t][seg[i]];
  }
  free(seg);
}

static void* chunk_thread(void* arg)
{
  sieve_chunk((chunk_job*) arg);
  return NULL;
}

/* Counts of primes, twins, triplets, quadruplets, quintuplets and sextuplets
 * in [start, stop], with the reference's semantics.  threads <= 0 means 1. */
void psg_oracle_count6(uint64_t start, uint64_t stop, uint64_t out[6], int threads)
{
  static const struct { uint64_t first, last; int index; } small[8] = {
    {2, 2, 0}, {3, 3, 0}, {5, 5, 0}, {3, 5, 1},
    {5, 7, 1}, {5, 11, 2}, {5, 13, 3}, {5, 17, 4},
  };
  memset(out, 0, 6 * sizeof(uint64_t));
  if (start > stop) return;
  if (start <= 5)
    for (int i = 0; i < 8; i++)
      if (small[i].first >= start && small[i].last <= stop)
        out[small[i].index]++;
  if (stop < 7) return;

  uint64_t lo = start < 7 ? 7 : start;
  uint64_t hi = stop;
  uint32_t sq = (uint32_t) psg_oracle_isqrt(hi);
  size_t np = psg_oracle_small_primes(sq, NULL);
  uint32_t* all = (uint32_t*) malloc((np + 1) * sizeof(uint32_t));
  psg_oracle_small_primes(sq, all);
  size_t first = 0;
  while (first < np && all[first] < 7) first++;

  if (threads <= 0) threads = 1;
  if (threads > 256) threads = 256;
  uint64_t dist = hi - lo;
  if (dist < (uint64_t) threads * 3000000ull) threads = 1;

  chunk_job jobs[256];
  pthread_t tids[256];
  /* Interior chunk ends are == 1 (mod 30), the last value of a byte, so no
   * byte is split between two chunks.  An empty chunk has lo > hi. */
  uint64_t a = lo;
  for (int t = 0; t < threads; t++) {
    uint64_t b = hi;
    if (t != threads - 1) {
      b = lo + (dist / (uint64_t) threads) * (uint64_t) (t + 1);
      b = b - (b % 30) + 1;
      if (b > hi) b = hi;
    }
    jobs[t].primes = all + first;
    jobs[t].nprimes = np - first;
    jobs[t].lo = a;
    jobs[t].hi = b;
    if (b >= a) a = b + 1;
  }
  for (int t = 0; t < threads; t++)
    pthread_create(&tids[t], NULL, chunk_thread, &jobs[t]);
  for (int t = 0; t < threads; t++) {
    pthread_join(tids[t], NULL);
    for (int k = 0; k < 6; k++) out[k] += jobs[t].counts[k];
  }
  free(all);
}

/* ---------------------------------------------------------------------- */
/* Independent brute force: explicit primality (trial division by a small
 * prime table) and explicit k-tuplet patterns.  For small ranges only. */

static int is_prime_td(uint64_t n, const uint32_t* pr, size_t np)
{
  if (n < 2) return 0;
  for (size_t i = 0; i < np; i++) {
    uint64_t p = pr[i];
    if (p * p > n) return 1;
    if (n % p == 0) return n == p;
  }
  return 1;
}

/* Admissible patterns (offsets from the smallest member). */
static const int kPat[6][2][6] = {
  {{0, -1}, {-1}},
  {{0, 2, -1}, {-1}},
  {{0, 2, 6, -1}, {0, 4, 6, -1}},
  {{0, 2, 6, 8, -1}, {-1}},
  {{0, 2, 6, 8, 12, -1}, {0, 4, 6, 10, 12, -1}},
  {{0, 4, 6, 10, 12, 16}, {-1}},
};

uint64_t psg_oracle_brute_count(uint64_t start, uint64_t stop, int type)
{
  if (start > stop || type < 0 || type > 5) return 0;
  uint64_t ext = stop + 16;
  uint32_t lim = (uint32_t) psg_oracle_isqrt(ext) + 1;
  size_t np = psg_oracle_small_primes(lim, NULL);
  uint32_t* pr = (uint32_t*) malloc((np + 1) * sizeof(uint32_t));
  psg_oracle_small_primes(lim, pr);
  uint64_t c = 0;
  for (uint64_t p = start; p <= stop; p++) {
    if (!is_prime_td(p, pr, np)) { if (p == UINT64_MAX) break; continue; }
    for (int v = 0; v < 2; v++) {
      const int* pat = kPat[type][v];
      if (pat[0] < 0) continue;
      int ok = 1;
      for (int m = 0; m < 6 && pat[m] >= 0; m++) {
        uint64_t x = p + (uint64_t) pat[m];
        if (x > stop || !is_prime_td(x, pr, np)) { ok = 0; break; }
      }
      c += (uint64_t) ok;
    }
    if (p == UINT64_MAX) break;
  }
  free(pr);
  return c;
}

/* ---------------------------------------------------------------------- */
/* Closed-form number of crossings: sum over sieving primes p in [pmin, pmax]
 * (p >= 7) of #{q : max(p, ceil(lo/p)) <= q <= floor(hi/p), gcd(q, 30) = 1}.
 * cop(n) = #{1 <= q <= n : gcd(q,30)=1} = 8*floor(n/30) + R[n mod 30]. */

static uint64_t cop(uint64_t n)
{
  static const uint8_t R[30] = {0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 4, 4, 4,
                                4, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 7, 7, 8};
  return 8 * (n / 30) + R[n % 30];
}

uint64_t psg_oracle_crossings(uint64_t lo, uint64_t hi, uint32_t pmin, uint32_t pmax)
{
  if (lo > hi) return 0;
  uint32_t sq = (uint32_t) psg_oracle_isqrt(hi);
  if (pmax > sq) pmax = sq;
  if (pmin < 7) pmin = 7;
  if (pmax < pmin) return 0;
  size_t np = psg_oracle_small_primes(pmax, NULL);
  uint32_t* pr = (uint32_t*) malloc((np + 1) * sizeof(uint32_t));
  psg_oracle_small_primes(pmax, pr);
  uint64_t c = 0;
  for (size_t i = 0; i < np; i++) {
    uint64_t p = pr[i];
    if (p < pmin) continue;
    uint64_t qa = (lo + p - 1) / p;
    if (qa < p) qa = p;
    uint64_t qb = hi / p;
    if (qb >= qa) c += cop(qb) - cop(qa - 1);
  }
  free(pr);
  return c;
}

/* Brute-force crossing counter for small ranges (pins psg_oracle_crossings). */
uint64_t psg_oracle_crossings_brute(uint64_t lo, uint64_t hi, uint32_t pmin, uint32_t pmax)
{
  if (lo > hi) return 0;
  uint32_t sq = (uint32_t) psg_oracle_isqrt(hi);
  if (pmax > sq) pmax = sq;
  if (pmin < 7) pmin = 7;
  uint64_t c = 0;
  size_t np = psg_oracle_small_primes(pmax, NULL);
  uint32_t* pr = (uint32_t*) malloc((np + 1) * sizeof(uint32_t));
  psg_oracle_small_primes(pmax, pr);
  for (size_t i = 0; i < np; i++) {
    uint64_t p = pr[i];
    if (p < pmin) continue;
    for (uint64_t v = lo; v <= hi; v++)
      if (v % p == 0 && v / p >= p && v % 2 && v % 3 && v % 5) c++;
  }
  free(pr);
  return c;
}
