// psg_sieve.hip -- gfx950 segment kernel of the wheel-30 prime / k-tuplet
// counter.  Built with `hipcc --genco --offload-arch=gfx950` into a code
// object that the host library loads with hipModuleLoadData.
//
// One workgroup sieves one segment of `segBytes` global bytes held in LDS.
// Bit b of global byte g stands for 30*g + {7,11,13,17,19,23,29,31}[b]
// (the reference layout, src/gpu/GpuWheel.hpp:125).  The work of the
// reference kernel ps_count (src/gpu/cuda/sieve.cu:157-320) is reorganised
// for CDNA4:
//
//   1. fill:   five pre-sieve tiles remove the multiples of 7..47 (groups
//              {7,11,13,17} {19,23} {29,31} {37,41} {43,47}; each pattern
//              repeats every product-of-the-group bytes and is read as
//              sliding 32-bit windows), then the tile primes' own bits are
//              put back;
//   2. strike: every sieving prime p >= 53 is split into its 8 residue
//              classes.  In class j the multiples p*q (q coprime to 30) lie at
//              global bytes c_j + t*p, so a class is an arithmetic progression
//              with byte stride p and a constant bit -- no wheel state
//              machine and no per-crossing table reads.  Tiers:
//                tier 0 (p <= 256): each wave strikes all of them over its
//                        own slice of the segment (even load); lane =
//                        (class, hit), stride 8p, loop-invariant word mask;
//                tier 1 (256 < p <= 2048): lane = (prime, class), 8 primes
//                        per work item; each class as four progressions of
//                        stride 4p with loop-invariant masks;
//                tier 2 (2048 < p <= S): 2 primes per lane, 128 per item,
//                        the same progressions for each of their classes;
//                tier 3 (S < p <= bucket threshold): 4 primes per lane, 256
//                        per item, at most one hit per class;
//              the items of tiers 1-3 form one queue per segment that the
//              waves take from: tiers 1-2 heaviest first, the light tier-3
//              items spread evenly between them;
//              plus the bucket list of the segment, filled by
//              psg_bucket_fill for larger primes (one LDS atomic per entry),
//              whose entries are loaded while the queue's items are struck;
//   3. count:  popcount and byte-lane SWAR masks give all six counts in one
//              pass; wave sums use ds_swizzle butterflies.  The same pass
//              writes the next segment's fill (step 1) into each chunk it
//              has counted.  In list mode the segment's primes are written
//              out instead (the device's sieving-prime cache;
//              psg_prime_magics adds their magics).
//
// In LDS a set bit marks a composite (the counts read the complement), so a
// strike is a ds_or_b32 atomic with a one-bit mask: several lanes may hit
// the same word at once.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

#include "../kernel_abi.h"

namespace {

// Device addresses travel as uint64_t in the argument blocks; these views
// make their loads and stores global (not flat) memory instructions, which
// keeps them off the LDS wait counter.
#define PSG_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const PSG_GLOBAL T* gptr(uint64_t a) { return (const PSG_GLOBAL T*) (uintptr_t) a; }
template <typename T>
__device__ __forceinline__ PSG_GLOBAL T* gptr_rw(uint64_t a) { return (PSG_GLOBAL T*) (uintptr_t) a; }

// An argument block of kernel_abi.h read from device memory at a uniform
// address: constant-address-space loads, so its fields land in scalar
// registers once per workgroup.
#define PSG_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ T load_args(uint64_t addr)
{
  static_assert(sizeof(T) % 4 == 0, "argument blocks are whole 32-bit words");
  union { T v; uint32_t w[sizeof(T) / 4]; } u;
  const PSG_CONST uint32_t* src = (const PSG_CONST uint32_t*) (uintptr_t) addr;
#pragma unroll
  for (uint32_t k = 0; k < sizeof(T) / 4; k++)
    u.w[k] = src[k];
  return u.v;
}

__device__ __constant__ uint32_t kOffsets[8] = {7, 11, 13, 17, 19, 23, 29, 31};
// the residues coprime to 30, ascending, and as a bit set (bit q for q = 0..29)
__device__ __constant__ uint32_t kCoprimes[8] = {1, 7, 11, 13, 17, 19, 23, 29};
constexpr uint32_t kCoprime30 = 0x208A2882u;
// Slots per row of the cofactor-ordered class table: the 8 coprime residues,
// then the first 5 again.  13 slots (26 words) put the rows of the 8 residues
// used at 8 different LDS banks (26 (b >> 1) mod 32), so lanes reading rows of
// different residues rarely meet in one bank.
constexpr uint32_t kCtabqSlots = 13;
// Words per row of the class-row table (8 entries, 4 words of padding): rows
// indexed by the residue's rank start at 8 different banks (12 r mod 32).
constexpr uint32_t kCtabrStride = 12;

// L = floor(log2 p) of a division magic M = ceil(2^(64+L)/p) (0 for p = 0)
__device__ __forceinline__ uint32_t magic_shift(uint32_t p)
{
  return 31u - (uint32_t) __clz((int) (p | 1u));
}

// R = B mod p for B < 2^63: floor(B/p) = mulhi(B, M) >> L.  Only the low 32
// bits of B - q*p matter.
__device__ __forceinline__ uint32_t mod_magic(uint64_t B, uint32_t p, uint64_t M)
{
  const uint64_t q = __umul64hi(B, M) >> magic_shift(p);
  return (uint32_t) B - (uint32_t) q * p;
}

// The same for p < 2^31 with three of the four partial products of B * M:
// H = B_hi M_hi + hi(B_hi M_lo) + hi(B_lo M_hi) drops B_lo M_lo / 2^64 and
// two fractional parts, each below 1, so it is at most 2 below the exact
// high word.  With L >= 1 the quotient H >> L is then floor(B / p) or one
// less, and B - q p lies in [0, 2p) (below 2^32): one subtraction of p fixes
// it (min(r, r - p): r - p wraps above r when r < p).
__device__ __forceinline__ uint32_t mod_magic31(uint64_t B, uint32_t p, uint64_t M)
{
  const uint32_t Bl = (uint32_t) B, Bh = (uint32_t) (B >> 32);
  const uint32_t Ml = (uint32_t) M, Mh = (uint32_t) (M >> 32);
  const uint64_t H = (uint64_t) Bh * Mh + __umulhi(Bh, Ml) + __umulhi(Bl, Mh);
  const uint32_t r = Bl - (uint32_t) (H >> magic_shift(p)) * p;
  return min(r, r - p);
}

__device__ __forceinline__ uint32_t rotl32(uint32_t x, uint32_t r)
{
  return (x << r) | (x >> (32u - r));
}

// B mod p for p >= 2^16 and B < 2^60 without a division magic.  Bf is the
// double nearest B (off by at most 2^6).  The hardware reciprocal is only an
// estimate, so two Newton steps r(2 - p r) (each squares the relative error)
// bring it to ~2^-51 from any estimate better than ~2^-13; then
// q = floor(Bf * r) is within one of floor(B / p), since
// B / p < 2^44 keeps the product's error far below 1.  The remainder is
// computed exactly in 64-bit integers and corrected by up to two p each way.
__device__ __forceinline__ double to_double(uint64_t B)
{
  return fma((double) (uint32_t) (B >> 32), 0x1p32, (double) (uint32_t) B);   // one rounding
}

// A wave-uniform double moved into scalar registers (VALU instructions take
// a scalar pair as an operand), so that it occupies no vector registers.
__device__ __forceinline__ double uniform_f64(double x)
{
  const uint64_t u = (uint64_t) __double_as_longlong(x);
  const uint32_t lo = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) u);
  const uint32_t hi = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) (u >> 32));
  return __longlong_as_double((long long) (((uint64_t) hi << 32) | lo));
}

__device__ __forceinline__ double recip(uint32_t p)
{
  const double pd = (double) p;
  const double r0 = __builtin_amdgcn_rcp(pd);
  const double r1 = fma(r0, fma(-pd, r0, 1.0), r0);
  return fma(r1, fma(-pd, r1, 1.0), r1);
}

// rp = recip(p)
__device__ __forceinline__ uint32_t mod_big(uint64_t B, double Bf, uint32_t p, double rp)
{
  const double qf = floor(Bf * rp);
  const double qh = floor(qf * 0x1p-32);
  const uint32_t qhi = (uint32_t) qh;
  const uint32_t qlo = (uint32_t) (qf - qh * 0x1p32);      // exact: qf < 2^53
  const uint64_t qp = (uint64_t) qlo * p + (((uint64_t) qhi * p) << 32);
  int64_t r = (int64_t) (B - qp);
  if (r < 0) r += p;
  if (r < 0) r += p;
  if (r >= (int64_t) p) r -= p;
  if (r >= (int64_t) p) r -= p;
  return (uint32_t) r;
}

// Class table entry: ctab[(p % 30) * 8 + j] packs q_j (the cofactor residue
// mod 30 of class j) in bits 0..15 and e_j + 1 in bits 16..31, where
// c_j = (p/30)*q_j + e_j is the global byte of p*q_j, so
// c_j + 1 = (p/30) * (entry & 0xffff) + (entry >> 16).
//
// (c - R) mod p for c, R < p: the unsigned difference wraps above 2^32 - p
// when c < R, and then adding p brings it back below p.
__device__ __forceinline__ uint32_t sub_mod(uint32_t c, uint32_t R, uint32_t p)
{
  const uint32_t u = c - R;
  return min(u, u + p);
}

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// (c_j - R) mod p for p/30 < 2^16, from a1 = (p/30) | 1 << 16 and nR = ~R:
// one unsigned 2-way dot product of the entry's halves with a1's,
// q_j * (p/30) + (e_j + 1) * 1 + ~R = c_j - R (mod 2^32), then the wrap.
__device__ __forceinline__ uint32_t class_offset(uint32_t ent, uint32_t a1, uint32_t nR, uint32_t p)
{
  const uint32_t u = __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, ent),
                                            __builtin_bit_cast(u16x2, a1), nR, false);
  return min(u, u + p);
}

// Start (byte index relative to segment base B) of residue class j of prime
// p, given R = B mod p.  The result is < p, except that the prime itself
// (q_j == 1, first period) is skipped; that can only happen when B < p, so
// callers that know B >= p pass kNear = false.
template <bool kNear = true>
__device__ __forceinline__ uint32_t class_start(uint32_t p, uint32_t R, uint64_t B,
                                                uint32_t j, const uint32_t* ctab)
{
  const uint32_t a = p / 30u;
  const uint32_t b = p - a * 30u;
  const uint32_t ent = ctab[b * 8u + j];
  const uint32_t q = ent & 0xffu;
  uint32_t i = sub_mod(__umul24(a, ent & 0xffffu) + (ent >> 16), R + 1u, p);   // a < 2^24 for every caller
  if (kNear && q == 1u && B + i < (uint64_t) p) {
    const uint64_t ii = (uint64_t) i + p;
    i = (ii > 0xffffffffull) ? 0xffffffffu : (uint32_t) ii;
  }
  return i;
}

// The segment kernel declares no static LDS, so its dynamic LDS -- the sieve
// first -- starts at LDS address 0 (the host checks the kernel's static LDS
// size when it loads the code object).  A sieve byte index rounded down to a
// multiple of 4 is then the LDS address of its 32-bit word, and the strikes
// use it as is: no base address is added per strike.
typedef __attribute__((address_space(3))) uint32_t lds_u32;

__device__ __forceinline__ lds_u32* lds_word(uint32_t byte)
{
  return (lds_u32*) (uintptr_t) (byte & ~3u);
}

__device__ __forceinline__ void lds_or(lds_u32* w, uint32_t mask)
{
  __hip_atomic_fetch_or(w, mask, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// The 8 class entries of p's residue b = p mod 30: two 16-byte LDS reads
// from the class-row table (the same entries as the class table, row r =
// rank of b among the coprime residues, kCtabrStride words apart).
struct ClassRow { uint32_t e[8]; };
__device__ __forceinline__ ClassRow class_row(const uint32_t* ctabr, uint32_t b)
{
  const uint4* row = (const uint4*) &ctabr[__popc(kCoprime30 & ((1u << b) - 1u)) * kCtabrStride];
  const uint4 r0 = row[0], r1 = row[1];
  return ClassRow{{ r0.x, r0.y, r0.z, r0.w, r1.x, r1.y, r1.z, r1.w }};
}

// All hits of residue class c (bit c) of a prime p <= S from byte i of the
// segment on: the progression i, i+p, i+2p, ... is struck as four
// progressions with stride 4p (a multiple of 4 bytes), so each keeps one
// word mask: the masks of consecutive hits differ by a rotation of
// 8 * (p mod 4) bits.
__device__ __forceinline__ void strike_class(uint32_t i, uint32_t p, uint32_t c, uint32_t S)
{
  const lds_u32* end = lds_word(S);
  const uint32_t rot = (p & 3u) << 3;            // 8 or 24: p is odd
  const uint32_t m0 = 1u << (((i & 3u) << 3) + c);
  const uint32_t m1 = rotl32(m0, rot), m2 = rotl32(m1, rot), m3 = rotl32(m2, rot);
  // word pointers; 4p bytes = p words per step
  lds_u32* q0 = lds_word(i);
  lds_u32* q1 = lds_word(i + p);
  lds_u32* q2 = lds_word(i + 2u * p);
  lds_u32* q3 = lds_word(i + 3u * p);
  for (; q3 < end; q0 += p, q1 += p, q2 += p, q3 += p) {
    lds_or(q0, m0);
    lds_or(q1, m1);
    lds_or(q2, m2);
    lds_or(q3, m3);
  }
  if (q0 < end) lds_or(q0, m0);
  if (q1 < end) lds_or(q1, m1);
  if (q2 < end) lds_or(q2, m2);
}

// All hits of a prime p <= S in the segment at B (R = B mod p, a = p/30,
// row = its class entries): its 8 residue classes, each an arithmetic
// progression of bytes with stride p.
template <bool kNear>
__device__ __forceinline__ void strike_medium(uint32_t p, uint32_t a, uint32_t R, uint64_t B,
                                              uint32_t S, const ClassRow& row)
{
  const uint32_t a1 = a | 0x10000u, nR = ~R;                       // p <= S: a < 2^16
#pragma unroll
  for (uint32_t c = 0; c < 8; c++) {
    const uint32_t ent = row.e[c];
    uint32_t i = class_offset(ent, a1, nR, p);
    if (kNear && (ent & 0xffu) == 1u && B + i < (uint64_t) p)
      i += p;                                   // skip the prime itself
    strike_class(i, p, c, S);
  }
}

// strike_medium for S/4 < p <= S, where a class has at most nh <= 4 hits
// (nh is uniform over the wave): the hits of a class one after another, the
// word mask rotated by 8 * (p mod 4) bits from one hit to the next.  kNh > 0
// fixes nh at compile time (the loop unrolls).
template <bool kNear, uint32_t kNh = 0>
__device__ __forceinline__ void strike_few(uint32_t p, uint32_t a, uint32_t R, uint64_t B,
                                           uint32_t S, const ClassRow& row, uint32_t nh)
{
  if (kNh) nh = kNh;
  const uint32_t a1 = a | 0x10000u, nR = ~R;                       // p <= S: a < 2^16
  const uint32_t rot = (p & 3u) << 3;
#pragma unroll
  for (uint32_t c = 0; c < 8; c++) {
    const uint32_t ent = row.e[c];
    uint32_t i = class_offset(ent, a1, nR, p);
    if (kNear && (ent & 0xffu) == 1u && B + i < (uint64_t) p)
      i += p;                                   // skip the prime itself
    uint32_t m = 1u << (((i & 3u) << 3) + c);
#pragma unroll
    for (uint32_t k = 0; k < (kNh ? kNh : nh); k++) {
      // the first hit i < p <= S is inside the segment unless it was moved
      // past the prime itself
      if ((!kNear && k == 0) || i < S) lds_or(lds_word(i), m);
      i += p;
      m = rotl32(m, rot);
    }
  }
}

// A tier-2 prime: the few-hit path when the smallest prime of its half of the
// item (pMin, uniform) is above S/4, else the progressions.
__device__ __forceinline__ void strike_tier2(uint32_t p, uint32_t a, uint32_t R, uint64_t B,
                                             uint32_t S, const ClassRow& row, uint32_t pMin,
                                             bool nearSmall)
{
  if (4u * pMin > S) {
    const uint32_t nh = (2u * pMin > S) ? 2u : (3u * pMin > S) ? 3u : 4u;
    if (nearSmall) strike_few<true>(p, a, R, B, S, row, nh);
    else if (nh == 2u) strike_few<false, 2>(p, a, R, B, S, row, nh);   // p > S/2: the most primes
    else if (nh == 3u) strike_few<false, 3>(p, a, R, B, S, row, nh);
    else strike_few<false, 4>(p, a, R, B, S, row, nh);
  } else {
    if (nearSmall) strike_medium<true>(p, a, R, B, S, row);
    else strike_medium<false>(p, a, R, B, S, row);
  }
}

// The <= 8 hits of a prime S < p in the segment at B (R = B mod p, a = p/30,
// row = its class entries): one per residue class at most.  kSmall:
// p < 30 * 2^16, so that class_offset applies.
template <bool kNear, bool kSmall>
__device__ __forceinline__ void strike_big(uint32_t p, uint32_t a, uint32_t R, uint64_t B,
                                           uint32_t S, const ClassRow& row)
{
  const uint32_t a1 = a | 0x10000u, nR = ~R;
#pragma unroll
  for (uint32_t c = 0; c < 8; c++) {
    const uint32_t ent = row.e[c];
    uint32_t i;
    if (kSmall) {
      i = class_offset(ent, a1, nR, p);
    } else {
      const uint32_t cc = a * (ent & 0xffffu) + (ent >> 16) - 1u;
      i = (cc >= R) ? cc - R : cc - R + p;
    }
    if (kNear && (ent & 0xffu) == 1u && B + i < (uint64_t) p)
      i = 0xffffffffu;                  // the prime itself; its next hit is beyond S
    if (i < S) lds_or(lds_word(i), 1u << (((i & 3u) << 3) + c));
  }
}

// strike_big for the kN primes of a lane (p[m] = 0: none).  All their class
// rows are read before the first strike: the LDS operations of a wave
// complete in order, so a row read issued after strikes would wait for all
// of them.
template <bool kNear, bool kSmall, uint32_t kN>
__device__ __forceinline__ void strike_big_lane(const uint32_t (&p)[kN], const uint32_t (&R)[kN],
                                                uint64_t B, uint32_t S, const uint32_t* ctab)
{
  uint32_t a[kN];
  ClassRow row[kN];
#pragma unroll
  for (uint32_t m = 0; m < kN; m++) {
    a[m] = p[m] / 30u;
    row[m] = class_row(ctab, p[m] - a[m] * 30u);
  }
#pragma unroll
  for (uint32_t m = 0; m < kN; m++)
    if (p[m]) strike_big<kNear, kSmall>(p[m], a[m], R[m], B, S, row[m]);
}

// The hits of the kN primes of a lane (p[m] = 0: none) with p < 30 * 2^16
// in a segment at B >= p (R = B mod p, M the division magic), for primes
// with 8p > 30S (kK = 3) or 6p > 30S (kK = 2).  The segment's 30S values
// hold at most 8 (6) consecutive multiples of such a prime, and any 8 (6)
// consecutive integers hold at most 3 (2) coprime to 30.  With B = Qp + R,
// the first multiple at or after the segment's first value 30B + 7 is
// p * (30Q + t), t = ceil((30R + 7) / p) in [1, 30], so its hits are among
// the kK slots of its cofactor-ordered row from the first coprime residue
// >= t (t = 30 is residue 0, the row's slot 8, which repeats slot 0); each is
// tested exactly, as in strike_big.  kK adjacent table slots and kK strikes
// per prime instead of two row reads and eight strikes, and no class offsets
// for the other classes.  A slot holds the class's bit, so a strike's word
// mask is that bit shifted by 8 * (i mod 4): (i << 3) mod 32.
template <uint32_t kN, uint32_t kK>
__device__ __forceinline__ void strike_cands_lane(const uint32_t (&p)[kN], const uint32_t (&R)[kN],
                                                  const uint64_t (&M)[kN], uint32_t S, const uint2* ctabq)
{
  uint32_t a[kN];
  uint64_t e[kN][kK];         // slots read whole (8-byte LDS reads, two per instruction)
#pragma unroll
  for (uint32_t m = 0; m < kN; m++) {
    a[m] = p[m] / 30u;
    const uint32_t b = p[m] - a[m] * 30u;
    // t = ceil((30R + 7) / p) = floor((30R + 7 + p - 1) / p) from the magic's
    // high word Mh = floor(2^(32+L) / p): floor(y Mh / 2^(32+L)) for
    // y = 30R + 7 + p (one more than the dividend) is exact for y < 2^26 and
    // p < 2^21, since y Mh / 2^(32+L) falls short of y / p by less than 1/p
    // (and by more than 0: p is odd).
    const uint32_t t = __umulhi(30u * R[m] + 7u + p[m], (uint32_t) (M[m] >> 32)) >> magic_shift(p[m]);
    const uint64_t* row = (const uint64_t*) ctabq + ((b >> 1) * kCtabqSlots + __popc(kCoprime30 & ((1u << t) - 1u)));
#pragma unroll
    for (uint32_t k = 0; k < kK; k++) e[m][k] = row[k];
  }
#pragma unroll
  for (uint32_t m = 0; m < kN; m++) {
    if (!p[m]) continue;
    const uint32_t a1 = a[m] | 0x10000u, nR = ~R[m];
#pragma unroll
    for (uint32_t k = 0; k < kK; k++) {
      const uint32_t i = class_offset((uint32_t) e[m][k], a1, nR, p[m]);
      if (i < S) lds_or(lds_word(i), (uint32_t) (e[m][k] >> 32) << ((i << 3) & 31u));
    }
  }
}

// Strike bit `v & 31` of word `v >> 5`: v = byte * 8 + bit within the segment.
__device__ __forceinline__ void strike_bit(uint32_t v)
{
  lds_or(lds_word(v >> 3), 1u << (v & 31u));
}

// Strike one class progression from byte `pos` with word stride `step`
// (a multiple of 4, so the word mask is loop-invariant) up to byte `hi`
// (a multiple of 4); four atomics per loop trip.
__device__ __forceinline__ void strike_run(uint32_t pos, uint32_t step, uint32_t hi, uint32_t bit)
{
  const uint32_t mask = 1u << (((pos & 3u) << 3) + bit);
  lds_u32* q = lds_word(pos);
  const lds_u32* end = lds_word(hi);
  const uint32_t ws = step >> 2;
  for (; q + 3u * ws < end; q += 4u * ws) {
    lds_or(q, mask);
    lds_or(q + ws, mask);
    lds_or(q + 2u * ws, mask);
    lds_or(q + 3u * ws, mask);
  }
  for (; q < end; q += ws)
    lds_or(q, mask);
}

// Sum over the 64 lanes of a wave: xor butterflies inside each 32-lane half
// with ds_swizzle (bitmask mode: and_mask 0x1f, xor_mask 1..16), then the two
// halves through readlane.
__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
  v += (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, 0x041f);
  v += (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, 0x081f);
  v += (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, 0x101f);
  v += (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, 0x201f);
  v += (uint32_t) __builtin_amdgcn_ds_swizzle((int) v, 0x401f);
  return (uint32_t) __builtin_amdgcn_readlane((int) v, 0) +
         (uint32_t) __builtin_amdgcn_readlane((int) v, 32);
}

// Zero the bytes of an 8-byte sieve word outside [gLo, gHi] and apply the
// partial-byte masks of the range ends.  gw = global byte of the word's byte 0.
__device__ uint64_t mask_edges(uint64_t x, uint64_t gw, const PsgRun& r)
{
  if (gw > r.gHi || gw + 7 < r.gLo)
    return 0;
  uint64_t out = 0;
  for (uint32_t k = 0; k < 8; k++) {
    const uint64_t g = gw + k;
    uint32_t byte = (uint32_t) (x >> (8 * k)) & 0xffu;
    if (g < r.gLo || g > r.gHi) byte = 0;
    if (g == r.gLo) byte &= r.loMask;
    if (g == r.gHi) byte &= r.hiMask;
    out |= (uint64_t) byte << (8 * k);
  }
  return out;
}


// Residue-class table entry, derived from first principles: for b = p mod 30
// and class j (offset o_j), q_j = o_j * b^-1 (mod 30) and
// e_j = (b*q_j - o_j)/30, packed as q_j | (e_j + 1) << 16 (e_j is -1 only
// for o_j = 31, b = 1); 0 for b not coprime to 30.
__device__ __forceinline__ uint32_t class_entry(uint32_t b, uint32_t j)
{
  const uint32_t o = kOffsets[j];
  if ((b % 2u) && (b % 3u) && (b % 5u)) {
    for (uint32_t q = 1; q < 30; q++)
      if ((b * q) % 30u == o % 30u) {
        const int e = ((int) (b * q) - (int) o) / 30;   // exact, may be -1
        return q | ((uint32_t) (e + 1) << 16);
      }
  }
  return 0u;
}

__device__ __forceinline__ void init_ctab(uint32_t* ctab)
{
  for (uint32_t t = threadIdx.x; t < 240; t += blockDim.x) ctab[t] = class_entry(t >> 3, t & 7u);
}

// The class-row table: row r holds the 8 entries of the r-th coprime residue.
__device__ __forceinline__ void init_ctabr(uint32_t* ctabr)
{
  for (uint32_t t = threadIdx.x; t < 8u * kCtabrStride; t += blockDim.x) {
    const uint32_t r = t / kCtabrStride, j = t % kCtabrStride;
    ctabr[t] = (j < 8u) ? class_entry(kCoprimes[r], j) : 0u;
  }
}

// The same entries in cofactor order: ctabq[(b >> 1) * 13 + s] = (the entry
// of the class whose cofactor residue is the (s mod 8)-th coprime residue,
// the bit of that class) for b = p mod 30 coprime to 30 (one row per odd b;
// the rows of odd b sharing a factor with 30 are zero).
// Consecutive multiples of p with cofactors coprime to 30 are consecutive
// slots of its row; the row repeats its 8 slots once (kCtabqSlots), so that
// the slots from any s <= 8 on need no wrap.

__device__ __forceinline__ void init_ctabq(uint2* ctabq)
{
  for (uint32_t t = threadIdx.x; t < 15u * kCtabqSlots; t += blockDim.x) {
    const uint32_t b = 2u * (t / kCtabqSlots) + 1u, q = kCoprimes[(t % kCtabqSlots) & 7u];
    uint2 e = make_uint2(0u, 0u);
    for (uint32_t j = 0; j < 8 && ((kCoprime30 >> b) & 1u); j++) {
      const uint32_t o = kOffsets[j];
      if ((b * q) % 30u == o % 30u) {
        const int ej = ((int) (b * q) - (int) o) / 30;   // exact, may be -1
        e = make_uint2(q | ((uint32_t) (ej + 1) << 16), 1u << j);
      }
    }
    ctabq[t] = e;
  }
}

// Window index (B + 16 tid) mod P of each tile for the 16-byte run of thread
// tid of the segment at B.  B is the same for the whole workgroup, so its
// remainders are taken once, in scalar registers.
__device__ __forceinline__ void tile_start(uint64_t B, uint32_t tid, uint32_t (&idx)[PSG_NTILES])
{
  constexpr uint32_t kPeriod[PSG_NTILES] = PSG_TILE_PERIODS;
  const uint32_t hi = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) (B >> 32));
  const uint32_t lo = (uint32_t) __builtin_amdgcn_readfirstlane((int) (uint32_t) B);
  const uint64_t Bu = ((uint64_t) hi << 32) | lo;
#pragma unroll
  for (int t = 0; t < PSG_NTILES; t++) {
    const uint32_t x = (uint32_t) (Bu % kPeriod[t]) + (16u * tid) % kPeriod[t];
    idx[t] = (x >= kPeriod[t]) ? x - kPeriod[t] : x;
  }
}

// Fill the segment from the pre-sieve tiles (stored complemented: a set bit
// marks a multiple of a tile prime): word w is the OR over the tiles of
// window (B + 4w) mod P.  Tile 0 comes from global memory, the small
// tiles from their LDS copy (kLds) or global memory.  A thread fills runs of
// 4 consecutive words: their windows are 4 adjacent tile words (4 loads that
// merge into wide ones, one index update per run) and the run is one 16-byte
// LDS store.  Two runs per thread per trip, so that all their loads are in
// flight before the first store.
template <bool kLds>
__device__ __forceinline__ void fill_tiles(uint32_t* sieve, uint32_t words, uint64_t B,
                                           const PSG_GLOBAL uint32_t* tiles, const uint32_t* small)
{
  constexpr uint32_t kPeriod[PSG_NTILES] = PSG_TILE_PERIODS;
  constexpr uint32_t kRows[PSG_NTILES] = PSG_TILE_ROWS;
  constexpr uint32_t kOffset[PSG_NTILES] = PSG_TILE_OFFSETS;
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = blockDim.x;
  uint32_t idx[PSG_NTILES], stp[PSG_NTILES];
  tile_start(B, tid, idx);
#pragma unroll
  for (int t = 0; t < PSG_NTILES; t++)
    stp[t] = (16u * nthreads) % kPeriod[t];
  const auto run = [&]() {
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int t = 0; t < PSG_NTILES; t++) {
      const uint32_t pos = (idx[t] & 3u) * kRows[t] + (idx[t] >> 2);
      if (t == 0 || !kLds) {
        const PSG_GLOBAL uint32_t* q = tiles + kOffset[t] + pos;
        v.x |= q[0]; v.y |= q[1]; v.z |= q[2]; v.w |= q[3];
      } else {
        const uint32_t* q = small + (kOffset[t] - kOffset[1]) + pos;
        v.x |= q[0]; v.y |= q[1]; v.z |= q[2]; v.w |= q[3];
      }
      idx[t] += stp[t];
      idx[t] = min(idx[t], idx[t] - kPeriod[t]);     // below 0 wraps to a large value
    }
    return v;
  };
  uint4* s4 = (uint4*) sieve;
  const uint32_t runs = words >> 2;
  uint32_t r = tid;
  for (; r + nthreads < runs; r += 2u * nthreads) {
    const uint4 a = run();
    const uint4 b = run();
    s4[r] = a;
    s4[r + nthreads] = b;
  }
  for (; r < runs; r += nthreads) s4[r] = run();
}

// Tier-3 primes per lane: a tier-3 item holds 64 * kT3 primes, so that the
// per-item costs (the queue ticket, the wait for the previous item's strikes
// before its class rows can be read) are shared by kT3 primes.
constexpr uint32_t kT3 = 4;

// LDS after the segment: the class table, the workgroup's counters (the six
// sums, segment and run numbers, the queue counter, the list mode's wave
// totals), the tier-0 primes and their magics (read once per workgroup),
// then the class-row table (8 rows of kCtabrStride words) and the
// cofactor-ordered class table (15 rows of kCtabqSlots uint2).
constexpr uint32_t kCtabWords = 240;
constexpr uint32_t kVarWords = 32;
constexpr uint32_t kTier0Slots = PSG_TIER0_SLOTS;
constexpr uint32_t kCtabqWords = 2u * 15u * kCtabqSlots;
constexpr uint32_t kCtabrWords = 8u * kCtabrStride;
static_assert((kCtabWords + kVarWords + 3 * kTier0Slots + kCtabqWords + kCtabrWords) * 4 == PSG_SIEVE_LDS_TAIL,
              "LDS tail layout");

// The primes of work item `it` of a segment's queue (see strike_segment):
// primes[kb + i] for i < n, with kb and n uniform (n = 0 past the queue).
// Items [0, u1) hold 8 primes of [t0, t1) each (lane = prime * 8 + class:
// slot i = lane / 8), [u1, u2) 128 primes of [t1, t2) (2 per lane: slots
// lane, lane + 64), [u2, u3) 64 * kT3 primes of [t2, t3) (lane's m-th prime:
// slot lane + 64 m).  `it` is uniform.
__device__ __forceinline__ void item_span(uint32_t it, uint32_t t0, uint32_t t1, uint32_t t2,
                                          uint32_t t3, uint32_t u1, uint32_t u2, uint32_t u3,
                                          uint32_t& kb, uint32_t& n)
{
  uint32_t end;
  if (it < u1) {
    kb = t0 + (it << 3);
    end = t1;
  } else if (it < u2) {
    kb = t1 + (it - u1) * 128u;
    end = t2;
  } else if (it < u3) {
    kb = t2 + (it - u2) * (64u * kT3);
    end = t3;
  } else {
    kb = 0;
    end = 0;
  }
  n = end - min(kb, end);
}

// List mode of step 3: write the segment's primes (ascending) instead of
// counting them.  Thread t owns words64 / nthreads consecutive 64-bit words;
// an exclusive scan of the per-thread counts (waves by shuffles, the wave
// totals through LDS) gives each thread its first output index.  The run is
// one segment, so its output starts at outOffsets[run.out].
__device__ __forceinline__ void list_primes(const PsgSieveArgs& args, const PsgRun& run, uint64_t B,
                            const uint32_t* sieve, uint32_t* waveTot)
{
  const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
  const uint32_t lane = tid & 63u, wave = tid >> 6;
  const uint64_t* s64 = (const uint64_t*) sieve;
  const uint32_t per = (args.segBytes >> 3) / nthreads;     // the host checks divisibility
  const uint32_t w0 = tid * per;
  const auto word = [&](uint32_t w) {
    uint64_t x = ~s64[w];
    const uint64_t gw = B + 8ull * w;
    if (gw <= run.gLo || gw + 7 >= run.gHi)
      x = mask_edges(x, gw, run);
    return x;
  };
  uint32_t c = 0;
  for (uint32_t k = 0; k < per; k++) c += __popcll(word(w0 + k));
  uint32_t x = c;
  for (uint32_t d = 1; d < 64u; d <<= 1) {
    const uint32_t y = (uint32_t) __shfl_up((int) x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63u) waveTot[wave] = x;
  __syncthreads();
  uint32_t pos = gptr<uint32_t>(args.outOffsets)[run.out] + x - c;
  for (uint32_t v = 0; v < wave; v++) pos += waveTot[v];
  auto* out = gptr_rw<uint32_t>(args.primesOut);
  constexpr uint64_t kOffsetBytes = 0x1F1D1713110D0B07ull;   // 7, 11, 13, 17, 19, 23, 29, 31
  for (uint32_t k = 0; k < per; k++) {
    uint64_t m = word(w0 + k);
    const uint64_t gw = B + 8ull * (w0 + k);
    while (m) {
      const uint32_t b = (uint32_t) __builtin_ctzll(m);
      m &= m - 1u;
      const uint64_t v = 30ull * (gw + (b >> 3)) + ((kOffsetBytes >> ((b & 7u) * 8u)) & 0xffu);
      out[pos++] = (uint32_t) v;
    }
  }
}

// The tiles also mark the tile primes 7..47 themselves: all of global byte 0
// (7..31) and bits 0-3 of global byte 1 (37..47).  A segment that starts at
// byte 0 or 1 has them in its word 0, which thread 0 wrote and restores.
__device__ __forceinline__ void restore_tile_primes(uint32_t* sieve, uint64_t B)
{
  if (B < 2u && threadIdx.x == 0)
    sieve[0] &= (B == 0) ? ~0x0fffu : ~0x000fu;
}

// A run read from global memory at a uniform index, its fields made
// uniform (scalar registers) so that they do not occupy vector registers.
__device__ __forceinline__ PsgRun load_run(const PSG_GLOBAL PsgRun* runs, uint32_t i)
{
  union { PsgRun r; uint32_t w[sizeof(PsgRun) / 4]; } u;
  u.r = runs[i];
#pragma unroll
  for (uint32_t k = 0; k < sizeof(PsgRun) / 4; k++)
    u.w[k] = (uint32_t) __builtin_amdgcn_readfirstlane((int) u.w[k]);
  return u.r;
}

// Step 1 for the segment at B: the pre-sieve fill (or zeros without tiles).
// No barrier.
__device__ __forceinline__ void fill_segment(const PsgSieveArgs& args, uint64_t B, uint32_t* sieve,
                                             const uint32_t* small)
{
  const uint32_t tid = threadIdx.x;
  const uint32_t words = args.segBytes >> 2;
  const auto* __restrict__ tiles = gptr<uint32_t>(args.tiles);
  if (!tiles) {
    // pre-sieving disabled: every prime >= 7 is struck in step 2
    for (uint32_t w = tid; w < words; w += blockDim.x)
      sieve[w] = 0u;
  } else {
    if (args.tilesInLds) fill_tiles<true>(sieve, words, B, tiles, small);
    else fill_tiles<false>(sieve, words, B, tiles, small);
    restore_tile_primes(sieve, B);
  }
}

// Step 2 for segment `seg` at B of `run`: strike every sieving prime and the
// segment's bucket list.  The queue counter must be 0; no barrier.
// take() (called by thread 0 after tier 0) claims the workgroup's next
// segment: its ticket and run lookup then overlap the other waves' strikes
// instead of holding the whole workgroup at the barrier after them.
template <typename Take>
__device__ __forceinline__ void strike_segment(const PsgSieveArgs& args, const PsgRun& run,
                                               uint32_t seg, uint64_t B, uint32_t* sieve,
                                               const uint32_t* ctab, uint32_t* wsum, Take take)
{
  const uint32_t* ctabr = ctab + kCtabWords + kVarWords + 3 * kTier0Slots;     // 16-byte aligned
  const uint2* ctabq = (const uint2*) (ctabr + kCtabrWords);
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = blockDim.x;
  const uint32_t lane = tid & 63u;
  const uint32_t S = args.segBytes;

  const auto* __restrict__ primes = gptr<uint32_t>(args.primes);
  const auto* __restrict__ magics = gptr<uint64_t>(args.magics);
  uint32_t* sWork = wsum + PSG_NCOUNTS + 2;        // the work queue's counter

  // The first kBucketPre bucket entries of every thread are loaded now (after
  // the fill: its tile loads would otherwise wait for them), so that their
  // HBM latency passes during tier 0, which waits on no global load.  They
  // are struck when the queue starts.
  constexpr uint32_t kBucketPre = 8;
  uint32_t bpre[kBucketPre];
  uint32_t bE = 0, bE1 = 0;
  uint32_t bW = 0;                  // bE - tid, the same for every thread (scalar)
  if (args.buckets) {
    const auto* __restrict__ offs = gptr<uint32_t>(args.listOff);
    const auto* __restrict__ ent = gptr<uint32_t>(args.buckets);
    const uint32_t e0 = offs[seg];
    const uint32_t room = offs[seg + 1u] - e0;
    bE1 = e0 + min(gptr<uint32_t>(args.listCount)[seg], room);
    bW = (uint32_t) __builtin_amdgcn_readfirstlane((int) e0);
    bE = bW + tid;
#pragma unroll
    for (uint32_t u = 0; u < kBucketPre; u++) {
      const uint32_t idx = bE + u * nthreads;
      bpre[u] = (idx < bE1) ? ent[idx] : 0u;
    }
  }

  // ---- 2. strike the sieving primes --------------------------------------
  {
    const uint32_t np = run.nprimes;
    const uint32_t t0 = min(args.tier0End, np);
    const uint32_t t1 = max(t0, min(args.tier1End, np));
    const uint32_t t2 = max(t1, min(args.tier2End, np));
    const uint32_t t3 = max(t2, min(args.tier3End, np));
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = nthreads >> 6;
    const uint32_t j = lane & 7u;
    const uint32_t slot = lane >> 3;
    const bool nearSmall = B < (uint64_t) S;   // tiers 0-2 hold primes <= S

    // tier 0 (p <= 256): every wave strikes all of them over its own slice
    // of the segment, so the heaviest primes are shared evenly.  Their primes
    // and magics come from the workgroup's LDS copy: no global load waits
    // here, so the bucket entries loaded above arrive meanwhile.
    {
      const uint32_t* t0p = ctab + kCtabWords + kVarWords;
      const uint64_t* t0m = (const uint64_t*) (t0p + kTier0Slots);
      const uint32_t sliceBytes = S / nwaves;          // S is a multiple of 256
      const uint32_t sliceLo = wave * sliceBytes;
      const uint32_t sliceHi = sliceLo + sliceBytes;
      const uint64_t Bs = B + sliceLo;
      for (uint32_t b = 0; b < t0; b += 8) {
        const uint32_t k = b + slot;                     // t0 <= kTier0Slots
        const uint32_t p = (k < t0) ? t0p[k] : 0u;
        const uint32_t R = p ? mod_magic31(Bs, p, t0m[k]) : 0u;
        const uint32_t st = !p ? 0u : nearSmall ? class_start<true>(p, R, Bs, j, ctab)
                                               : class_start<false>(p, R, Bs, j, ctab);
        // all eight permutes first: one wait for the batch, not one per prime
        uint32_t sq[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) sq[q] = (uint32_t) __shfl((int) st, (int) (q * 8u + j));
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) {
          const uint32_t ps = (uint32_t) __builtin_amdgcn_readlane((int) p, (int) (q * 8u));
          if (!ps) break;
          strike_run(sliceLo + sq[q] + slot * ps, 8u * ps, sliceHi, j);
        }
      }
    }

    if (tid == 0) take();

    // tiers 1-3: a queue of work items shared by the waves of the workgroup
    // (an LDS counter), so that no wave is left with a long tail of work
    // while the others wait at the barrier:
    //   tier 1 (256 < p <= 2048): 8 primes per item, lane = (prime, class),
    //     each lane strikes one class over the whole segment (the primes of
    //     an item are close, so its lanes finish together);
    //   tier 2 (2048 < p <= S): 128 primes per item, 2 per lane;
    //   tier 3 (S < p <= bucket threshold): 64 * kT3 primes per item, kT3 per
    //     lane, <= 1 hit per class; the prime itself can only be in the
    //     segment while B < p (uniform test against the largest tier-3 prime).
    // Ticket t maps to an item so that tiers 1-2 keep their order (heaviest
    // first) and the light tier-3 items are spread evenly between them: waves
    // on LDS-bound strikes then share the CU with waves computing tier-3
    // class starts.  With g(t) = floor(t * n3 / u3) tier-3 items before
    // ticket t (n3 = u3 - u2; computed as (t * r) >> 32 with
    // r = ceil(n3 * 2^32 / u3), which is exact at t = u3 and steps by at most
    // one), ticket t is tier-3 item g(t) when g(t + 1) > g(t), otherwise item
    // t - g(t) of tiers 1-2.
    // While an item is struck, the next item's number and primes are in
    // flight, and so is the thread's next batch of 8 bucket entries: a batch
    // is struck at the start of the item after the one it was loaded in.
    {
      const uint32_t u1 = (t1 - t0 + 7u) >> 3;
      const uint32_t u2 = u1 + (t2 - t1 + 127u) / 128u;
      const uint32_t u3 = u2 + (t3 - t2 + 64u * kT3 - 1u) / (64u * kT3);
      const uint32_t pTop = (t3 > t2) ? primes[t3 - 1u] : 0u;
      const bool near = B < (uint64_t) pTop;
      const bool small = pTop < (30u << 16);
      // the candidate bands: 6p > 30S (two classes), 8p > 30S (three), as
      // p > floor(30S / 6) and p > floor(30S / 8) for integers p
      const uint32_t cand2 = 5u * S, cand3 = (30u * S) >> 3;
      const uint64_t rmix = (((uint64_t) (u3 - u2) << 32) + u3 - 1u) / max(u3, 1u);
      const auto itemOf = [=](uint32_t t) -> uint32_t {
        if (t >= u3) return t;
        const uint32_t g0 = (uint32_t) (((uint64_t) t * rmix) >> 32);
        const uint32_t g1 = (uint32_t) (((uint64_t) (t + 1u) * rmix) >> 32);
        return (g1 > g0) ? u2 + g0 : t - g0;
      };
      // An item's primes and magics, lane slot first (p = 0, M = 0: none).
      // The item's first prime index is uniform, so the loads take scalar
      // bases and the lane's fixed slot offset.
      const auto load_item = [&](uint32_t it, uint32_t& p0, uint64_t& M0, uint32_t (&pm)[kT3 - 1],
                                 uint64_t (&Mm)[kT3 - 1]) {
        uint32_t kb, n;
        item_span(it, t0, t1, t2, t3, u1, u2, u3, kb, n);
        const auto* __restrict__ pp = primes + kb;
        const auto* __restrict__ mp = magics + kb;
        const uint32_t sl = (it < u1) ? (lane >> 3) : lane;
        const uint32_t more = (it >= u2) ? kT3 - 1u : (it >= u1) ? 1u : 0u;   // further primes per lane
        p0 = (sl < n) ? pp[sl] : 0u;
        M0 = (sl < n) ? mp[sl] : 0ull;
#pragma unroll
        for (uint32_t m = 0; m < kT3 - 1; m++) {
          const bool ok = m < more && (int) sl < (int) n - (int) (64u * (m + 1u));   // n < 2^31
          pm[m] = ok ? pp[sl + 64u * (m + 1u)] : 0u;
          Mm[m] = ok ? mp[sl + 64u * (m + 1u)] : 0ull;
        }
      };
      const auto ticket = [=]() -> uint32_t { return (lane == 0) ? atomicAdd(sWork, 1u) : 0u; };
      const auto uniform = [](uint32_t v) { return (uint32_t) __builtin_amdgcn_readfirstlane((int) v); };
      uint32_t cur = itemOf(uniform(ticket()));
      uint32_t nxt = itemOf(uniform(ticket()));
      // the lane's first prime of the current item and its further ones
      // (slot + 64 m: one more in tier 2, kT3 - 1 more in tier 3)
      uint32_t p;
      uint64_t M;
      uint32_t pX[kT3 - 1];
      uint64_t MX[kT3 - 1];
      load_item(cur, p, M, pX, MX);
      // the bucket entries prefetched before tier 0 (arrived by now), then
      // the first batch in flight; a lane loads min(index, last) so that its
      // loads stay inside the list.  A batch that lies inside the list for
      // every lane of the wave (all but a list's last ones) is loaded and
      // struck without the per-entry tests.
      const auto* __restrict__ bent = gptr<uint32_t>(args.buckets);
      const auto strike_batch = [&](const uint32_t (&v)[8]) {
        if (__all(bE + 7u * nthreads < bE1)) {
#pragma unroll
          for (uint32_t u = 0; u < 8; u++) strike_bit(v[u]);
        } else {
#pragma unroll
          for (uint32_t u = 0; u < 8; u++)
            if (bE + u * nthreads < bE1) strike_bit(v[u]);
        }
      };
      const auto load_batch = [&](uint32_t (&v)[8]) {
        if (__all(bE + 7u * nthreads < bE1)) {
          // scalar bases plus the thread's fixed offset: no vector address
          // arithmetic per load
#pragma unroll
          for (uint32_t u = 0; u < 8; u++) v[u] = (bent + (bW + u * nthreads))[tid];
        } else {
#pragma unroll
          for (uint32_t u = 0; u < 8; u++) v[u] = bent[min(bE + u * nthreads, bE1 - 1u)];
        }
      };
      static_assert(kBucketPre == 8, "the prefetched batch is struck as one batch");
      strike_batch(bpre);
      bE += kBucketPre * nthreads;
      bW += kBucketPre * nthreads;
      uint32_t bq[8];
      if (bE < bE1) load_batch(bq);
      while (cur < u3) {
        const uint32_t after = ticket();                 // the item after nxt, in flight
        if (bE < bE1) {
          strike_batch(bq);
          bE += 8u * nthreads;
          bW += 8u * nthreads;
          if (bE < bE1) load_batch(bq);
        }
        uint32_t pN;
        uint64_t MN;
        uint32_t pNX[kT3 - 1];
        uint64_t MNX[kT3 - 1];
        load_item(nxt, pN, MN, pNX, MNX);
        if (p) {
          if (cur < u1) {
            const uint32_t R = mod_magic31(B, p, M);             // p <= S < 2^31
            const uint32_t i = nearSmall ? class_start<true>(p, R, B, j, ctab)
                                         : class_start<false>(p, R, B, j, ctab);
            strike_class(i, p, j, S);
          } else if (cur < u2) {
            // two primes per lane (k, k + 64), both class rows read before
            // the first strike; lane 0 holds each half's smallest prime
            const uint32_t p2 = pX[0];
            const uint32_t R = mod_magic31(B, p, M);
            const uint32_t R2 = p2 ? mod_magic31(B, p2, MX[0]) : 0u;
            const uint32_t pMin = uniform(p), pMin2 = uniform(p2);
            const uint32_t a = p / 30u, a2 = p2 / 30u;
            const ClassRow row = class_row(ctabr, p - a * 30u);
            const ClassRow row2 = class_row(ctabr, p2 - a2 * 30u);
            strike_tier2(p, a, R, B, S, row, pMin, nearSmall);
            if (p2) strike_tier2(p2, a2, R2, B, S, row2, pMin2, nearSmall);
          } else {
            // remainders (p < 30 * 2^16 < 2^31 when `small`; lanes without a
            // prime, p = 0 and M = 0, get an unused value)
            uint32_t pa[kT3], Ra[kT3];
            uint64_t Ma[kT3];
            pa[0] = p;
            Ma[0] = M;
#pragma unroll
            for (uint32_t m = 0; m < kT3 - 1; m++) {
              pa[m + 1] = pX[m];
              Ma[m + 1] = MX[m];
            }
            if (small) {
#pragma unroll
              for (uint32_t m = 0; m < kT3; m++) Ra[m] = mod_magic31(B, pa[m], Ma[m]);
            } else {
#pragma unroll
              for (uint32_t m = 0; m < kT3; m++) Ra[m] = mod_magic(B, pa[m], Ma[m]);
            }
            // lane 0 holds the item's smallest prime
            if (near) strike_big_lane<true, false, kT3>(pa, Ra, B, S, ctabr);
            else if (small && uniform(p) > cand2) strike_cands_lane<kT3, 2>(pa, Ra, Ma, S, ctabq);
            else if (small && uniform(p) > cand3) strike_cands_lane<kT3, 3>(pa, Ra, Ma, S, ctabq);
            else if (small) strike_big_lane<false, true, kT3>(pa, Ra, B, S, ctabr);
            else strike_big_lane<false, false, kT3>(pa, Ra, B, S, ctabr);
          }
        }
        cur = nxt;
        nxt = itemOf(uniform(after));
        p = pN;
        M = MN;
#pragma unroll
        for (uint32_t m = 0; m < kT3 - 1; m++) {
          pX[m] = pNX[m];
          MX[m] = MNX[m];
        }
      }
      if (bE < bE1) {
        strike_batch(bq);
        bE += 8u * nthreads;
      }
    }
  }

  // ---- 2b. the rest of the segment's bucket list -------------------------
  // Entries are bit indices (byte * 8 + bit).  Most were struck during the
  // queue; the rest (short queues) with eight independent loads per thread in
  // flight before their strikes.
  if (args.buckets) {
    const auto* __restrict__ ent = gptr<uint32_t>(args.buckets);
    const uint32_t e1 = bE1;
    uint32_t e = bE;
    for (; e + 7u * nthreads < e1; e += 8u * nthreads) {
      uint32_t v[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) v[u] = ent[e + u * nthreads];
#pragma unroll
      for (uint32_t u = 0; u < 8; u++) strike_bit(v[u]);
    }
    for (; e < e1; e += nthreads) strike_bit(ent[e]);
  }
}

// Step 3 for the segment at B of `run`, fused with step 1 of the next one
// (at nB, when hasNext): one pass over the LDS sieve in 16-byte chunks, each
// read, counted (all six types) and overwritten with the next segment's
// pre-sieve fill by the same thread.  The count's VALU work then covers the
// fill's tile loads.  Adds this thread's counts to c[].  No barrier.
__device__ __forceinline__ void count_and_fill(const PsgSieveArgs& args, const PsgRun& run,
                                               uint64_t B, bool hasNext, uint64_t nB,
                                               uint32_t* sieve, const uint32_t* small,
                                               uint32_t (&c)[PSG_NCOUNTS])
{
  constexpr uint32_t kPeriod[PSG_NTILES] = PSG_TILE_PERIODS;
  constexpr uint32_t kRows[PSG_NTILES] = PSG_TILE_ROWS;
  constexpr uint32_t kOffset[PSG_NTILES] = PSG_TILE_OFFSETS;
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = blockDim.x;
  const uint32_t S = args.segBytes;
  const auto* __restrict__ tiles = gptr<uint32_t>(args.tiles);
  const bool fill = hasNext && tiles;              // else the next segment starts from zeros
  const bool lds = args.tilesInLds != 0;
  uint32_t idx[PSG_NTILES] = {0u, 0u, 0u, 0u, 0u}, stp[PSG_NTILES];
  if (fill) tile_start(nB, tid, idx);
#pragma unroll
  for (int t = 0; t < PSG_NTILES; t++)
    stp[t] = (16u * nthreads) % kPeriod[t];
  const auto next_run = [&]() {                    // as in fill_tiles
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
    for (int t = 0; t < PSG_NTILES; t++) {
      const uint32_t pos = (idx[t] & 3u) * kRows[t] + (idx[t] >> 2);
      if (t == 0 || !lds) {
        const PSG_GLOBAL uint32_t* q = tiles + kOffset[t] + pos;
        v.x |= q[0]; v.y |= q[1]; v.z |= q[2]; v.w |= q[3];
      } else {
        const uint32_t* q = small + (kOffset[t] - kOffset[1]) + pos;
        v.x |= q[0]; v.y |= q[1]; v.z |= q[2]; v.w |= q[3];
      }
      idx[t] += stp[t];
      idx[t] = min(idx[t], idx[t] - kPeriod[t]);     // below 0 wraps to a large value
    }
    return v;
  };
  const auto add = [&](uint64_t x) {
    const uint64_t t1 = x & (x >> 1);
    const uint64_t t2 = t1 & (x >> 2);
    const uint64_t t3 = t2 & (x >> 3);
    const uint64_t quad = t3 & 0x0202020202020202ull;   // bits 1..4
    c[0] += __popcll(x);
    c[1] += __popcll(t1 & 0x4a4a4a4a4a4a4a4aull);   // (11,13) (17,19) (29,31)
    c[2] += __popcll(t2 & 0x0f0f0f0f0f0f0f0full);   // 3 consecutive bits from 0..3
    // Every quintuplet and sextuplet pattern contains the quadruplet bits
    // 1..4 of its byte, and quadruplets are rare: the last three counts are
    // skipped (per wave, when no lane has one) for words without one.
    if (quad) {
      const uint64_t t4 = t3 & (x >> 4);
      const uint64_t t5 = t4 & (x >> 5);
      c[3] += __popcll(quad);
      c[4] += __popcll(t4 & 0x0303030303030303ull);   // bits 0..4, 1..5
      c[5] += __popcll(t5 & 0x0101010101010101ull);   // bits 0..5
    }
  };
  // the whole segment inside the interval: no edge masks
  const bool inside = B > run.gLo && B + S <= run.gHi;
  const auto count = [&](const uint4& v, uint32_t r) {   // chunk r: 64-bit words 2r, 2r+1
    uint64_t x0 = ~((uint64_t) v.x | ((uint64_t) v.y << 32));
    uint64_t x1 = ~((uint64_t) v.z | ((uint64_t) v.w << 32));
    if (!inside) {
      const uint64_t g0 = B + 16ull * r, g1 = g0 + 8u;
      if (g0 <= run.gLo || g0 + 7 >= run.gHi) x0 = mask_edges(x0, g0, run);
      if (g1 <= run.gLo || g1 + 7 >= run.gHi) x1 = mask_edges(x1, g1, run);
    }
    add(x0);
    add(x1);
  };
  uint4* s4 = (uint4*) sieve;
  const uint32_t chunks = S >> 4;
  uint32_t r = tid;
  for (; r + nthreads < chunks; r += 2u * nthreads) {
    const uint4 a = s4[r], b = s4[r + nthreads];
    uint4 fa = make_uint4(0u, 0u, 0u, 0u), fb = fa;
    if (fill) {
      fa = next_run();
      fb = next_run();
    }
    count(a, r);
    count(b, r + nthreads);
    if (hasNext) {
      s4[r] = fa;
      s4[r + nthreads] = fb;
    }
  }
  for (; r < chunks; r += nthreads) {
    const uint4 a = s4[r];
    const uint4 fa = fill ? next_run() : make_uint4(0u, 0u, 0u, 0u);
    count(a, r);
    if (hasNext) s4[r] = fa;
  }
  if (fill) restore_tile_primes(sieve, nB);
}

// Last index i in [0, n) with key(i) <= x (keys ascending, key(0) <= x).
template <typename Key>
__device__ __forceinline__ uint32_t last_le(uint32_t n, uint32_t x, Key key)
{
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1u) >> 1;
    if (key(mid) <= x) lo = mid; else hi = mid - 1u;
  }
  return lo;
}

// Residues r = (v - 7) mod 30 of the 8 bit positions {0,4,6,10,12,16,22,24},
// mapped to their bit index: 3 bits per residue, r = 0..19 and r = 20..29.
constexpr uint64_t kBitLo = (1ull << 12) | (2ull << 18) | (3ull << 30) | (4ull << 36) | (5ull << 48);
constexpr uint64_t kBitHi = (6ull << 6) | (7ull << 12);

__device__ __forceinline__ uint32_t bit_of_residue(uint32_t r)
{
  return (uint32_t) ((r < 20u ? (kBitLo >> (3u * r)) : (kBitHi >> (3u * (r - 20u)))) & 7u);
}

// Primes a thread has loaded ahead of the one it walks: the current group
// (resident; taken in order, shifted down at every take) and the next group
// of kGroup (in flight).  The next group's loads are issued when the current
// group is refilled from it, so they have kGroup takes to arrive.  The refill
// copies them with DPP moves: a plain copy would be deferred to the end of the
// refill block, after the next loads, which would then land in temporaries
// copied into place at once -- a wait on the loads just issued at every refill
// (and, with a rotating queue of single primes, at every take).
constexpr uint32_t kGroup = 4;

struct BucketWalk {
  uint32_t k;        // index of the next group's first prime
  uint32_t q0, q1, q2, q3;  // the current group, next prime first (0: none left)
  uint32_t r0, r1, r2, r3;  // the next group, in flight
  uint32_t t;        // primes of the current group taken (uniform over the wave)
  uint32_t p;        // prime being walked (0: none)
  uint32_t R;        // dense: window base mod p
  uint32_t j;        // dense: current class; 8 marks a sparse prime
  uint32_t pos;      // dense: next byte of class j, relative to the window
  uint32_t qm;       // sparse: cofactor of the next multiple, mod 30
  uint32_t off;      // sparse: next multiple minus the window's first value (< 2^31)
  uint32_t self;     // sparse: offset of p itself when it lies in the window, else ~0
};

// The thread's next prime (0: none left), not yet taken.
__device__ __forceinline__ uint32_t walk_peek(const BucketWalk& s)
{
  return s.q0;
}

// Load a group: primes k + i * stride, the index clamped to end - 1 (end >= 1;
// unpredicated loads, so that they can write their registers directly; the
// primes past end are zeroed when the group moves in).
__device__ __forceinline__ void walk_load(uint32_t& g0, uint32_t& g1, uint32_t& g2, uint32_t& g3,
                                          const uint32_t* primes, uint32_t k, uint32_t stride,
                                          uint32_t end)
{
  const auto ld = [&](uint32_t i) { return primes[min(k + i * stride, end - 1u)]; };
  g0 = ld(0);
  g1 = ld(1);
  g2 = ld(2);
  g3 = ld(3);
}

// A register copy the compiler keeps where it is written (a DPP move with the
// identity lane pattern).
__device__ __forceinline__ uint32_t copy_now(uint32_t x)
{
  return (uint32_t) __builtin_amdgcn_update_dpp(0, (int) x, 0xE4, 0xF, 0xF, false);
}

// The thread's primes are first, first + stride, first + 2 stride, ...
__device__ __forceinline__ void walk_init(BucketWalk& s, const uint32_t* primes, uint32_t first,
                                          uint32_t stride, uint32_t end)
{
  walk_load(s.q0, s.q1, s.q2, s.q3, primes, first, stride, end);
  s.q0 = (first < end) ? s.q0 : 0u;
  s.q1 = (first + stride < end) ? s.q1 : 0u;
  s.q2 = (first + 2u * stride < end) ? s.q2 : 0u;
  s.q3 = (first + 3u * stride < end) ? s.q3 : 0u;
  s.k = first + kGroup * stride;
  walk_load(s.r0, s.r1, s.r2, s.r3, primes, s.k, stride, end);
  s.t = 0u;
}

// Step past the next prime (all lanes of a wave together, so t stays
// uniform); when the current group is used up the next one moves in and the
// group after it is loaded.
__device__ __forceinline__ void walk_advance(BucketWalk& s, const uint32_t* primes, uint32_t stride,
                                             uint32_t end)
{
  if (++s.t == kGroup) {
    s.t = 0u;
    s.q0 = (s.k < end) ? copy_now(s.r0) : 0u;
    s.q1 = (s.k + stride < end) ? copy_now(s.r1) : 0u;
    s.q2 = (s.k + 2u * stride < end) ? copy_now(s.r2) : 0u;
    s.q3 = (s.k + 3u * stride < end) ? copy_now(s.r3) : 0u;
    // nothing of the next loads (their zeroing included) moves above the copies
    __builtin_amdgcn_sched_barrier(0);
    s.k += kGroup * stride;
    walk_load(s.r0, s.r1, s.r2, s.r3, primes, s.k, stride, end);
  } else {
    s.q0 = s.q1;
    s.q1 = s.q2;
    s.q2 = s.q3;
  }
}

// A window's first byte and first value as doubles (off by at most 2^6 and
// 2^11), in scalar registers.
struct WindowBase {
  double baseF;
  double v0F;
};

__device__ __forceinline__ WindowBase window_base(const PsgWindow& w)
{
  return {uniform_f64(to_double(w.base)), uniform_f64(to_double(30ull * w.base + 7u))};
}

// emit(slot, byte * 8 + bit) for the stage's hits in order, with byte and
// bit within the window (so the window segment is the hit >> (segShift + 3))
// and slot the hit's LDS stage word: a thread's slots are PSG_BUCKET_THREADS
// words apart, from slot0.  Returns the number of hits of this stage.
template <typename Emit>
__device__ __forceinline__ uint32_t bucket_walk_stage(const PsgBucketArgs& a, const PsgWindow& w,
                                                      const WindowBase& wb, BucketWalk& s,
                                                      uint32_t stride, const uint32_t* ctab,
                                                      lds_u32* slot0, uint32_t limit, Emit emit)
{
  constexpr uint32_t kStep = PSG_BUCKET_THREADS;
  const auto* __restrict__ primes = gptr<uint32_t>(a.primes);
  const uint32_t bytes = w.segs << a.segShift;
  const double baseF = wb.baseF;
  const uint32_t V = 30u * bytes;                 // values covered by the window (< 2^31)
  const uint64_t v0 = 30ull * w.base + 7u;        // its first value (< 2^64 - 8)
  const double v0F = wb.v0F;
  lds_u32* slot = slot0;
  lds_u32* const slotEnd = slot0 + limit * kStep;
  while (slot < slotEnd) {
    // The lanes of a wave take their next primes together, once every active
    // lane has finished its current one: the set-up (remainder, first
    // multiple or class start) then runs once per prime for the whole wave
    // instead of interleaving with the hit loops of other lanes.
    if (__all(s.p == 0u)) {
      const uint32_t p = walk_peek(s);
      if (__all(p == 0u)) break;
      walk_advance(s, primes, stride, w.bigEnd);
      if (p == 0u) continue;                  // this lane is done; others take primes
      const double rp = recip(p);
      s.p = p;
      if ((uint64_t) p * 8u > V) {
        // The first multiple of p at or after v0, from one quotient:
        // q = floor(v0F * rp) is within one of v0 / p (v0 / p < 2^48, and the
        // product's relative error is below 2^-50), r = v0 - q p exactly in
        // 64-bit integers, corrected into [0, p); the multiple is m p with
        // cofactor m = q + (r != 0), at offset m p - v0 = p - r (or 0).
        const double qf = floor(v0F * rp);
        const double qh = floor(qf * 0x1p-32);
        const uint32_t qlo = (uint32_t) (qf - qh * 0x1p32);   // exact: qf < 2^53
        const uint32_t qhi = (uint32_t) qh;
        int64_t r = (int64_t) (v0 - (uint64_t) qlo * p - ((uint64_t) (qhi * p) << 32));
        double m = qf;
        if (r < 0) { r += p; m -= 1.0; }
        if (r >= (int64_t) p) { r -= p; m += 1.0; }
        if (r) m += 1.0;
        s.off = r ? p - (uint32_t) r : 0u;
        // m mod 30: m * (1/30) is within 2^-7 of m / 30 (m < 2^49), so the
        // floor is exact except that a multiple of 30 may give 30 instead of 0
        const uint32_t qm = (uint32_t) fma(floor(m * (1.0 / 30.0)), -30.0, m);
        s.qm = (qm == 30u) ? 0u : qm;
        s.self = (p >= v0 && p - v0 < V) ? (uint32_t) (p - v0) : 0xffffffffu;
        s.j = 8u;
      } else {
        s.R = mod_big(w.base, baseF, p, rp);                 // bucket primes are > 2^16
        s.j = 0u;
        s.pos = class_start(p, s.R, w.base, 0u, ctab);
      }
    }
    if (s.p == 0u) continue;                  // idle until the wave's next primes
    if (s.j == 8u) {
      while (slot < slotEnd && s.off < V) {
        if (((kCoprime30 >> s.qm) & 1u) && s.off != s.self) {
          const uint32_t byte = s.off / 30u;
          const uint32_t bit = bit_of_residue(s.off - 30u * byte);
          emit(slot, (byte << 3) | bit);
          slot += kStep;
        }
        s.qm = (s.qm == 29u) ? 0u : s.qm + 1u;
        s.off = (V - s.off > s.p) ? s.off + s.p : V;       // no 32-bit overflow
      }
      if (s.off >= V) s.p = 0u;
    } else {
      if (s.pos < bytes) {
        // one exit test per hit: stop at the window's end or when the stage is
        // full.  The loop steps the staged value byte * 8 + class by 8p
        // (dense primes have 8p <= 30 * bytes < 2^31; bytes <= 2^25).
        const uint32_t room = (uint32_t) (slotEnd - slot) / kStep;
        const uint32_t stop = ((uint64_t) room * s.p >= bytes - s.pos) ? bytes : s.pos + room * s.p;
        uint32_t v = (s.pos << 3) | s.j;
        const uint32_t v1 = stop << 3, p8 = s.p << 3;
        for (; v < v1; v += p8, slot += kStep) emit(slot, v);
        s.pos = v >> 3;
      }
      if (s.pos >= bytes) {
        if (++s.j == 8u) s.p = 0u;
        else s.pos = class_start(s.p, s.R, w.base, s.j, ctab);
      }
    }
  }
  return (uint32_t) (slot - slot0) / kStep;
}

// The window of fill work item `item`: wi on entry is the window of the
// workgroup's previous item (w, or ~0 before the first), and items only grow,
// so the window is w itself or is searched for after it.
__device__ __forceinline__ void window_of_item(const PsgBucketArgs& a, uint32_t item, uint32_t& wi,
                                               PsgWindow& w)
{
  if (wi != 0xffffffffu && item < w.wgBase + w.nChunks) return;
  const auto* win = gptr<PsgWindow>(a.windows);
  const uint32_t lo = wi + 1u;                     // 0 before the first item
  wi = lo + last_le(a.nWindows - lo, item, [&](uint32_t i) { return win[lo + i].wgBase; });
  w = win[wi];
}

}  // namespace

// The scratch memory of count calls that must not allocate: the host
// resolves its address once per loaded code object (hipModuleGetGlobal) and
// carves the call's descriptors, counters and bucket lists out of it.  It
// has external linkage so that the module loader can find it by name.
extern "C" {
__device__ __attribute__((used)) __align__(256) uint32_t psg_scratch[PSG_SCRATCH_BYTES / 4];
}

// Zero the ranges of a PsgClearArgs block (counts, work counters, bucket
// list counters, flags) before a slab series -- a kernel rather than
// memset calls, so that a captured launch series holds kernel nodes only.
extern "C" __global__ void __launch_bounds__(256)
psg_clear(uint64_t argsAddr)
{
  const PsgClearArgs c = load_args<PsgClearArgs>(argsAddr);
  const uint32_t stride = gridDim.x * blockDim.x;
#pragma unroll
  for (uint32_t r = 0; r < PSG_CLEAR_RANGES; r++) {
    auto* p = gptr_rw<uint32_t>(c.addr[r]);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < c.words[r]; i += stride)
      p[i] = 0u;
  }
}

// Persistent segment kernel: each workgroup takes slab segments from
// *workCounter (work stealing) until all nSegs are done.  Its argument
// block is in device memory (PsgSieveArgs at argsAddr).
extern "C" __global__ void __launch_bounds__(1024)
psg_sieve_count(uint64_t argsAddr)
{
  const PsgSieveArgs args = load_args<PsgSieveArgs>(argsAddr);
  // the sieve starts at LDS address 0 (strike addresses need no base); the
  // class table and the block's small variables follow it
  extern __shared__ uint32_t lds[];
  uint32_t* sieve = lds;
  uint32_t* ctab = lds + (args.segBytes >> 2);                 // 16-byte aligned
  uint32_t* wsum = ctab + kCtabWords;                           // PSG_NCOUNTS
  uint32_t& sSeg = wsum[PSG_NCOUNTS];
  uint32_t& sRun = wsum[PSG_NCOUNTS + 1];
  // the small pre-sieve tiles, copied once per workgroup when they fit
  uint32_t* small = lds + ((args.segBytes + PSG_SIEVE_LDS_TAIL) >> 2);
  if (args.tilesInLds) {
    constexpr uint32_t kOffset[PSG_NTILES] = PSG_TILE_OFFSETS;
    const auto* tiles = gptr<uint32_t>(args.tiles);
    for (uint32_t i = threadIdx.x; i < PSG_SMALL_TILE_WORDS; i += blockDim.x)
      small[i] = tiles[kOffset[1] + i];
  }

  const auto* runs = gptr<PsgRun>(args.runs);
  init_ctab(ctab);
  init_ctabr(ctab + kCtabWords + kVarWords + 3 * kTier0Slots);
  init_ctabq((uint2*) (ctab + kCtabWords + kVarWords + 3 * kTier0Slots + kCtabrWords));
  {
    uint32_t* t0p = ctab + kCtabWords + kVarWords;
    uint64_t* t0m = (uint64_t*) (t0p + kTier0Slots);
    const auto* pr = gptr<uint32_t>(args.primes);
    const auto* mg = gptr<uint64_t>(args.magics);
    const uint32_t n0 = min(args.tier0End, kTier0Slots);
    for (uint32_t i = threadIdx.x; i < kTier0Slots; i += blockDim.x) {
      t0p[i] = (i < n0) ? pr[i] : 0u;
      t0m[i] = (i < n0) ? mg[i] : 0ull;
    }
  }
  // Per segment: [strike] barrier [count this one + fill the next] barrier
  // [publish the counts].  The first segment is filled on its own.
  uint32_t* sWork = wsum + PSG_NCOUNTS + 2;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63u;
  const auto take = [&]() {                        // thread 0: the next segment and its run
    const uint32_t seg = atomicAdd((uint32_t*) args.workCounter, 1u);
    sSeg = seg;
    sRun = (seg < args.nSegs)
         ? last_le(args.nRuns, seg, [&](uint32_t i) { return runs[i].segStart; }) : 0u;
  };
  if (tid == 0) {
    take();
    *sWork = 0u;
  }
  if (tid < PSG_NCOUNTS) wsum[tid] = 0;
  __syncthreads();
  uint32_t seg = sSeg;
  if (seg >= args.nSegs) return;
  PsgRun run = load_run(runs, sRun);
  uint64_t B = run.B0 + (uint64_t) (seg - run.segStart) * args.segBytes;
  bool filled = false;                             // by the previous segment's count pass
  for (;;) {
    if (!filled) {
      fill_segment(args, B, sieve, small);
      __syncthreads();
    }
    strike_segment(args, run, seg, B, sieve, ctab, wsum, take);   // thread 0 takes the next one
    __syncthreads();                               // strikes done, next segment known
    const uint32_t next = sSeg;
    const bool hasNext = next < args.nSegs;
    PsgRun nrun = run;
    uint64_t nB = 0;
    if (hasNext) {
      nrun = load_run(runs, sRun);
      nB = nrun.B0 + (uint64_t) (next - nrun.segStart) * args.segBytes;
    }
    if (tid == 0) *sWork = 0u;                     // every wave has left the queue
    if (args.primesOut) {
      // list mode: write the primes; the next segment is filled at the top
      list_primes(args, run, B, sieve, wsum + PSG_NCOUNTS + 3);
      __syncthreads();
      filled = false;
    } else {
      uint32_t c[PSG_NCOUNTS] = {0u, 0u, 0u, 0u, 0u, 0u};
      count_and_fill(args, run, B, hasNext, nB, sieve, small, c);
#pragma unroll
      for (int t = 0; t < PSG_NCOUNTS; t++) c[t] = wave_sum(c[t]);
      if (lane == 0) {
#pragma unroll
        for (int t = 0; t < PSG_NCOUNTS; t++) atomicAdd(&wsum[t], c[t]);
      }
      __syncthreads();                             // counts summed, next segment filled
      if (tid < PSG_NCOUNTS) {
        if (wsum[tid])
          atomicAdd((unsigned long long*) args.counts + (uint64_t) run.out * PSG_NCOUNTS + tid,
                    (unsigned long long) wsum[tid]);
        wsum[tid] = 0;                             // summed again after the next strikes
      }
      filled = true;
    }
    if (!hasNext) break;
    seg = next;
    run = nrun;
    B = nB;
  }
}

// ---------------------------------------------------------------------------
// Big-prime bucket tier (see PsgBucketArgs)
// ---------------------------------------------------------------------------

constexpr uint32_t kStageF = PSG_BUCKET_HITS;   // hits per thread per stage

// Persistent: the workgroups take (window, chunk) work items from
// *workCounter until all nItems are done.  One pass per item, in stages.
// In a stage every thread walks up to kStageF hits into its LDS stage slots
// while an LDS histogram counts them per window segment.  Then each
// segment's run of the stage takes its place in the segment's list with one
// global atomic (issued early, used late), the hits are sorted by segment in
// place (every thread first holds its hits in registers), and the sorted
// runs are written out contiguously: coalesced stores instead of one
// scattered store per hit.  Its argument block is in device memory
// (PsgBucketArgs at argsAddr).
extern "C" __global__ void __launch_bounds__(PSG_BUCKET_THREADS) __attribute__((amdgpu_waves_per_eu(4)))
psg_bucket_fill(uint64_t argsAddr)
{
  const PsgBucketArgs a = load_args<PsgBucketArgs>(argsAddr);
  extern __shared__ uint32_t lds[];
  const uint32_t tid = threadIdx.x;
  const uint32_t nthreads = PSG_BUCKET_THREADS;        // the launch's block size
  uint32_t* stage = lds;                               // [kStageF][nthreads], then sorted
  // [W][C] counts per segment and copy (copy = thread mod C), then scatter cursors
  uint32_t* hist = stage + PSG_BUCKET_STAGE;
  // [W] end of the segment's written slots (first sorted slot + fit) | (the
  // list index of its first slot minus that slot, mod 2^32) << 32: sorted slot
  // i < end goes to list index base + i
  uint64_t* place = (uint64_t*) (hist + PSG_BUCKET_HIST_WORDS);
  __shared__ __align__(16) uint32_t ctab[240];
  __shared__ uint32_t waveTotal[PSG_BUCKET_THREADS / 64u];
  __shared__ uint32_t sItem;

  init_ctab(ctab);
  auto* entries = gptr_rw<uint32_t>(a.entries);
  const uint32_t lane = tid & 63u, wave = tid >> 6, nwaves = nthreads >> 6;
  const uint32_t eshift = a.segShift + 3u;             // window bit index -> window segment
  const uint32_t segBits = (1u << eshift) - 1u;        // window bit index -> segment bit index
  uint64_t appended = 0;
  uint32_t wi = 0xffffffffu;                           // window of the previous item
  PsgWindow w{};
  for (;;) {
    // The next item.  Every thread has read the previous item's number before
    // thread 0 overwrites it: the item's stages end with a barrier.
    if (tid == 0) sItem = atomicAdd((uint32_t*) a.workCounter, 1u);
    __syncthreads();
    const uint32_t item = sItem;
    if (item >= a.nItems) break;
    window_of_item(a, item, wi, w);
    const uint32_t c = item - w.wgBase;
    const uint32_t W = w.segs;                           // <= PSG_MAX_WINDOW_SEGS <= blockDim
    const auto* __restrict__ listOff = gptr<uint32_t>(a.listOff) + w.firstSeg;
    uint32_t* listCount = (uint32_t*) a.listCount + w.firstSeg;
    // C = 2^cs histogram copies per segment (uniform over the item)
    const uint32_t cs = (W <= PSG_BUCKET_HIST_WORDS / 4u) ? 2u : 1u;
    const uint32_t copy = tid & ((1u << cs) - 1u);
    // this thread's histogram copy as an LDS byte address; segment s's
    // counter is at histc + (s << cs2)
    const uint32_t histc = (uint32_t) (uintptr_t) (lds_u32*) (hist + copy);
    const uint32_t cs2 = cs + 2u;
    const uint32_t HW = W << cs;
    for (uint32_t i = tid; i < HW; i += nthreads) hist[i] = 0;

    BucketWalk st;
    const uint32_t stride = nthreads * w.nChunks;
    walk_init(st, gptr<uint32_t>(a.primes), a.bigBegin + c * nthreads + tid, stride, w.bigEnd);
    st.p = 0u; st.R = 0u; st.j = 0u; st.pos = 0u; st.qm = 0u; st.off = 0u; st.self = 0u;
    const WindowBase wb = window_base(w);
    __syncthreads();
    for (;;) {
      const uint32_t n = bucket_walk_stage(a, w, wb, st, stride, ctab, (lds_u32*) (stage + tid), kStageF,
          [&](lds_u32* slot, uint32_t hit) {
            *slot = hit;
            // the hit's window segment (< 256) and its counter in this copy
            const uint32_t seg = (hit >> eshift) & (PSG_MAX_WINDOW_SEGS - 1u);
            __hip_atomic_fetch_add((lds_u32*) (histc + (seg << cs2)), 1u, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WORKGROUP);
          });
      const int more = __syncthreads_or((st.p != 0u || walk_peek(st) != 0u) ? 1 : 0);

      // exclusive prefix of the counts; thread i owns segment i and its copies
      uint32_t cc[4] = {0u, 0u, 0u, 0u};
      if (tid < W) {
        if (cs == 2u) {
          const uint4 h4 = *(const uint4*) &hist[tid << 2];
          cc[0] = h4.x; cc[1] = h4.y; cc[2] = h4.z; cc[3] = h4.w;
        } else {
          const uint2 h2 = *(const uint2*) &hist[tid << 1];
          cc[0] = h2.x; cc[1] = h2.y;
        }
      }
      const uint32_t cnt = cc[0] + cc[1] + cc[2] + cc[3];
      uint32_t x = cnt;
      for (uint32_t d = 1; d < 64u; d <<= 1) {
        const uint32_t y = (uint32_t) __shfl_up((int) x, d);
        if (lane >= d) x += y;
      }
      if (lane == 63u) waveTotal[wave] = x;
      __syncthreads();
      uint32_t before = 0, total = 0;
      for (uint32_t v = 0; v < nwaves; v++) {
        const uint32_t t = waveTotal[v];
        before += (v < wave) ? t : 0u;
        total += t;
      }
      const uint32_t f = before + x - cnt;
      uint32_t at = 0;
      if (tid < W) {
        // each copy's cursor starts after the segment's earlier copies
        if (cs == 2u)
          *(uint4*) &hist[tid << 2] = make_uint4(f, f + cc[0], f + cc[0] + cc[1], f + cc[0] + cc[1] + cc[2]);
        else
          *(uint2*) &hist[tid << 1] = make_uint2(f, f + cc[0]);
        if (cnt) at = atomicAdd(&listCount[tid], cnt);   // consumed after the scatter
      }
      // The stage passes below cover only the slots in use: a wave reads
      // back, and scatters, the groups of eight in which any of its threads
      // staged a hit, and the write-out stops at the stage's total.  The last
      // stage of an item is often far from full, and on the batch of short
      // intervals such stages are most of the work.
      uint32_t e[kStageF];
#pragma unroll
      for (uint32_t g = 0; g < kStageF; g += 8) {
        if (__any(g < n)) {
#pragma unroll
          for (uint32_t q = 0; q < 8; q++) e[g + q] = stage[(g + q) * nthreads + tid];
        } else {
#pragma unroll
          for (uint32_t q = 0; q < 8; q++) e[g + q] = 0u;
        }
      }
      __syncthreads();                                   // cursors set, stage read

      // Scatter by segment, eight hits at a time: their cursor atomics first,
      // then their stores, so that eight atomics are in flight at once.  A
      // wave whose threads all staged a full stage (every stage of an item
      // but its last) skips the per-hit test.
      const auto scatter = [&](auto full) {
#pragma unroll
        for (uint32_t g = 0; g < kStageF; g += 8) {
          if (!full.value && !__any(g < n)) break;
          uint32_t slot[8];
#pragma unroll
          for (uint32_t q = 0; q < 8; q++)
            slot[q] = (full.value || g + q < n) ? atomicAdd(&hist[((e[g + q] >> eshift) << cs) | copy], 1u) : 0u;
#pragma unroll
          for (uint32_t q = 0; q < 8; q++)
            if (full.value || g + q < n) stage[slot[q]] = e[g + q];
        }
      };
      if (__all(n == kStageF)) scatter(std::true_type{});
      else scatter(std::false_type{});
      if (tid < W) {
        const uint32_t room = listOff[tid + 1u] - listOff[tid];
        if ((uint64_t) at + cnt > room) atomicOr((uint32_t*) a.overflow, 1u);
        const uint32_t fit = (at >= room) ? 0u : min(cnt, room - at);
        place[tid] = (uint64_t) (f + fit) | ((uint64_t) (listOff[tid] + at - f) << 32);
      }
      __syncthreads();

      // Write-out, eight slots at a time: their stage reads, then their
      // segments' placements, then the stores.  Slot i holds an entry of its
      // segment's run (which starts at or before i), so it is written when
      // it lies below the run's end.  Slots past `total` hold stale values;
      // their segment index is masked into range and they store nothing.
#pragma unroll
      for (uint32_t g = 0; g < kStageF; g += 8) {
        if (g * nthreads >= total) break;                // total is uniform
        uint32_t v[8];
        uint64_t pl[8];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) v[q] = stage[(g + q) * nthreads + tid];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) pl[q] = place[(v[q] >> eshift) & (PSG_MAX_WINDOW_SEGS - 1u)];
#pragma unroll
        for (uint32_t q = 0; q < 8; q++) {
          const uint32_t i = (g + q) * nthreads + tid;
          const uint32_t end = (uint32_t) pl[q], base = (uint32_t) (pl[q] >> 32);
          if (i < min(total, end)) entries[base + i] = v[q] & segBits;
        }
      }
      appended += total;
      for (uint32_t i = tid; i < HW; i += nthreads) hist[i] = 0;
      __syncthreads();
      if (!more) break;
    }
  }
  if (tid == 0 && appended)
    atomicAdd((unsigned long long*) a.total, (unsigned long long) appended);
}

// Division magics of a prime list (the device-resident sieving-prime cache):
// M = ceil(2^(64+L) / p), L = floor(log2 p), for odd primes p < 2^32, by
// long division in two 32-bit digits: 2^(32+L) = q1 * p + r (q1 < 2^32,
// since 2^L < p), then r * 2^32 = q0 * p + r'.  p does not divide 2^(64+L),
// so the ceiling is the floor plus one.
extern "C" __global__ void __launch_bounds__(256)
psg_prime_magics(uint64_t primes, uint64_t magics, uint32_t n)
{
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = gptr<uint32_t>(primes)[i];
  const uint32_t L = 31u - (uint32_t) __clz((int) p);
  const uint64_t num = 1ull << (32u + L);
  const uint64_t q1 = num / p;
  const uint64_t r = num - q1 * p;
  const uint64_t q0 = (r << 32) / p;
  gptr_rw<uint64_t>(magics)[i] = (q1 << 32) + q0 + 1u;
}
