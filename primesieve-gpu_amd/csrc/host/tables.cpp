///
/// @file  tables.cpp
/// @brief See tables.hpp.
///

#include "tables.hpp"

#include <primesieve/primesieve_error.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

namespace psg {

const uint32_t kBitOffset[8] = { 7, 11, 13, 17, 19, 23, 29, 31 };

uint64_t isqrt64(uint64_t n)
{
  uint64_t r = (uint64_t) std::sqrt((double) n);
  // correct the floating-point estimate exactly
  while (r > 0 && (r > 0xffffffffull || r * r > n)) r--;
  while (r < 0xffffffffull && (r + 1) * (r + 1) <= n) r++;
  return r;
}

ByteRange byteRange(uint64_t lo, uint64_t hi)
{
  ByteRange r;
  const uint64_t remLo = byteRemainder(lo);
  const uint64_t remHi = byteRemainder(hi);
  r.gLo = (lo - remLo) / 30;
  r.gHi = (hi - remHi) / 30;
  for (int b = 0; b < 8; b++)
  {
    if (kBitOffset[b] >= remLo) r.loMask |= 1u << b;
    if (kBitOffset[b] <= remHi) r.hiMask |= 1u << b;
  }
  return r;
}

uint64_t divisionMagic(uint32_t p)
{
  if (p < 3)
    throw primesieve::primesieve_error("divisionMagic: p must be >= 3");
  const uint32_t L = 31 - (uint32_t) __builtin_clz(p);
  const unsigned __int128 num = (unsigned __int128) 1 << (64 + L);
  unsigned __int128 q = num / p;
  if (q * p != num) q++;
  return (uint64_t) q;
}

std::vector<uint32_t> tileWindows(const std::vector<uint32_t>& primes)
{
  uint64_t period = 1;
  for (uint32_t p : primes)
  {
    if (p < 7 || p % 2 == 0 || p % 3 == 0 || p % 5 == 0)
      throw primesieve::primesieve_error("tile primes must be >= 7 and coprime to 30");
    period *= p;
  }
  std::vector<uint8_t> pattern((size_t) period, 0xff);
  for (uint64_t i = 0; i < period; i++)
    for (int b = 0; b < 8; b++)
    {
      const uint64_t v = 30 * i + kBitOffset[b];
      for (uint32_t p : primes)
        if (v % p == 0) { pattern[(size_t) i] &= (uint8_t) ~(1u << b); break; }
    }
  std::vector<uint32_t> win((size_t) period);
  for (uint64_t i = 0; i < period; i++)
  {
    uint32_t w = 0;
    for (int k = 0; k < 4; k++)
      w |= (uint32_t) pattern[(size_t) ((i + k) % period)] << (8 * k);
    win[(size_t) i] = w;
  }
  return win;
}

std::vector<uint32_t> primesBetween(uint32_t lo, uint32_t hi)
{
  std::vector<uint32_t> out;
  if (hi < 2 || lo > hi)
    return out;
  if (lo <= 2 && hi >= 2)
    out.push_back(2);

  // base primes up to sqrt(hi) < 2^16
  const uint32_t sq = (uint32_t) isqrt64(hi);
  std::vector<uint8_t> small(sq + 1, 1);
  std::vector<uint32_t> base;
  for (uint32_t i = 3; i <= sq; i += 2)
    if (small[i])
    {
      base.push_back(i);
      for (uint64_t j = (uint64_t) i * i; j <= sq; j += 2 * i)
        small[(size_t) j] = 0;
    }

  // odd numbers only: index k <-> 2k+1
  const uint64_t kFirst = std::max<uint64_t>(lo, 3) / 2;
  const uint64_t kLast = (hi - 1) / 2;
  const uint64_t segLen = 1 << 18;
  std::vector<uint8_t> seg(segLen);
  for (uint64_t k0 = kFirst; k0 <= kLast; k0 += segLen)
  {
    const uint64_t n = std::min<uint64_t>(segLen, kLast - k0 + 1);
    std::fill(seg.begin(), seg.begin() + (ptrdiff_t) n, 1);
    const uint64_t vLo = 2 * k0 + 1;
    const uint64_t vHi = 2 * (k0 + n - 1) + 1;
    for (uint32_t p : base)
    {
      const uint64_t pp = (uint64_t) p * p;
      if (pp > vHi) break;
      uint64_t m = std::max<uint64_t>(pp, (vLo + p - 1) / p * p);
      if (m % 2 == 0) m += p;
      for (; m <= vHi; m += 2ull * p)
        seg[(size_t) ((m - 1) / 2 - k0)] = 0;
    }
    for (uint64_t i = 0; i < n; i++)
      if (seg[(size_t) i])
      {
        const uint64_t v = 2 * (k0 + i) + 1;
        if (v >= lo && v > 1) out.push_back((uint32_t) v);
      }
  }
  return out;
}

namespace {

uint64_t coprime30UpTo(uint64_t n)
{
  // #{1 <= q <= n : gcd(q, 30) = 1}
  static const uint8_t R[30] = { 0, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 4, 4, 4,
                                 4, 5, 5, 6, 6, 6, 6, 7, 7, 7, 7, 7, 7, 8 };
  return 8 * (n / 30) + R[n % 30];
}

} // namespace

uint64_t crossings(const std::vector<uint32_t>& primes, uint64_t lo, uint64_t hi,
                   uint32_t pmin, uint32_t pmax)
{
  if (lo > hi) return 0;
  uint64_t c = 0;
  for (uint32_t p32 : primes)
  {
    const uint64_t p = p32;
    if (p < pmin || p > pmax) continue;
    if (p * p > hi) break;
    uint64_t qa = lo / p + (lo % p != 0);
    if (qa < p) qa = p;
    const uint64_t qb = hi / p;
    if (qb >= qa)
      c += coprime30UpTo(qb) - coprime30UpTo(qa - 1);
  }
  return c;
}

} // namespace psg
