///
/// @file  tables.hpp
/// @brief Host-side tables and arithmetic of the wheel-30 count path:
///        range geometry (byteRemainder, partial-byte masks), the pre-sieve
///        tiles, division magics, sieving primes and the closed-form work
///        counts.  Everything here is derived from first principles; the
///        reference equivalents are cited at each declaration.
///
#ifndef PSG_TABLES_HPP
#define PSG_TABLES_HPP

#include <stdint.h>
#include <vector>

namespace psg {

/// Offsets of the 8 bits of a byte: bit b of byte g is 30*g + kBitOffset[b]
/// (reference src/gpu/GpuWheel.hpp:125).
extern const uint32_t kBitOffset[8];

/// floor(sqrt(n)) for any 64-bit n.
uint64_t isqrt64(uint64_t n);

/// n mod 30 in the classes 7..36 (reference src/Erat.cpp:231-237); n >= 7.
inline uint64_t byteRemainder(uint64_t n) { return (n - 7) % 30 + 7; }

/// Global byte range and edge masks of the integers [lo, hi], lo >= 7.
/// loMask keeps the bits of byte gLo whose value is >= lo, hiMask the bits of
/// byte gHi whose value is <= hi (reference src/Erat.cpp:30-47 unsetSmaller /
/// unsetLarger, applied at :263-290).
struct ByteRange
{
  uint64_t gLo = 0;
  uint64_t gHi = 0;
  uint32_t loMask = 0;
  uint32_t hiMask = 0;
};
ByteRange byteRange(uint64_t lo, uint64_t hi);

/// ceil(2^(64+L) / p), L = floor(log2 p): floor(x / p) = mulhi(x, M) >> L for
/// every x < 2^63 (reference contract src/gpu/GpuTables.hpp:65-75).
uint64_t divisionMagic(uint32_t p);

/// Sliding-window pre-sieve tile for the given primes (>= 7, coprime to 30):
/// window[i] packs pattern bytes i..i+3 (mod period) little-endian, where the
/// pattern byte at i has bit b clear iff some tile prime divides 30*i + off[b].
/// The period is the product of the primes (in bytes).
std::vector<uint32_t> tileWindows(const std::vector<uint32_t>& primes);

/// All primes in [lo, hi], hi < 2^32, ascending (segmented sieve on the host).
std::vector<uint32_t> primesBetween(uint32_t lo, uint32_t hi);

/// Closed-form crossing count: sum over primes p in `primes` with
/// p*p <= hi of #{q : max(p, ceil(lo/p)) <= q <= floor(hi/p), gcd(q,30)=1}.
uint64_t crossings(const std::vector<uint32_t>& primes, uint64_t lo, uint64_t hi,
                   uint32_t pmin, uint32_t pmax);

} // namespace psg

#endif
