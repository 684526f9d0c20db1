// psg_bench.hip -- roofline probes for gfx950, used to state the measured
// ceilings that the sieve kernels are compared against.
//
//   psg_lds_and:   every lane issues `iters` LDS atomics into a `words`-word
//                  LDS array: ds_and_b32, or ds_or_b32 (the sieve's strike)
//                  when mode bit 1 is set.  Mode bit 0 clear: lane-consecutive
//                  words (conflict-free); set: a per-lane LCG address (random,
//                  the access pattern of scattered strikes).
//   psg_lds_read:  ds_read_b64 streaming over the LDS array.
//   psg_hbm_copy:  coalesced 16-byte-per-lane copy, src -> dst (HBM stream).
//
// Every kernel ends by writing a value that depends on the LDS contents, so
// no access can be removed by the compiler.
#include <hip/hip_runtime.h>
#include <stdint.h>

template <bool kOr>
__device__ __forceinline__ void lds_rmw(uint32_t* lds, uint32_t mask, uint32_t iters, uint32_t mode)
{
  const uint32_t tid = threadIdx.x;
  if ((mode & 1u) == 0) {
    uint32_t a = tid;
    const uint32_t step = blockDim.x;
#pragma unroll 8
    for (uint32_t i = 0; i < iters; i++) {
      if (kOr) atomicOr(&lds[a & mask], 1u << (i & 31u));
      else atomicAnd(&lds[a & mask], ~(1u << (i & 31u)));
      a += step;
    }
  } else {
    uint32_t x = tid * 2654435761u + blockIdx.x * 40503u + 1u;
#pragma unroll 8
    for (uint32_t i = 0; i < iters; i++) {
      x = x * 1664525u + 1013904223u;
      if (kOr) atomicOr(&lds[(x >> 8) & mask], 1u << (i & 31u));
      else atomicAnd(&lds[(x >> 8) & mask], ~(1u << (i & 31u)));
    }
  }
}

extern "C" __global__ void __launch_bounds__(1024)
psg_lds_and(uint32_t words, uint32_t iters, uint32_t mode, uint32_t* out)
{
  extern __shared__ uint32_t lds[];
  const uint32_t tid = threadIdx.x;
  for (uint32_t w = tid; w < words; w += blockDim.x) lds[w] = 0xffffffffu;
  __syncthreads();
  const uint32_t mask = words - 1;   // words is a power of two
  if (mode & 2u) lds_rmw<true>(lds, mask, iters, mode);
  else lds_rmw<false>(lds, mask, iters, mode);
  __syncthreads();
  if (tid == 0) out[blockIdx.x] = lds[blockIdx.x & mask];
}

extern "C" __global__ void __launch_bounds__(1024)
psg_lds_read(uint32_t words, uint32_t iters, uint32_t* out)
{
  extern __shared__ uint32_t lds[];
  const uint32_t tid = threadIdx.x;
  for (uint32_t w = tid; w < words; w += blockDim.x) lds[w] = w * 2654435761u;
  __syncthreads();
  const uint64_t* l64 = (const uint64_t*) lds;
  const uint32_t mask = (words >> 1) - 1;
  uint64_t acc = 0;
  uint32_t a = tid;
#pragma unroll 8
  for (uint32_t i = 0; i < iters; i++) {
    acc ^= l64[a & mask];
    a += blockDim.x;
  }
  if (acc == 0x1234567ull) out[blockIdx.x] = (uint32_t) acc;
}

extern "C" __global__ void __launch_bounds__(256)
psg_hbm_copy(const uint4* __restrict__ src, uint4* __restrict__ dst, uint64_t n16)
{
  const uint64_t stride = (uint64_t) gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t) blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
    dst[i] = src[i];
}
