/*
 * primesieve_gpu.h -- C ABI of the MI355X (gfx950) HIP counting backend.
 *
 * Plain C types only (no C++, no torch): this is what a ctypes / cgo / JNI /
 * N-API binding, or the reference's own C API, binds.  Every function takes
 * the backend's global lock, like the reference's GPU API
 * (src/gpu/GpuSieve.cpp:27,186-343).
 *
 * Error convention of the reference C API (include/primesieve.h:28-31,
 * src/api-c.cpp:143-225): functions returning uint64_t return PSG_ERROR
 * (~0) and set errno = EDOM on failure, after printing the reason to stderr;
 * int-returning functions return 0 on success / false and a negative value
 * on error, with the reason available from psg_gpu_last_error().
 */
#ifndef PRIMESIEVE_GPU_H
#define PRIMESIEVE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSG_ERROR ((uint64_t) ~((uint64_t) 0))

/* ---- device selection and configuration --------------------------------
 * Replace gpu_available / gpu_devices / set_gpu_device / get_gpu_device /
 * set_gpu_enabled / get_gpu_enabled and the four tuning knobs
 * (reference include/primesieve/gpu.hpp:32-56, src/gpu/GpuSieve.cpp:184-254).
 */
int  psg_gpu_available(void);
int  psg_gpu_device_count(void);
/* Writes DeviceInfo::describe() of device `index` (reference
 * src/gpu/GpuSieve.cpp:43-55) into buf; returns its length or -1. */
int  psg_gpu_device_describe(int index, char* buf, size_t len);
void psg_set_gpu_device(int device_index);
int  psg_get_gpu_device(void);
void psg_set_gpu_enabled(int enable);
int  psg_get_gpu_enabled(void);
void psg_set_gpu_sieve_size(int kilobytes);
void psg_set_gpu_work_group_size(int size);
void psg_set_gpu_segments_per_launch(int segments);
void psg_set_gpu_presieve(int enable);

/* New: the full tuning vector (the reference's four knobs plus the bucket
 * tier's).  0 in a field means "derive from the device"; bucket_factor
 * 0xffffffff disables the bucket tier.  Setting recreates the backend on the
 * next call, like the reference's setters (src/gpu/GpuSieve.cpp:228-254);
 * psg_get_gpu_tuning reports the values the active backend uses. */
typedef struct psg_tuning {
  uint32_t sieve_bytes;       /* LDS bytes per segment (multiple of 256) */
  uint32_t work_group_size;   /* threads per workgroup: 64 * waves */
  uint32_t segments_per_launch;
  uint32_t slab_segments;     /* segments per big-prime bucket slab */
  uint32_t bucket_factor;     /* primes > bucket_factor * sieve_bytes use buckets; 0 = auto (7, batches 12) */
  uint32_t presieve;          /* 1 = pre-sieve tiles for 7..47 */
  uint32_t bucket_margin;     /* room of a segment's bucket list beyond the expected
                                 count, in standard deviations; 0 = auto (6) */
} psg_tuning;
int psg_set_gpu_tuning(const psg_tuning* t);
int psg_get_gpu_tuning(psg_tuning* t);

/* gpu_active_device / gpu_last_error (reference gpu.hpp:58-63): copy the
 * string into buf (truncated, NUL-terminated); return its full length. */
int  psg_gpu_active_device(char* buf, size_t len);
int  psg_gpu_last_error(char* buf, size_t len);

/* gpu_worthwhile (reference src/gpu/GpuSieve.cpp:272-277): stop - start >= 1e8 */
int  psg_gpu_worthwhile(uint64_t start, uint64_t stop);

/* ---- counts -------------------------------------------------------------
 * gpu_count_primes .. gpu_count_sextuplets (reference
 * src/gpu/GpuSieve.cpp:279-337).  PSG_ERROR + errno = EDOM when no GPU is
 * usable or the device fails.
 */
uint64_t psg_gpu_count_primes(uint64_t start, uint64_t stop);
uint64_t psg_gpu_count_twins(uint64_t start, uint64_t stop);
uint64_t psg_gpu_count_triplets(uint64_t start, uint64_t stop);
uint64_t psg_gpu_count_quadruplets(uint64_t start, uint64_t stop);
uint64_t psg_gpu_count_quintuplets(uint64_t start, uint64_t stop);
uint64_t psg_gpu_count_sextuplets(uint64_t start, uint64_t stop);

/* gpu_try_count (reference src/gpu/GpuSieve.cpp:341-365): 1 and *result when
 * the GPU counted; 0 when it is disabled, unavailable, the range is below
 * the 1e8 threshold, or the device failed (the caller then sieves on the
 * CPU).  countType 0..5 = primes..sextuplets.  Never fails loudly. */
int psg_gpu_try_count(uint64_t start, uint64_t stop, int count_type, uint64_t* result);

/* New: all six counts of [start, stop] from one sieve pass.  0 on success,
 * -1 on error (see psg_gpu_last_error). */
int psg_gpu_count_all(uint64_t start, uint64_t stop, uint64_t counts[6]);

/* ---- batches ------------------------------------------------------------
 * New (no reference counterpart; the nearest are the random-interval loops of
 * test/count_primes3.cpp:44-50).  out[i] = count of type k_i over
 * [start_i, stop_i], k = 1 primes, 2 twins, ..., 6 sextuplets.
 * PSG_BATCH_PERSISTENT packs all queries into persistent work-stealing
 * launches; PSG_BATCH_PLAIN runs one count call (launch series + sync) per
 * query, the baseline the batch is compared with; PSG_BATCH_GRAPH is below.
 * 0 on success, -1 on error. */
typedef struct psg_query {
  uint64_t start;
  uint64_t stop;
  uint32_t k;
  uint32_t reserved;
} psg_query;
#define PSG_BATCH_PERSISTENT 0
#define PSG_BATCH_PLAIN 1
/* persistent, through a captured launch series (hipGraph): the series of a
 * batch (clear, then per slab bucket fill and segment kernel) takes all its
 * arguments from one device-resident descriptor upload, so it is captured
 * once per shape (slab count, which slabs fill buckets) and replayed for
 * every later batch of that shape with new descriptors. */
#define PSG_BATCH_GRAPH 2
int psg_gpu_count_batch(const psg_query* queries, size_t n, uint64_t* out, int mode);

/* ---- no device allocation ------------------------------------------------
 * No reference counterpart.  For callers that must not allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc): psg_gpu_count_all with every per-call device buffer
 * (descriptors, counters, bucket lists, counts) carved out of one static
 * array of the backend's code object, psg_scratch, PSG_SCRATCH_BYTES
 * (64 MiB) per device, held while the backend lives.  The call makes no
 * hipMalloc / hipFree: the backend and its sieving-prime cache must already
 * exist (any earlier count call, or psg_gpu_sieving_primes(isqrt(stop)),
 * creates them); otherwise, or when the interval needs more scratch than
 * there is (about 1.6e13 integers at most, fewer at high starts), it returns
 * -1 with the reason in psg_gpu_last_error.  Calls are serialised by the
 * library's lock; overlapping calls on one GPU from two processes each use
 * their own copy of the array.  0 on success. */
int psg_gpu_count_all_noalloc(uint64_t start, uint64_t stop, uint64_t counts[6]);

/* ---- device-resident sieving primes -------------------------------------
 * No reference counterpart: the reference regenerates its sieving primes on
 * the CPU whenever sqrt(stop) grows (src/gpu/cuda/CudaBackend.cpp:346-388).
 * Makes the backend's device cache hold every sieving prime up to `limit`
 * (from 53 with the pre-sieve, else from 7; a cache reaching 2^24 is
 * generated on the GPU: the segment kernel counts and then lists the primes
 * of every segment, and psg_prime_magics adds their division magics) and
 * copies the cached primes <= limit, ascending, and their magics
 * ceil(2^(64+L)/p), L = floor(log2 p), read back from the device, to the
 * caller's arrays (at most cap of each; either may be NULL).  Returns the
 * number of cached primes <= limit, or PSG_ERROR with errno = EDOM. */
uint64_t psg_gpu_sieving_primes(uint32_t limit, uint32_t* primes, uint64_t* magics, uint64_t cap);

/* ---- measurement --------------------------------------------------------
 * No reference counterpart.  Statistics of the last count call, measured
 * with HIP events on the backend's own stream. */
typedef struct psg_stats {
  double   kernel_ms;        /* sum of the segment kernel launch durations */
  double   fill_ms;          /* sum of the big-prime bucket kernel durations */
  double   device_ms;        /* first device event to last */
  double   wall_ms;          /* host wall time of the call */
  uint64_t segments;         /* segments sieved */
  uint64_t bucket_entries;   /* big-prime bucket entries written and read */
  uint32_t launches;         /* sieve kernel launches */
  uint32_t sieve_bytes;      /* LDS bytes per segment */
  uint32_t work_group_size;  /* threads per workgroup */
  uint32_t sieving_primes;   /* sieving primes struck in LDS (from 53; from 7 without pre-sieve) with p*p <= stop */
  uint32_t bucket_retries;   /* calls repeated because a bucket list outgrew its room */
  uint32_t graph;            /* 0 launched one by one (kernel_ms / fill_ms timed per launch);
                                1 replayed a captured launch series, 2 captured it first
                                (then only device_ms is timed) */
} psg_stats;
int psg_gpu_last_stats(psg_stats* out);

/* Algorithmic work of sieving [start, stop] with segments of seg_bytes:
 * crossings by sieving primes p >= 53 (the primes struck in LDS; 7..47 are
 * removed by the pre-sieve tiles), the 32-bit words of the
 * bit array, and the crossings by primes above 8*seg_bytes ("big" primes).
 * Closed form on the host; no device needed. */
typedef struct psg_work {
  uint64_t crossings;        /* C: sum over 53 <= p <= isqrt(stop) */
  uint64_t words32;          /* W = ceil(bytes / 4) */
  uint64_t crossings_big;    /* crossings by p > 8 * seg_bytes */
  uint64_t sieving_primes;   /* primes in [53, isqrt(stop)] */
} psg_work;
int psg_algorithmic_work(uint64_t start, uint64_t stop, uint32_t seg_bytes, psg_work* out);

/* Measured ceilings of the active device (microbenchmark kernels):
 * ds_and_b32 (and ds_or_b32, the sieve's strike) throughput on LDS with
 * conflict-free and with random word
 * addresses (GB/s counted as 8 bytes -- one 4-byte read plus one 4-byte
 * write -- per atomic, the convention of the sieve's LDS roofline), LDS
 * ds_read_b64 streaming, and HBM copy (read + write bytes).  0 on success. */
typedef struct psg_roofline {
  double and_conflict_free_gbs;
  double and_random_gbs;
  double and_conflict_free_gops;   /* 1e9 atomics per second */
  double and_random_gops;
  double lds_read_gbs;
  double hbm_copy_gbs;
  uint32_t compute_units;
  uint32_t reserved;
  double or_conflict_free_gbs;     /* the same for ds_or_b32, the sieve's strike */
  double or_random_gbs;
} psg_roofline;
int psg_measure_roofline(psg_roofline* out);

/* ---- cost model ---------------------------------------------------------
 * No reference counterpart (the reference tunes by grid search only,
 * tools/gpu_check.cpp:164-204).  An FP32 6 -> 64 -> 64 -> 1 MLP (ReLU, ReLU,
 * linear) with hand-written HIP forward, backward and Adam kernels; inputs
 * are 6 normalised features per sample (row-major, n x 6), outputs one float
 * per sample.  Parameters are one flat array of psg_mlp_param_count() floats:
 * W1[64][6], b1[64], W2[64][64], b2[64], W3[64], b3.  Loss: mean squared
 * error.  Adam: beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected.
 * int-returning calls: 0 on success, -1 on error (psg_gpu_last_error). */
typedef struct psg_mlp psg_mlp;
int      psg_mlp_param_count(void);
psg_mlp* psg_mlp_create(void);        /* parameters zero; NULL on error */
void     psg_mlp_destroy(psg_mlp* m);
int      psg_mlp_set_params(psg_mlp* m, const float* params);   /* resets Adam state */
int      psg_mlp_get_params(psg_mlp* m, float* params);
int      psg_mlp_forward(psg_mlp* m, const float* x, size_t n, float* y);
/* gradients of the mean squared error over the n samples, and that loss */
int      psg_mlp_gradients(psg_mlp* m, const float* x, const float* t, size_t n,
                           float* grads, float* loss);
/* one Adam update of the parameters with the given gradients */
int      psg_mlp_adam_step(psg_mlp* m, const float* grads, float lr);
/* minibatch training: `epochs` shuffled passes (seeded), one Adam step per
 * batch; epoch_loss[e] (optional) = mean batch loss of epoch e */
int      psg_mlp_train(psg_mlp* m, const float* x, const float* t, size_t n, uint32_t epochs,
                       uint32_t batch, float lr, uint32_t seed, float* epoch_loss);

#ifdef __cplusplus
}
#endif

#endif
