///
/// @file  GpuSieve.hpp
/// @brief Internal hooks of GpuSieve.cpp shared with the C ABI layer.
///
#ifndef PSG_GPU_SIEVE_HPP
#define PSG_GPU_SIEVE_HPP

#include <stdint.h>
#include <mutex>
#include <string>

#include <primesieve/gpu.hpp>

namespace psg { class HipBackend; struct HipTuning; }

namespace primesieve {
namespace gpu {

class GpuBackend;
struct GpuConfig;

/// Replace the whole configuration (the reference's four tunables and the
/// HIP backend's own); the backend is recreated on next use.
void setTuning(const GpuConfig& config, const psg::HipTuning& tuning);

/// The global lock every public GPU entry point holds.
std::mutex& gpuMutex();
/// Backend of the selected device or nullptr (caller holds gpuMutex()).
GpuBackend* activeBackend();
psg::HipBackend* activeHipBackend();
/// The backend if one exists, without creating it (caller holds gpuMutex()).
psg::HipBackend* existingHipBackend();
/// g_lastError (caller holds gpuMutex()).
std::string& lastErrorRef();
/// Drop the backend after a device failure (caller holds gpuMutex()).
void resetBackend();
/// Six counts of [start, stop] including 2/3/5; false when no GPU is usable.
bool countAllOnGpu(uint64_t start, uint64_t stop, uint64_t counts[6]);
/// countAllOnGpu with no device allocation (the backend must exist);
/// false when there is no backend yet.
bool countAllNoAllocOnGpu(uint64_t start, uint64_t stop, uint64_t counts[6]);
/// Batch counts (mode 0 persistent, 1 plain, 2 persistent through a captured
/// launch series); false when no GPU is usable.
bool countBatchOnGpu(const GpuQuery* q, size_t n, uint64_t* out, int mode);

} // namespace gpu
} // namespace primesieve

#endif
