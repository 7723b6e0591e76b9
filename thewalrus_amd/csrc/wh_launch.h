// wh_launch.h -- host-side launchers exported by the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <mutex>

// A kernel's dynamic-LDS limit, raised only when a launch needs more than the kernel was last
// given on the current device: hipFuncSetAttribute costs about as much as the launch itself, and
// the limit never has to shrink.  One instance per kernel (a function-local static).
struct wh_lds_limit {
    std::atomic<size_t> bytes[64];  // per device; zero in static storage
    std::mutex mu;
};
inline hipError_t wh_raise_lds_limit(const void* fn, size_t lds, wh_lds_limit& lim) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= 64) return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (lim.bytes[dev].load(std::memory_order_acquire) >= lds) return hipSuccess;
    std::lock_guard<std::mutex> g(lim.mu);
    if (lim.bytes[dev].load(std::memory_order_relaxed) >= lds) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e == hipSuccess) lim.bytes[dev].store(lds, std::memory_order_release);
    return e;
}

struct wh_job_dev;
struct wh_lhb_job;

size_t wh_generic_lds_bytes(int MM, int NQ);
hipError_t wh_launch_generic(const wh_job_dev* jobs, int njobs, const void* pool, void* partials,
                             int64_t nblocks, int64_t block_offset, int MM, int NQ, hipStream_t stream);
size_t wh_pm1_lds_bytes(int T, bool real, int nw);
hipError_t wh_launch_pm1(int T, bool half, bool real, bool loop, const wh_job_dev* jobs, int njobs,
                         const void* pool, void* partials, int64_t nblocks, int64_t block_offset,
                         int nw, hipStream_t stream);
size_t wh_pm1m_lds_bytes(int T, int nw);
hipError_t wh_launch_pm1m(int T, bool loop, bool mtl, const wh_job_dev* jobs, int njobs, const void* pool,
                          void* partials, int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream);
size_t wh_pm1s_lds_bytes(int T, int nw, bool real, bool loop);
// Workgroups of nw waves that one CU holds for that instance (registers, LDS, wave slots), from
// the runtime's occupancy calculator on the current device; 0 when it cannot say.
int wh_pm1s_resident(int T, int nw, bool real, bool loop);
struct wh_mixed_group;
hipError_t wh_launch_pm1s_persistent(const wh_mixed_group* groups, int ngroups, const wh_job_dev* jobs, int njobs,
                                     const void* pool, void* partials, unsigned* counters, void* out, int nwg,
                                     hipStream_t stream);
hipError_t wh_launch_pm1s(int T, bool loop, bool real, const wh_job_dev* jobs, int njobs, const void* pool,
                          void* partials, int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream);
size_t wh_rep_lds_bytes(int T, bool real, int nw, int H);
hipError_t wh_launch_rep(int T, int H, bool real, bool loop, const wh_job_dev* jobs, int njobs, const void* pool,
                         void* partials, int64_t nblocks, int64_t block_offset, int nw, hipStream_t stream);
hipError_t wh_launch_fold(const wh_job_dev* jobs, int njobs, const void* partials, void* out,
                          hipStream_t stream);
size_t wh_lhb_lds_bytes(int MM);
hipError_t wh_launch_lhb(const wh_lhb_job* jobs, int njobs, const void* pool, void* partials,
                         int64_t nblocks, int MM, void* out, hipStream_t stream);
hipError_t wh_launch_hermite(int n, const int* cut, int renorm, const void* R, const void* y, const double* sq,
                             void* G, hipStream_t stream);
