// adgpu_launch.h -- host-side launch helpers shared by the kernel translation
// units (grid sizing for persistent kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <map>
#include <utility>

namespace adgpu {

// Workgroups of `threads` threads that fit on the device at once for kernel
// kfn with lds bytes of dynamic LDS (occupancy x compute units), cached per
// (kernel, lds, threads).  ADGPU_LS_GROUPS_PER_CU caps the workgroups per
// compute unit of the persistent local-search launches (a launch-shape
// parameter for tuning; the default is the occupancy limit).
inline int resident_groups(const void* kfn, size_t lds, int threads = 256) {
    static std::mutex mu;
    static std::map<std::pair<const void*, size_t>, int> cache;
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_pair(kfn, lds * 1024 + (size_t)threads);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, threads, lds) != hipSuccess) return 0;
    if (const char* cap = std::getenv("ADGPU_LS_GROUPS_PER_CU")) {
        const int c = std::atoi(cap);
        if (c > 0 && c < per_cu) per_cu = c;
    }
    int n = per_cu * cus;
    cache[key] = n;
    return n;
}

// Workgroup size of the persistent local-search launches: smaller groups
// release their compute-unit slots sooner when a launch drains, so another
// run group's launch can move in.  ADGPU_LS_BLOCK = 64, 128 or 256.
inline int ls_group_threads() {
    static const int t = [] {
        const char* e = std::getenv("ADGPU_LS_BLOCK");
        const int v = e ? std::atoi(e) : 256;
        return (v == 64 || v == 128) ? v : 256;
    }();
    return t;
}

// Resident workgroups a persistent launch may take when `share` launches run
// side by side: 1/share of the device.  ADGPU_LS_SHARE=k divides by k
// instead (0: every launch may take the whole device).
inline int persistent_cap(const void* kfn, size_t lds, int threads, int share) {
    static const int fixed = [] {
        const char* e = std::getenv("ADGPU_LS_SHARE");
        return e ? std::atoi(e) : -1;
    }();
    const int n = resident_groups(kfn, lds, threads);
    const int d = fixed >= 0 ? fixed : share;
    return d > 1 ? n / d : n;
}

}  // namespace adgpu
