// adgpu_launch.h -- host-side launch helpers shared by the kernel translation
// units (grid sizing for persistent kernels).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>
#include <map>
#include <utility>

namespace adgpu {

// Workgroups of 256 threads that fit on the device at once for kernel kfn
// with lds bytes of dynamic LDS (occupancy x compute units), cached per
// (kernel, lds) pair.  ADGPU_LS_GROUPS_PER_CU caps the workgroups per compute
// unit of the persistent local-search launches (a launch-shape parameter for
// tuning; the default is the occupancy limit).
inline int resident_groups(const void* kfn, size_t lds) {
    static std::mutex mu;
    static std::map<std::pair<const void*, size_t>, int> cache;
    std::lock_guard<std::mutex> g(mu);
    auto key = std::make_pair(kfn, lds);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kfn, 256, lds) != hipSuccess) return 0;
    if (const char* cap = std::getenv("ADGPU_LS_GROUPS_PER_CU")) {
        const int c = std::atoi(cap);
        if (c > 0 && c < per_cu) per_cu = c;
    }
    int n = per_cu * cus;
    cache[key] = n;
    return n;
}

}  // namespace adgpu
