// GpuData.h -- device state of the MI355X backend: the HIP counterpart of the
// reference's cuda/GpuData.h (GpuData / GpuTempData, cuda/GpuData.h:74-149).
// The reference's main.cpp touches only devnum, preallocated_gridsize,
// pMem_fgrids, device_name and device_busy (host/src/main.cpp:219-232,
// 316-325, 407, 471); the rest is this backend's own.
#ifndef GPUDATADOTH
#define GPUDATADOTH

#include "adgpu.h"

struct GpuData {
    int       devnum;                  // requested device (-1: first); -2 once set up
    int       preallocated_gridsize;   // unused: map sets stay resident in the context
    int       devid;                   // device in use
    adgpu_ctx ctx;                     // library context (stream, resident maps)
    int       map_id;                  // resident map set of the last receptor, 0 = none
    char      map_key[1024];           // which receptor grid map_id holds
};

struct GpuTempData {
    float* pMem_fgrids;                // unused (kept for main.cpp)
    char*  device_name;
    bool   device_busy;
};

#endif
