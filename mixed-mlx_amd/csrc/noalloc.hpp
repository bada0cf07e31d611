// Internal interface of the allocation-free Jacobi entry point
// (msq_jacobi_batched_noalloc): the per-device static scratch and the kernels
// that plan its launches.  Not exported; the C ABI is include/msq_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/msq_hip.h"
#include "kernels.hpp"

namespace msq {

constexpr uint32_t kNoallocMaxMatrices = MSQ_NOALLOC_MAX_MATRICES;
// Launch classes: padded orders 16, 32, ..., 256 (orders 1..256).
constexpr int kNoallocClasses = 16;

__host__ __device__ constexpr int noalloc_class(uint32_t order) {
  return static_cast<int>((order + 15u) / 16u) - 1;
}

// Position of each class's first job in the plan (by class index).
struct NoallocClasses {
  uint32_t start[kNoallocClasses];
};

// Device addresses of the scratch regions on one GPU (null when the device
// was never registered).
struct NoallocScratch {
  uint32_t* orders;
  msq_jacobi_options* options;
  JacobiJob* jobs;
  JacobiStatus* status;
  msq_jacobi_report* reports;
};

// Once per device, with that device current: looks up the scratch array's
// address on the device (hipGetSymbolAddress) and keeps it.
hipError_t noalloc_register(int device);
NoallocScratch noalloc_scratch(int device);

// Jobs of `count` matrices from the orders and options already copied into
// the scratch.
hipError_t launch_noalloc_plan(const NoallocScratch& s, uint32_t options_count, uint32_t count,
                               const NoallocClasses& classes, hipStream_t stream);
// status -> reports (C ABI records) for matrices 0 .. count - 1.
hipError_t launch_noalloc_reports(const NoallocScratch& s, uint32_t count, hipStream_t stream);

}  // namespace msq
