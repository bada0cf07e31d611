// Device scratch of the allocation-free entry point (msq_jacobi_batched_noalloc)
// and the kernels that plan its launches on the device.
//
// The scratch is one file-scope static device array.  Every GPU the library
// runs on holds its own copy of it, at its own address, until the process
// ends.  Host code never passes the array's name to a kernel or a copy (in
// host code that name is only a stand-in, not a device address): the address
// on each device comes from hipGetSymbolAddress, called once per device with
// that device current (msq_runtime_create calls noalloc_register), and is
// kept in a per-device table.
//
// Layout (256-byte aligned regions, kNoallocMaxMatrices entries each):
//   orders    uint32            the caller's orders, copied in
//   options   msq_jacobi_options (1 or count entries)
//   jobs      JacobiJob          launch plan, class-major (largest padded order first)
//   status    JacobiStatus       written by the Jacobi kernel, by matrix index
//   reports   msq_jacobi_report  status converted to the C ABI's record, by matrix index
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/msq_hip.h"
#include "kernels.hpp"
#include "noalloc.hpp"

namespace msq {
namespace {

constexpr size_t align256(size_t b) { return (b + 255) / 256 * 256; }

constexpr size_t kOrdersOffset = 0;
constexpr size_t kOptionsOffset = kOrdersOffset + align256(kNoallocMaxMatrices * sizeof(uint32_t));
constexpr size_t kJobsOffset =
    kOptionsOffset + align256(kNoallocMaxMatrices * sizeof(msq_jacobi_options));
constexpr size_t kStatusOffset = kJobsOffset + align256(kNoallocMaxMatrices * sizeof(JacobiJob));
constexpr size_t kReportsOffset =
    kStatusOffset + align256(kNoallocMaxMatrices * sizeof(JacobiStatus));
constexpr size_t kScratchBytes =
    kReportsOffset + align256(kNoallocMaxMatrices * sizeof(msq_jacobi_report));
static_assert(kScratchBytes == MSQ_NOALLOC_SCRATCH_BYTES, "scratch size of the C ABI");

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
char* g_scratch[kMaxDevices] = {};

}  // namespace
}  // namespace msq

// The scratch array itself (see the file comment: host code uses the
// per-device addresses from noalloc_register, never this name).
static __device__ __align__(256) unsigned char scratch[msq::kScratchBytes];

namespace msq {
namespace {

constexpr int kPlanThreads = 1024;
constexpr int kPerThread = kNoallocMaxMatrices / kPlanThreads;
static_assert(kPerThread * kPlanThreads == kNoallocMaxMatrices, "whole chunks per thread");

// One workgroup plans the whole call.  Thread t takes matrices
// [t kPerThread, (t + 1) kPerThread): their batch-layout offsets are the
// exclusive prefix sums of ld^2 and ld, and each matrix's job goes to its
// class's first position plus the number of earlier matrices of the class
// (stable, in matrix order, as the allocating entry point plans them).  The
// per-thread sums and class counts are scanned across threads in LDS
// (Hillis-Steele, inclusive).
__global__ __launch_bounds__(kPlanThreads) void noalloc_plan_kernel(
    const uint32_t* __restrict__ orders, const msq_jacobi_options* __restrict__ options,
    uint32_t options_count, uint32_t count, NoallocClasses classes, JacobiJob* __restrict__ jobs) {
  __shared__ uint64_t s_off[kPlanThreads];
  __shared__ uint32_t s_voff[kPlanThreads];
  __shared__ uint32_t s_cls[kNoallocClasses][kPlanThreads];
  const int t = threadIdx.x;
  const uint32_t first = static_cast<uint32_t>(t) * kPerThread;
  for (int c = 0; c < kNoallocClasses; ++c) s_cls[c][t] = 0;
  uint64_t off = 0;
  uint32_t voff = 0;
  for (int k = 0; k < kPerThread; ++k) {
    const uint32_t i = first + k;
    if (i >= count) break;
    const uint32_t o = orders[i];
    const uint32_t ld = (o + 31u) & ~31u;
    off += static_cast<uint64_t>(ld) * ld;
    voff += ld;
    s_cls[noalloc_class(o)][t] += 1;  // this thread's own column
  }
  s_off[t] = off;
  s_voff[t] = voff;
  __syncthreads();
  for (int d = 1; d < kPlanThreads; d <<= 1) {
    const uint64_t a = t >= d ? s_off[t - d] : 0;
    const uint32_t v = t >= d ? s_voff[t - d] : 0;
    uint32_t cc[kNoallocClasses];
#pragma unroll
    for (int c = 0; c < kNoallocClasses; ++c) cc[c] = t >= d ? s_cls[c][t - d] : 0;
    __syncthreads();
    s_off[t] += a;
    s_voff[t] += v;
#pragma unroll
    for (int c = 0; c < kNoallocClasses; ++c) s_cls[c][t] += cc[c];
    __syncthreads();
  }
  // Exclusive prefixes of this thread's chunk: the inclusive sums of the
  // thread before it.
  uint64_t a_off = t > 0 ? s_off[t - 1] : 0;
  uint32_t s_off_v = t > 0 ? s_voff[t - 1] : 0;
  for (int k = 0; k < kPerThread; ++k) {
    const uint32_t i = first + k;
    if (i >= count) break;
    const uint32_t o = orders[i];
    const uint32_t ld = (o + 31u) & ~31u;
    const int c = noalloc_class(o);
    // Rank of matrix i in its class: earlier threads' count, then this
    // thread's earlier matrices of the class (counted again here).
    uint32_t rank = t > 0 ? s_cls[c][t - 1] : 0;
    for (int q = 0; q < k; ++q)
      if (noalloc_class(orders[first + q]) == c) ++rank;
    const msq_jacobi_options& op = options[options_count == 1 ? 0 : i];
    JacobiJob j{};
    j.a_offset = a_off;
    j.v_offset = a_off;
    j.s_offset = s_off_v;
    j.order = o;
    j.ld = ld;
    j.max_sweeps = op.max_sweeps ? op.max_sweeps : 30u;
    j.fixed_sweeps = op.fixed_sweeps;
    j.off_tol = op.off_tol;
    j.off_tol_lambda = op.off_tol_lambda;
    j.stagnation = op.stagnation > 0.0f ? op.stagnation : 0.5f;
    j.floor_tol = op.floor_tol > 0.0f ? op.floor_tol : 1e-5f;
    j.correction_rate = op.correction_rate;
    j.status_index = i;
    jobs[classes.start[c] + rank] = j;
    a_off += static_cast<uint64_t>(ld) * ld;
    s_off_v += ld;
  }
}

// The Jacobi kernel's per-matrix status as the C ABI's report record
// (fill_jacobi_report's conversion, on the device).
__global__ void noalloc_reports_kernel(const JacobiStatus* __restrict__ status,
                                       msq_jacobi_report* __restrict__ reports, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const JacobiStatus& s = status[i];
  msq_jacobi_report r;
  r.sweeps = s.sweeps;
  r.invalid_spectrum = (s.flags & kJacobiInvalidSpectrum) ? 1u : 0u;
  r.off_norm = s.off_norm;
  r.matrix_norm = s.matrix_norm;
  r.min_diagonal = s.min_diagonal;
  r.correction_rate = s.correction_rate;
  for (uint32_t k = 0; k < MSQ_JACOBI_HISTORY; ++k)
    r.off_history[k] = k < s.sweeps && k < kJacobiHistory ? s.off_history[k] : 0.0f;
  reports[i] = r;
}

}  // namespace

hipError_t noalloc_register(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = static_cast<char*>(p);
  return hipSuccess;
}

NoallocScratch noalloc_scratch(int device) {
  NoallocScratch s{};
  if (device < 0 || device >= kMaxDevices) return s;
  char* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    base = g_scratch[device];
  }
  if (!base) return s;
  s.orders = reinterpret_cast<uint32_t*>(base + kOrdersOffset);
  s.options = reinterpret_cast<msq_jacobi_options*>(base + kOptionsOffset);
  s.jobs = reinterpret_cast<JacobiJob*>(base + kJobsOffset);
  s.status = reinterpret_cast<JacobiStatus*>(base + kStatusOffset);
  s.reports = reinterpret_cast<msq_jacobi_report*>(base + kReportsOffset);
  return s;
}

hipError_t launch_noalloc_plan(const NoallocScratch& s, uint32_t options_count, uint32_t count,
                               const NoallocClasses& classes, hipStream_t stream) {
  if (count == 0 || count > kNoallocMaxMatrices) return hipErrorInvalidValue;
  noalloc_plan_kernel<<<1, kPlanThreads, 0, stream>>>(s.orders, s.options, options_count, count,
                                                      classes, s.jobs);
  return hipGetLastError();
}

hipError_t launch_noalloc_reports(const NoallocScratch& s, uint32_t count, hipStream_t stream) {
  if (count == 0 || count > kNoallocMaxMatrices) return hipErrorInvalidValue;
  noalloc_reports_kernel<<<(count + 255) / 256, 256, 0, stream>>>(s.status, s.reports, count);
  return hipGetLastError();
}

}  // namespace msq
