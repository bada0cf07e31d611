// C ABI of libmsq_hip.so (include/msq_hip.h): runtime, codecs, Jacobi,
// products and the batched GPU-resident solve.
//
// The solve follows solveGpuResidentSpd (src/metal_solver.cpp:92-336):
// Jacobi -> frozen factor -> reconstruction -> residual -> corrections, with
// three differences of design: the whole batch shares each launch, the
// residual is binary64 (the reference's float-float, :1268-1317 of the .metal),
// and the per-matrix stopping decision is taken on the device (the reference
// reads the norm back after every correction, src/metal_runtime.cpp:950-965),
// so the launch sequence is fixed and no host round trip sits between
// corrections.
#include "../../include/msq_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <array>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "noalloc.hpp"

namespace {

thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

// On failure: record the message and clear HIP's per-thread last error, so a
// later, unrelated HIP call in the caller does not report this one.
#define HIP_TRY(expr)                                                                     \
  do {                                                                                    \
    hipError_t e_ = (expr);                                                               \
    if (e_ != hipSuccess) {                                                               \
      (void)hipGetLastError();                                                            \
      return set_error(MSQ_ERROR_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                       __FILE__, __LINE__);                                               \
    }                                                                                     \
  } while (0)

#define MSQ_TRY(expr)          \
  do {                         \
    int s_ = (expr);           \
    if (s_ != MSQ_OK) return s_; \
  } while (0)

uint32_t round_up(uint32_t v, uint32_t m) { return (v + m - 1) / m * m; }

// Device scratch owned by the runtime; grows, never shrinks, reused by calls.
struct Arena {
  char* base = nullptr;
  size_t capacity = 0;
};

}  // namespace

// Streams on which independent launches of one call overlap (the order
// classes of a batched Jacobi); forked from and joined back into `stream`.
constexpr int kLanes = 4;
// Captured solve graphs kept per runtime (one per batch signature).
constexpr size_t kMaxGraphs = 4;
// GEMM tile table: 5 operand types x ld classes 32, 64, ..., 256, and > 256.
constexpr int kGemmTypes = 5;
constexpr int kTileClasses = 9;
constexpr int kTileEdges[3] = {32, 64, 128};
static_assert(kGemmTypes * kTileClasses == MSQ_TILE_TABLE_BYTES, "tile table size");
struct GraphEntry;

struct msq_runtime {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  hipStream_t lanes[kLanes] = {};
  hipEvent_t fork = nullptr;
  hipEvent_t joins[kLanes] = {};
  hipEvent_t graph_in = nullptr, graph_out = nullptr;
  Arena arena;
  Arena large;  // blocked Jacobi workspace (orders above 256)
  Arena refine;  // second Jacobi pass of unresolved matrices (eager solves)
  bool graphs_enabled = false;
  // Autotuned GEMM tile edge per (GemmType, ld class); 0: the default.
  uint8_t tiles[kGemmTypes][kTileClasses] = {};
  std::vector<std::unique_ptr<GraphEntry>> graphs;
  float* last_vectors = nullptr;
  float* last_roots = nullptr;
  // RoundingMode and seed of the reference's runtime (msq_runtime_set_rounding).
  int rounding_mode = MSQ_ROUNDING_NEAREST;
  uint32_t rounding_seed = 0;
  // Job counter of the persistent Jacobi launch (device memory, allocated
  // with the runtime) and the device's CU count (its grid).
  uint32_t* jacobi_counter = nullptr;
  int cu_count = 0;
  // Eigensolver of the persistent family (msq_runtime_set_jacobi_method).
  int jacobi_method = MSQ_JACOBI_BLOCKED_MFMA;
};

namespace {

int activate(msq_runtime* rt) {
  if (!rt) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "runtime is null");
  HIP_TRY(hipSetDevice(rt->device));
  return MSQ_OK;
}

// Reserves `bytes` of runtime scratch for one call (256-byte aligned).  All
// reservations of a call are made before any kernel is enqueued, and the
// arena is only regrown between calls, after a synchronize.
struct Scratch {
  msq_runtime* rt;
  Arena* target = nullptr;  // null: the runtime's arena
  std::vector<std::pair<void**, size_t>> wants;
  template <typename T>
  void want(T*& ptr, size_t count) {
    wants.push_back({reinterpret_cast<void**>(&ptr), count * sizeof(T)});
  }
  int commit() {
    size_t total = 0;
    for (auto& w : wants) total += (w.second + 255) / 256 * 256;
    Arena& a = target ? *target : rt->arena;
    if (total > a.capacity) {
      const size_t grow = std::max(total, a.capacity * 3 / 2);
      HIP_TRY(hipStreamSynchronize(rt->stream));
      // msq_last_factors' pointers into this arena die with it.
      if (rt->last_vectors && rt->last_vectors >= reinterpret_cast<float*>(a.base) &&
          rt->last_vectors < reinterpret_cast<float*>(a.base + a.capacity)) {
        rt->last_vectors = nullptr;
        rt->last_roots = nullptr;
      }
      if (a.base) HIP_TRY(hipFree(a.base));
      a.base = nullptr;
      a.capacity = 0;
      hipError_t e = hipMalloc(&a.base, grow);
      if (e != hipSuccess)
        return set_error(MSQ_ERROR_OUT_OF_MEMORY, "device scratch of %zu bytes: %s", grow,
                         hipGetErrorString(e));
      a.capacity = grow;
    }
    size_t off = 0;
    for (auto& w : wants) {
      *w.first = a.base + off;
      off += (w.second + 255) / 256 * 256;
    }
    return MSQ_OK;
  }
};

struct Layout {
  std::vector<uint32_t> order, ld;
  std::vector<uint64_t> off, voff;
  uint64_t total = 0, vtotal = 0;
  uint32_t max_ld = 0;
};

int make_layout(uint32_t count, const uint32_t* orders, Layout& L) {
  if (count == 0) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "empty batch");
  if (!orders) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "orders is null");
  L.order.assign(orders, orders + count);
  L.ld.resize(count);
  L.off.resize(count);
  L.voff.resize(count);
  for (uint32_t i = 0; i < count; ++i) {
    if (orders[i] < 1 || orders[i] > 65504)
      return set_error(MSQ_ERROR_INVALID_ARGUMENT, "order %u of matrix %u is outside [1, 65504]",
                       orders[i], i);
    L.ld[i] = round_up(orders[i], 32);
    L.off[i] = L.total;
    L.voff[i] = L.vtotal;
    L.total += static_cast<uint64_t>(L.ld[i]) * L.ld[i];
    L.vtotal += L.ld[i];
    L.max_ld = std::max(L.max_ld, L.ld[i]);
  }
  return MSQ_OK;
}

msq::Format format_of(int precision, bool* ok) {
  *ok = true;
  switch (precision) {
    case MSQ_BF16: return msq::Format::Bf16;
    case MSQ_FP8_E4M3FN: return msq::Format::Fp8E4M3Fn;
    case MSQ_FP8_E5M2: return msq::Format::Fp8E5M2;
    case MSQ_MXFP8_E4M3FN: return msq::Format::Mxfp8E4M3Fn;
    case MSQ_MXFP8_E5M2: return msq::Format::Mxfp8E5M2;
    case MSQ_FP4_E2M1: return msq::Format::Fp4E2M1;
    case MSQ_MXFP4_E2M1: return msq::Format::Mxfp4E2M1;
    default: *ok = false; return msq::Format::Fp32;
  }
}

bool is_mx(int precision) {
  return precision == MSQ_MXFP8_E4M3FN || precision == MSQ_MXFP8_E5M2 ||
         precision == MSQ_MXFP4_E2M1;
}

bool is_log_spectrum(int precision) {
  return precision == MSQ_FP8_E4M3FN_LOG_S || precision == MSQ_FP4_E2M1_LOG_S;
}

// Products supported by the MFMA GEMMs.
bool product_supported(int p) {
  return p == MSQ_FP32 || p == MSQ_BF16 || p == MSQ_FP8_E4M3FN_SCALED || p == MSQ_MXFP4_E2M1;
}

msq::GemmType gemm_type(int p) {
  switch (p) {
    case MSQ_BF16: return msq::GemmType::Bf16;
    case MSQ_FP8_E4M3FN_SCALED: return msq::GemmType::Fp8;
    case MSQ_MXFP4_E2M1: return msq::GemmType::Mxfp4;
    default: return msq::GemmType::F32;
  }
}

// Bytes of the encoded operand for `elements` values in precision p.
size_t operand_bytes(int p, uint64_t elements) {
  switch (p) {
    case MSQ_BF16: return elements * 2;
    case MSQ_FP8_E4M3FN_SCALED: return elements;
    case MSQ_MXFP4_E2M1: return (elements + 1) / 2;
    default: return elements * 4;
  }
}

template <typename T>
int upload(msq_runtime* rt, T* dst, const std::vector<T>& src) {
  if (src.empty()) return MSQ_OK;
  HIP_TRY(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice,
                         rt->stream));
  return MSQ_OK;
}

// -------------------------------------------------------------- Jacobi plan
msq::JacobiJob make_job(const Layout& L, uint32_t i, const msq_jacobi_options& o) {
  msq::JacobiJob j{};
  j.a_offset = L.off[i];
  j.v_offset = L.off[i];
  j.s_offset = L.voff[i];
  j.order = L.order[i];
  j.ld = L.ld[i];
  j.max_sweeps = o.max_sweeps ? o.max_sweeps : 30;
  j.fixed_sweeps = o.fixed_sweeps;
  j.off_tol = o.off_tol;
  j.off_tol_lambda = o.off_tol_lambda;
  j.stagnation = o.stagnation > 0.0f ? o.stagnation : 0.5f;
  j.floor_tol = o.floor_tol > 0.0f ? o.floor_tol : 1e-5f;
  j.correction_rate = o.correction_rate;
  j.status_index = i;
  return j;
}

// Blocked Jacobi above the batched kernel's limit (orders > 256): block
// rounds of 64 x 64 subproblems (diagonalised by the batched kernel) whose
// rotations are applied to A and V as MFMA products.  Matrices of one padded
// order form a class and advance together; the FP64 stop-rule measures of the
// class are read back after each sweep and the members that stop leave the
// active list.
// Sweeps of each 64 x 64 subproblem per block round: one.  The outer sweeps
// converge at about the same count as with two inner sweeps (12.3 against
// 11.9 on orders 257..512), and each block round costs ~30% less.
constexpr uint32_t kBlockedInnerSweeps = 1;

// Device workspace of the large-order path, kept between calls (grown only
// after a synchronize).
int reserve_large(msq_runtime* rt, size_t bytes, char** out) {
  Arena& a = rt->large;
  if (bytes > a.capacity) {
    HIP_TRY(hipStreamSynchronize(rt->stream));
    for (int l = 0; l < kLanes; ++l) HIP_TRY(hipStreamSynchronize(rt->lanes[l]));
    if (a.base) HIP_TRY(hipFree(a.base));
    a.base = nullptr;
    a.capacity = 0;
    hipError_t e = hipMalloc(&a.base, bytes);
    if (e != hipSuccess)
      return set_error(MSQ_ERROR_OUT_OF_MEMORY, "blocked Jacobi workspace of %zu bytes: %s", bytes,
                       hipGetErrorString(e));
    a.capacity = bytes;
  }
  *out = a.base;
  return MSQ_OK;
}

struct LargeClass {
  uint32_t N = 0, np = 0;
  std::vector<uint32_t> members;  // matrix indices
  std::vector<msq::JacobiJob> opts;
  std::vector<msq::JacobiStatus> status;
  std::vector<double> previous;
  std::vector<uint32_t> active, next;  // member positions
  std::vector<double> h_off, h_g, h_norm;
  std::vector<float> h_min;
  std::vector<uint32_t> h_bad;
  hipStream_t stream = nullptr;
  // device
  msq::BlockedMember* dmem = nullptr;
  uint32_t* dactive = nullptr;
  float *A = nullptr, *Vw = nullptr, *U = nullptr, *pdiag = nullptr, *proots = nullptr;
  double *staging = nullptr, *off2 = nullptr, *g2 = nullptr, *norm2 = nullptr;
  float* dmin = nullptr;
  uint32_t* bad = nullptr;
  msq::JacobiStatus* pst = nullptr;
  msq::JacobiJob* pjobs = nullptr;  // subproblem jobs
};

int run_large_jacobi(msq_runtime* rt, const Layout& L, const double* a64, float* V, float* diag,
                     float* roots, const msq_jacobi_options* opts, uint32_t opts_count,
                     msq::JacobiStatus* jstatus) {
  std::map<uint32_t, std::vector<uint32_t>> by_padded;
  for (uint32_t i = 0; i < L.order.size(); ++i)
    if (L.order[i] > msq::kBatchedJacobiMaxOrder)
      by_padded[msq::blocked_padded_order(L.order[i])].push_back(i);
  if (by_padded.empty()) return MSQ_OK;
  std::vector<LargeClass> classes;
  for (auto it = by_padded.rbegin(); it != by_padded.rend(); ++it) {
    LargeClass c;
    c.N = it->first;
    c.np = c.N / msq::kBlockedWidth / 2;
    c.members = it->second;
    classes.push_back(std::move(c));
  }
  // Workspace: per class, the members' working copies, one round's
  // subproblems and the per-member measures.
  size_t bytes = 0;
  auto carve = [&](size_t b) {
    const size_t o = bytes;
    bytes += (b + 255) / 256 * 256;
    return o;
  };
  std::vector<std::array<size_t, 14>> offs(classes.size());
  for (size_t k = 0; k < classes.size(); ++k) {
    const LargeClass& c = classes[k];
    const size_t cnt = c.members.size(), sq = static_cast<size_t>(c.N) * c.N;
    const size_t subs = cnt * c.np;
    offs[k] = {carve(cnt * sizeof(msq::BlockedMember)), carve(cnt * sizeof(uint32_t)),
               carve(cnt * sq * sizeof(float)),         carve(cnt * sq * sizeof(float)),
               carve(subs * 4096 * sizeof(double)),     carve(subs * 4096 * sizeof(float)),
               carve(subs * 64 * sizeof(float)),        carve(subs * 64 * sizeof(float)),
               carve(subs * sizeof(msq::JacobiStatus)), carve(subs * sizeof(msq::JacobiJob)),
               carve(cnt * (3 * msq::kOffnormParts) * sizeof(double)),
               carve(cnt * msq::kOffnormParts * sizeof(float)),
               carve(cnt * sizeof(uint32_t)),           0};
  }
  const size_t o_status = carve(L.order.size() * sizeof(msq::JacobiStatus));
  const size_t o_index = carve(L.order.size() * sizeof(uint32_t));
  char* base = nullptr;
  MSQ_TRY(reserve_large(rt, bytes, &base));
  // The input and the caller's earlier work are complete before the lanes
  // start (the host synchronises every sweep anyway).
  HIP_TRY(hipStreamSynchronize(rt->stream));
  for (size_t k = 0; k < classes.size(); ++k) {
    LargeClass& c = classes[k];
    const auto& o = offs[k];
    c.stream = classes.size() > 1 ? rt->lanes[k % kLanes] : rt->stream;
    c.dmem = reinterpret_cast<msq::BlockedMember*>(base + o[0]);
    c.dactive = reinterpret_cast<uint32_t*>(base + o[1]);
    c.A = reinterpret_cast<float*>(base + o[2]);
    c.Vw = reinterpret_cast<float*>(base + o[3]);
    c.staging = reinterpret_cast<double*>(base + o[4]);
    c.U = reinterpret_cast<float*>(base + o[5]);
    c.pdiag = reinterpret_cast<float*>(base + o[6]);
    c.proots = reinterpret_cast<float*>(base + o[7]);
    c.pst = reinterpret_cast<msq::JacobiStatus*>(base + o[8]);
    c.pjobs = reinterpret_cast<msq::JacobiJob*>(base + o[9]);
    c.off2 = reinterpret_cast<double*>(base + o[10]);
    const size_t cnt = c.members.size();
    c.g2 = c.off2 + cnt * msq::kOffnormParts;
    c.norm2 = c.g2 + cnt * msq::kOffnormParts;
    c.dmin = reinterpret_cast<float*>(base + o[11]);
    c.bad = reinterpret_cast<uint32_t*>(base + o[12]);
    std::vector<msq::BlockedMember> hm(cnt);
    for (size_t j = 0; j < cnt; ++j) {
      const uint32_t i = c.members[j];
      hm[j] = msq::BlockedMember{L.off[i], L.voff[i], L.order[i], L.ld[i]};
      c.opts.push_back(make_job(L, i, opts[opts_count == 1 ? 0 : i]));
    }
    std::vector<msq::JacobiJob> jobs(cnt * c.np);
    for (size_t b = 0; b < jobs.size(); ++b) {
      msq::JacobiJob& jb = jobs[b];
      jb = msq::JacobiJob{};
      jb.a_offset = jb.v_offset = static_cast<uint64_t>(b) * 64 * 64;
      jb.s_offset = static_cast<uint64_t>(b) * 64;
      jb.order = 64;
      jb.ld = 64;
      jb.max_sweeps = kBlockedInnerSweeps;
      jb.fixed_sweeps = kBlockedInnerSweeps;
      jb.off_tol = 1e-7f;
      jb.stagnation = 0.5f;
      jb.floor_tol = 1e-5f;
      jb.status_index = static_cast<uint32_t>(b);
    }
    c.status.assign(cnt, msq::JacobiStatus{});
    c.previous.assign(cnt, INFINITY);
    c.h_off.resize(cnt * msq::kOffnormParts);
    c.h_g.resize(cnt * msq::kOffnormParts);
    c.h_norm.resize(cnt * msq::kOffnormParts);
    c.h_min.resize(cnt * msq::kOffnormParts);
    c.h_bad.resize(cnt);
    c.active.resize(cnt);
    for (size_t j = 0; j < cnt; ++j) c.active[j] = static_cast<uint32_t>(j);
    HIP_TRY(hipMemcpyAsync(c.dmem, hm.data(), cnt * sizeof(msq::BlockedMember),
                           hipMemcpyHostToDevice, c.stream));
    HIP_TRY(hipMemcpyAsync(c.pjobs, jobs.data(), jobs.size() * sizeof(msq::JacobiJob),
                           hipMemcpyHostToDevice, c.stream));
    HIP_TRY(msq::launch_blocked_init(c.dmem, static_cast<uint32_t>(cnt), a64, c.A, c.Vw, c.N,
                                     c.norm2, c.stream));
    HIP_TRY(hipMemcpyAsync(c.h_norm.data(), c.norm2, cnt * msq::kOffnormParts * sizeof(double),
                           hipMemcpyDeviceToHost, c.stream));
    // hm and jobs are host temporaries: their copies finish before they go.
    HIP_TRY(hipStreamSynchronize(c.stream));
    const uint32_t nparts = msq::blocked_offnorm_parts(static_cast<uint32_t>(cnt));
    for (size_t j = 0; j < cnt; ++j) {
      double s2 = 0.0;
      for (uint32_t k = 0; k < nparts; ++k) s2 += c.h_norm[j * msq::kOffnormParts + k];
      c.status[j].matrix_norm = std::sqrt(s2);
    }
  }
  for (uint32_t sweep = 0;; ++sweep) {
    bool any = false;
    for (LargeClass& c : classes) {
      if (c.active.empty()) continue;
      any = true;
      const uint32_t na = static_cast<uint32_t>(c.active.size());
      HIP_TRY(hipMemcpyAsync(c.dactive, c.active.data(), na * sizeof(uint32_t),
                             hipMemcpyHostToDevice, c.stream));
      const uint32_t nb = c.N / msq::kBlockedWidth;
      for (uint32_t t = 0; t + 1 < nb; ++t)
        HIP_TRY(msq::launch_blocked_round(c.dactive, na, c.A, c.Vw, c.N, t, c.staging, c.U,
                                          c.pjobs, c.pdiag, c.proots, c.pst, c.stream));
      HIP_TRY(msq::launch_blocked_offnorm(c.dactive, na, c.dmem, c.A, c.N, c.off2, c.g2, c.dmin,
                                          c.stream));
      const size_t slots = c.members.size() * msq::kOffnormParts;
      HIP_TRY(hipMemcpyAsync(c.h_off.data(), c.off2, slots * sizeof(double),
                             hipMemcpyDeviceToHost, c.stream));
      HIP_TRY(hipMemcpyAsync(c.h_g.data(), c.g2, slots * sizeof(double), hipMemcpyDeviceToHost,
                             c.stream));
      HIP_TRY(hipMemcpyAsync(c.h_min.data(), c.dmin, slots * sizeof(float),
                             hipMemcpyDeviceToHost, c.stream));
    }
    if (!any) break;
    for (LargeClass& c : classes) {
      if (c.active.empty()) continue;
      HIP_TRY(hipStreamSynchronize(c.stream));
      c.next.clear();
      const uint32_t parts = msq::blocked_offnorm_parts(static_cast<uint32_t>(c.active.size()));
      for (uint32_t j : c.active) {
        const msq::JacobiJob& job = c.opts[j];
        msq::JacobiStatus& st = c.status[j];
        double s2 = 0.0, g2 = 0.0;
        float d = INFINITY;
        for (uint32_t k = 0; k < parts; ++k) {
          const size_t slot = static_cast<size_t>(j) * msq::kOffnormParts + k;
          s2 += c.h_off[slot];
          g2 += c.h_g[slot];
          d = std::min(d, c.h_min[slot]);
        }
        const double offn = std::sqrt(s2);
        const double rate =
            d > 0.0f ? std::sqrt(g2) / std::sqrt(static_cast<double>(d)) : INFINITY;
        st.correction_rate = static_cast<float>(rate);
        st.sweeps = sweep + 1;
        st.off_norm = offn;
        st.min_diagonal = d;
        if (sweep < msq::kJacobiHistory)
          st.off_history[sweep] = static_cast<float>(offn / st.matrix_norm);
        const uint32_t cap = job.fixed_sweeps ? job.fixed_sweeps : job.max_sweeps;
        bool stop;
        if (job.fixed_sweeps) {
          stop = st.sweeps >= job.fixed_sweeps;
        } else {
          double tol = job.off_tol * st.matrix_norm;
          if (job.off_tol_lambda > 0.0f && d > 0.0f)
            tol = std::max(tol, double(job.off_tol_lambda) * d);
          stop = offn <= tol || st.sweeps >= cap ||
                 (job.correction_rate > 0.0f && rate <= job.correction_rate) ||
                 (offn > job.stagnation * c.previous[j] &&
                  offn <= job.floor_tol * st.matrix_norm);
        }
        c.previous[j] = offn;
        if (!stop) c.next.push_back(j);
      }
      std::swap(c.active, c.next);
    }
  }
  // Outputs and statuses.
  msq::JacobiStatus* dst = reinterpret_cast<msq::JacobiStatus*>(base + o_status);
  uint32_t* dindex = reinterpret_cast<uint32_t*>(base + o_index);
  std::vector<msq::JacobiStatus> all;
  std::vector<uint32_t> index;
  for (LargeClass& c : classes) {
    const uint32_t cnt = static_cast<uint32_t>(c.members.size());
    HIP_TRY(msq::launch_fill_u32(c.bad, cnt, 0u, c.stream));
    HIP_TRY(msq::launch_blocked_finish(c.dmem, cnt, c.A, c.Vw, c.N, V, diag, roots, c.bad,
                                       c.stream));
    HIP_TRY(hipMemcpyAsync(c.h_bad.data(), c.bad, cnt * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c.stream));
  }
  for (LargeClass& c : classes) {
    HIP_TRY(hipStreamSynchronize(c.stream));
    for (size_t j = 0; j < c.members.size(); ++j) {
      if (c.h_bad[j]) c.status[j].flags |= msq::kJacobiInvalidSpectrum;
      all.push_back(c.status[j]);
      index.push_back(c.members[j]);
    }
  }
  const uint32_t total = static_cast<uint32_t>(all.size());
  HIP_TRY(hipMemcpyAsync(dst, all.data(), total * sizeof(msq::JacobiStatus), hipMemcpyHostToDevice,
                         rt->stream));
  HIP_TRY(hipMemcpyAsync(dindex, index.data(), total * sizeof(uint32_t), hipMemcpyHostToDevice,
                         rt->stream));
  HIP_TRY(msq::launch_scatter_status(dst, dindex, total, jstatus, rt->stream));
  // all / index are host temporaries.
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

// Jacobi launch plan of a batch: jobs of the orders up to 256 grouped by
// padded order (largest first), as (first job, padded order) launches.
struct JacobiPlan {
  std::vector<msq::JacobiJob> jobs;
  std::vector<std::pair<size_t, uint32_t>> launches;  // ends with (jobs.size(), 0)
  std::vector<double> cost;                           // per launch: sum of n^3
  bool has_large = false;                             // orders above 256
};

void plan_jacobi(const Layout& L, const msq_jacobi_options* opts, uint32_t opts_count,
                 JacobiPlan& P) {
  const uint32_t count = static_cast<uint32_t>(L.order.size());
  P = JacobiPlan{};
  std::map<uint32_t, std::vector<uint32_t>> by_padded;
  for (uint32_t i = 0; i < count; ++i) {
    if (L.order[i] <= msq::kBatchedJacobiMaxOrder)
      by_padded[msq::batched_jacobi_padded_order(L.order[i])].push_back(i);
    else
      P.has_large = true;
  }
  // Largest orders first: their workgroups are the longest-running.
  for (auto it = by_padded.rbegin(); it != by_padded.rend(); ++it) {
    P.launches.push_back({P.jobs.size(), it->first});
    double cost = 0.0;
    for (uint32_t i : it->second) {
      P.jobs.push_back(make_job(L, i, opts[opts_count == 1 ? 0 : i]));
      cost += static_cast<double>(L.order[i]) * L.order[i] * L.order[i];
    }
    P.cost.push_back(cost);
  }
  P.launches.push_back({P.jobs.size(), 0});
}

// Launch classes of a batched Jacobi (padded orders 16..256: at most 16).
constexpr int kMaxJacobiClasses = 16;
struct JacobiClass {
  size_t first = 0;      // first job of the class
  uint32_t count = 0;    // jobs
  uint32_t order = 0;    // padded order
  double cost = 0.0;     // sum of n^3
};

// The batched launches of `classes` (jobs already on the device): device work
// only, no host synchronisation and no allocation (capturable in a graph).
// Orders 129..256 (padded 144..256, whose workgroups each fill a CU) run as
// one persistent launch over their jobs, largest order first: the classes
// lie in djobs largest first and contiguous (plan_jacobi and the
// allocation-free plan both lay them out so).  The smaller orders run one
// launch per class.  The launches run concurrently on the runtime's lanes,
// dealt longest first to the lane with the least work so far (work: sum of
// n^3), which keeps the lanes' totals level.
int enqueue_jacobi_classes(msq_runtime* rt, const JacobiClass* cls, int classes,
                           const msq::JacobiJob* djobs, const double* a64, float* V, float* diag,
                           float* roots, msq::JacobiStatus* jstatus) {
  if (classes <= 0) return MSQ_OK;
  if (classes > kMaxJacobiClasses)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "%d Jacobi classes", classes);
  // Launch list: the persistent family as one entry (order 0), then the
  // smaller classes.
  struct Launch {
    size_t first;
    uint32_t count, order;
    double cost;
  };
  Launch launches[kMaxJacobiClasses + 1];
  int nl = 0;
  size_t pfirst = 0, pend = 0;
  uint32_t pcount = 0;
  double pcost = 0.0;
  bool contiguous = true;
  // The persistent family: padded orders 144..256 on the scalar solver, from
  // msq::tile_jacobi_min_padded() on the blocked one.
  const uint32_t pmin = rt->jacobi_method == MSQ_JACOBI_BLOCKED_MFMA
                            ? msq::tile_jacobi_min_padded()
                            : msq::batched_jacobi_padded_order(msq::kPersistentMinOrder);
  for (int c = 0; c < classes; ++c) {
    if (cls[c].order < pmin) continue;
    if (pcount == 0) {
      pfirst = cls[c].first;
    } else if (cls[c].first != pend) {
      contiguous = false;
    }
    pend = cls[c].first + cls[c].count;
    pcount += cls[c].count;
    pcost += cls[c].cost;
  }
  const bool persistent = pcount > 0 && contiguous && rt->jacobi_counter && rt->cu_count > 0;
  if (persistent) launches[nl++] = {pfirst, pcount, 0u, pcost};
  for (int c = 0; c < classes; ++c)
    if (!persistent || cls[c].order < pmin)
      launches[nl++] = {cls[c].first, cls[c].count, cls[c].order, cls[c].cost};
  const int used = nl > 1 ? std::min(nl, kLanes) : 0;
  if (used) {
    HIP_TRY(hipEventRecord(rt->fork, rt->stream));
    for (int l = 0; l < used; ++l) HIP_TRY(hipStreamWaitEvent(rt->lanes[l], rt->fork, 0));
  }
  // Stable order by decreasing cost (insertion sort on the stack).
  int order[kMaxJacobiClasses + 1];
  for (int c = 0; c < nl; ++c) {
    int k = c;
    while (k > 0 && launches[order[k - 1]].cost < launches[c].cost) {
      order[k] = order[k - 1];
      --k;
    }
    order[k] = c;
  }
  double load[kLanes] = {};
  for (int q = 0; q < nl; ++q) {
    const Launch& c = launches[order[q]];
    int lane = 0;
    for (int l = 1; l < used; ++l)
      if (load[l] < load[lane]) lane = l;
    if (used) load[lane] += c.cost;
    hipStream_t s = used ? rt->lanes[lane] : rt->stream;
    if (c.order == 0 && rt->jacobi_method == MSQ_JACOBI_BLOCKED_MFMA) {
      HIP_TRY(msq::launch_tile_jacobi(djobs + c.first, c.count, rt->jacobi_counter, rt->cu_count,
                                      a64, V, diag, roots, jstatus, s));
    } else if (c.order == 0) {
      HIP_TRY(msq::launch_persistent_jacobi(djobs + c.first, c.count, rt->jacobi_counter,
                                            rt->cu_count, a64, V, diag, roots, jstatus, s));
    } else {
      HIP_TRY(msq::launch_batched_jacobi(djobs + c.first, c.count, c.order, a64, V, diag, roots,
                                         jstatus, s));
    }
  }
  for (int l = 0; l < used; ++l) {
    HIP_TRY(hipEventRecord(rt->joins[l], rt->lanes[l]));
    HIP_TRY(hipStreamWaitEvent(rt->stream, rt->joins[l], 0));
  }
  return MSQ_OK;
}

// The batched launches of a plan whose jobs are already on the device.
int enqueue_jacobi(msq_runtime* rt, const JacobiPlan& P, const msq::JacobiJob* djobs,
                   const double* a64, float* V, float* diag, float* roots,
                   msq::JacobiStatus* jstatus) {
  if (P.jobs.empty()) return MSQ_OK;
  const int classes = static_cast<int>(P.launches.size() - 1);
  JacobiClass cls[kMaxJacobiClasses];
  if (classes > kMaxJacobiClasses)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "%d Jacobi classes", classes);
  for (int c = 0; c < classes; ++c) {
    cls[c].first = P.launches[c].first;
    cls[c].count = static_cast<uint32_t>(P.launches[c + 1].first - P.launches[c].first);
    cls[c].order = P.launches[c].second;
    cls[c].cost = P.cost[c];
  }
  return enqueue_jacobi_classes(rt, cls, classes, djobs, a64, V, diag, roots, jstatus);
}

// Jacobi on a whole batch.  djobs: device room for `count` jobs; jstatus:
// device status array in matrix order.
int run_jacobi(msq_runtime* rt, const Layout& L, const double* a64, float* V, float* diag,
               float* roots, const msq_jacobi_options* opts, uint32_t opts_count,
               msq::JacobiStatus* jstatus, msq::JacobiJob* djobs) {
  JacobiPlan P;
  plan_jacobi(L, opts, opts_count, P);
  MSQ_TRY(upload(rt, djobs, P.jobs));
  MSQ_TRY(enqueue_jacobi(rt, P, djobs, a64, V, diag, roots, jstatus));
  return run_large_jacobi(rt, L, a64, V, diag, roots, opts, opts_count, jstatus);
}

void fill_jacobi_report(const msq::JacobiStatus& s, msq_jacobi_report& r) {
  r.sweeps = s.sweeps;
  r.invalid_spectrum = (s.flags & msq::kJacobiInvalidSpectrum) ? 1u : 0u;
  r.off_norm = s.off_norm;
  r.matrix_norm = s.matrix_norm;
  r.min_diagonal = s.min_diagonal;
  r.correction_rate = s.correction_rate;
  for (uint32_t k = 0; k < MSQ_JACOBI_HISTORY; ++k)
    r.off_history[k] = k < s.sweeps && k < msq::kJacobiHistory ? s.off_history[k] : 0.0f;
}

// ---------------------------------------------------------------- GEMM plan
// Tile lists of one group of matrices: 32 x 32 workgroup tiles for ld = 32,
// 64 x 64 for larger pitches; only tiles that meet the order are listed.
struct GemmPlan {
  std::vector<msq::GemmJob> jobs;        // one per matrix of the group
  std::vector<uint32_t> tile_job[3];     // per tile edge kTileEdges[t]
  std::vector<uint32_t> tile_pos[3];
  // Norm epilogues: per-wave slots, tile kind t's from slot_base[t]; job k's
  // slots are [norm_first[k], norm_first[k] + norm_count[k]).
  uint32_t slot_base[3] = {0, 0, 0};
  uint32_t slots = 0;
  std::vector<uint32_t> norm_first, norm_count;
  msq::GemmJob* d_jobs = nullptr;
  uint32_t* d_tile_job[3] = {nullptr, nullptr, nullptr};
  uint32_t* d_tile_pos[3] = {nullptr, nullptr, nullptr};
  uint32_t* d_norm_first = nullptr;
  uint32_t* d_norm_count = nullptr;
  double* d_partials = nullptr;
};

int tile_class(uint32_t ld) { return ld > 256 ? kTileClasses - 1 : static_cast<int>(ld / 32) - 1; }

// Tile edge index for matrices of pitch ld in products of this type: the
// autotuned choice if any, else 32-tiles for ld = 32, 64-tiles up to 256 and
// 128-tiles above (FP64: 64-tiles, whose 4 x 4 binary64 accumulators per
// wave would leave one wave per SIMD).
int tile_kind(const msq_runtime* rt, msq::GemmType type, uint32_t ld) {
  const uint8_t t = rt ? rt->tiles[static_cast<int>(type)][tile_class(ld)] : 0;
  for (int k = 0; k < 3; ++k)
    if (t == kTileEdges[k]) return k;
  return ld == 32 ? 0 : (ld > 256 && type != msq::GemmType::F64 ? 2 : 1);
}

// upper: only the tiles on and above the diagonal (tile row <= tile column),
// for symmetric products whose epilogue mirrors the off-diagonal tiles.
void build_plan(const Layout& L, const std::vector<uint32_t>& members, bool aux_is_matrix,
                GemmPlan& P, const msq_runtime* rt = nullptr,
                msq::GemmType type = msq::GemmType::F32, bool upper = false) {
  P.jobs.clear();
  P.norm_first.clear();
  P.norm_count.clear();
  std::vector<int> kind;
  std::vector<uint32_t> first_tile;
  for (int t = 0; t < 3; ++t) {
    P.tile_job[t].clear();
    P.tile_pos[t].clear();
  }
  for (uint32_t i : members) {
    msq::GemmJob j{};
    j.a_offset = j.b_offset = j.c_offset = L.off[i];
    j.aux_offset = aux_is_matrix ? L.off[i] : L.voff[i];
    j.order = L.order[i];
    j.ld = L.ld[i];
    j.status_index = i;
    const uint32_t jid = static_cast<uint32_t>(P.jobs.size());
    P.jobs.push_back(j);
    const int t = tile_kind(rt, type, L.ld[i]);
    const uint32_t tile = static_cast<uint32_t>(kTileEdges[t]);
    const uint32_t nt = (L.order[i] + tile - 1) / tile;
    kind.push_back(t);
    first_tile.push_back(static_cast<uint32_t>(P.tile_job[t].size()));
    for (uint32_t a = 0; a < nt; ++a)
      for (uint32_t b = upper ? a : 0; b < nt; ++b) {
        P.tile_job[t].push_back(jid);
        P.tile_pos[t].push_back((a << 16) | b);
      }
  }
  uint32_t slots = 0;
  for (int t = 0; t < 3; ++t) {
    P.slot_base[t] = slots;
    slots += static_cast<uint32_t>(P.tile_job[t].size()) * msq::gemm_waves(kTileEdges[t]);
  }
  P.slots = slots;
  for (size_t j = 0; j < P.jobs.size(); ++j) {
    const int t = kind[j];
    const uint32_t w = static_cast<uint32_t>(msq::gemm_waves(kTileEdges[t]));
    // Tiles of one job are consecutive in its kind's list; the next job of
    // the same kind starts where they end.
    uint32_t e = static_cast<uint32_t>(P.tile_job[t].size());
    for (size_t q = j + 1; q < P.jobs.size(); ++q)
      if (kind[q] == t) {
        e = first_tile[q];
        break;
      }
    P.norm_first.push_back(P.slot_base[t] + first_tile[j] * w);
    P.norm_count.push_back((e - first_tile[j]) * w);
  }
}

void plan_want(Scratch& S, GemmPlan& P) {
  S.want(P.d_jobs, std::max<size_t>(1, P.jobs.size()));
  for (int t = 0; t < 3; ++t) {
    S.want(P.d_tile_job[t], std::max<size_t>(1, P.tile_job[t].size()));
    S.want(P.d_tile_pos[t], std::max<size_t>(1, P.tile_pos[t].size()));
  }
  S.want(P.d_norm_first, std::max<size_t>(1, P.norm_first.size()));
  S.want(P.d_norm_count, std::max<size_t>(1, P.norm_count.size()));
  S.want(P.d_partials, std::max<size_t>(1, P.slots));
}

int plan_upload(msq_runtime* rt, GemmPlan& P) {
  MSQ_TRY(upload(rt, P.d_jobs, P.jobs));
  for (int t = 0; t < 3; ++t) {
    MSQ_TRY(upload(rt, P.d_tile_job[t], P.tile_job[t]));
    MSQ_TRY(upload(rt, P.d_tile_pos[t], P.tile_pos[t]));
  }
  MSQ_TRY(upload(rt, P.d_norm_first, P.norm_first));
  MSQ_TRY(upload(rt, P.d_norm_count, P.norm_count));
  return MSQ_OK;
}

// The plan's launches (one per tile edge in use); a Residual or Gram epilogue
// then sums its per-wave slots into args.norms in tile order.
int run_gemm(msq_runtime* rt, const GemmPlan& P, msq::GemmType type, msq::Epilogue epi,
             msq::GemmArgs args) {
  args.jobs = P.d_jobs;
  const bool norm = epi == msq::Epilogue::Residual || epi == msq::Epilogue::Gram;
  args.partials = P.d_partials;
  for (int t = 0; t < 3; ++t) {
    if (P.tile_job[t].empty()) continue;
    args.tile_job = P.d_tile_job[t];
    args.tile_pos = P.d_tile_pos[t];
    args.partial_base = P.slot_base[t];
    HIP_TRY(msq::launch_gemm(type, epi, kTileEdges[t], args,
                             static_cast<uint32_t>(P.tile_job[t].size()), rt->stream));
  }
  if (norm)
    HIP_TRY(msq::launch_gemm_norms(P.d_jobs, P.d_norm_first, P.d_norm_count,
                                   static_cast<uint32_t>(P.jobs.size()), P.d_partials,
                                   args.active, args.norms, rt->stream));
  return MSQ_OK;
}

// ------------------------------------------------------- operand encodings
// One operand of a product in its tier: float data, or codes (+ E8M0 scales
// for MXFP4, + per-matrix sigma for scaled FP8).
struct Operand {
  const void* data = nullptr;
  const uint8_t* scales = nullptr;
  const float* sigma = nullptr;
};

struct TierBuffers {
  void* codes = nullptr;      // operand_bytes(tier, total)
  uint8_t* scales = nullptr;  // total / 32 (MXFP4)
  float* sigma = nullptr;     // count (FP8)
  uint32_t* amax = nullptr;   // count (FP8)
};

void tier_want(Scratch& S, TierBuffers& T, int tier, const Layout& L) {
  const size_t count = L.order.size();
  if (tier == MSQ_FP32) return;
  S.want(reinterpret_cast<uint8_t*&>(T.codes), operand_bytes(tier, L.total));
  if (tier == MSQ_MXFP4_E2M1) S.want(T.scales, L.total / 32);
  if (tier == MSQ_FP8_E4M3FN_SCALED) {
    S.want(T.sigma, count);
    S.want(T.amax, count);
  }
}

// Encodes src (float, batch layout) for the segments of one group.  BF16 and
// MXFP4 are elementwise / 32-block codecs (tiles are whole 32-blocks): one
// launch over the group's span of the batch.  FP8 carries a per-matrix scale:
// one segmented launch.
int encode_operand(msq_runtime* rt, int tier, const Layout& L, const std::vector<uint32_t>& members,
                   const msq::Segment* dsegs, const float* src, TierBuffers& T, Operand& out) {
  if (tier == MSQ_FP32) {
    out.data = src;
    return MSQ_OK;
  }
  if (tier == MSQ_FP8_E4M3FN_SCALED) {
    HIP_TRY(msq::launch_fill_u32(T.amax, L.order.size(), 0u, rt->stream));
    HIP_TRY(msq::launch_segment_fp8(dsegs, static_cast<uint32_t>(members.size()), L.max_ld, src,
                                    T.amax, T.sigma, static_cast<uint8_t*>(T.codes), rt->stream));
    out.data = T.codes;
    out.sigma = T.sigma;
    return MSQ_OK;
  }
  const msq::Format f = tier == MSQ_BF16 ? msq::Format::Bf16 : msq::Format::Mxfp4E2M1;
  // One launch over the span from the first member's tile to the last one's.
  // Tiles of other matrices inside the span are encoded too (their codes are
  // never read): with members interleaved across tiers, one launch over the
  // span is far cheaper than one launch per contiguous run.
  if (!members.empty()) {
    uint64_t begin = UINT64_MAX, end = 0;
    for (uint32_t i : members) {
      begin = std::min<uint64_t>(begin, L.off[i]);
      end = std::max<uint64_t>(end, L.off[i] + static_cast<uint64_t>(L.ld[i]) * L.ld[i]);
    }
    void* dst = f == msq::Format::Bf16
                    ? static_cast<void*>(static_cast<uint16_t*>(T.codes) + begin)
                    : static_cast<void*>(static_cast<uint8_t*>(T.codes) + begin / 2);
    HIP_TRY(msq::launch_encode(f, src + begin, dst, T.scales ? T.scales + begin / 32 : nullptr,
                               end - begin, rt->stream));
  }
  out.data = T.codes;
  out.scales = T.scales;
  return MSQ_OK;
}

msq::GemmArgs product_args(const Operand& a, const Operand& b, void* c, const float* alpha,
                           const int* active) {
  msq::GemmArgs g{};
  g.a = a.data;
  g.b = b.data;
  g.a_scales = a.scales;
  g.b_scales = b.scales;
  g.alpha = alpha;
  g.active = active;
  g.c = c;
  return g;
}

}  // namespace

namespace {

// 32 x 32 tile pairs {(ti, tj), (tj, ti)}, ti <= tj, of a pitch.
uint32_t tile_pairs(uint32_t ld) {
  const uint32_t t = ld / 32;
  return t * (t + 1) / 2;
}

// Segments of the members, in member order; `reserved` counts the tile pairs
// of the segments before each one (the tile-pair passes' workgroup map).
std::vector<msq::Segment> segments_of(const Layout& L, const std::vector<uint32_t>& members) {
  std::vector<msq::Segment> segs;
  segs.reserve(members.size());
  uint32_t pairs = 0;
  for (uint32_t i : members) {
    msq::Segment g{};
    g.offset = L.off[i];
    g.vec_offset = L.voff[i];
    g.order = L.order[i];
    g.ld = L.ld[i];
    g.index = i;
    g.reserved = pairs;
    pairs += tile_pairs(L.ld[i]);
    segs.push_back(g);
  }
  return segs;
}

uint32_t pair_count(const std::vector<msq::Segment>& segs) {
  return segs.empty() ? 0u : segs.back().reserved + tile_pairs(segs.back().ld);
}
uint32_t pair_count(const Layout& L, const std::vector<uint32_t>& members) {
  uint32_t pairs = 0;
  for (uint32_t i : members) pairs += tile_pairs(L.ld[i]);
  return pairs;
}

std::vector<uint32_t> iota_members(uint32_t count) {
  std::vector<uint32_t> m(count);
  for (uint32_t i = 0; i < count; ++i) m[i] = i;
  return m;
}

// Products of one group share these device buffers.
struct CorrectionBuffers {
  float* t1 = nullptr;  // float, batch layout
  float* t2 = nullptr;
  TierBuffers enc_a, enc_b;
  float* alpha = nullptr;  // count
};

int fp8_alpha(msq_runtime* rt, int tier, const Operand& a, const Operand& b, float* alpha,
              uint32_t count, const float** out) {
  if (tier != MSQ_FP8_E4M3FN_SCALED) {
    *out = nullptr;
    return MSQ_OK;
  }
  HIP_TRY(msq::launch_alpha(a.sigma, b.sigma, alpha, count, rt->stream));
  *out = alpha;
  return MSQ_OK;
}

// One frozen-factor correction for a group (refineUpdateAndMeasure's product
// chain, src/metal_runtime.cpp:867-928):
//   W = V^T R ; Y = (W V) o H ; W2 = V Y ; E = W2 V^T   (E into cb.t2)
// With x_update (the solve's binary64 iterate X) and upper (the group's plan
// of the tiles on and above the diagonal): Y = (V^T R V) o H is symmetric, so
// only its upper tiles are computed and stored mirrored, and the last product
// adds its upper triangle to X at (i, j) and (j, i) instead of storing E
// (working_add_symmetric_correction's X += sym(E), .metal:1955-1970, fused
// into the product; E's two triangles differ only by the rounding of W2).
// ov / ovt: V and V^T encoded in the group's tier; r32: R as float; rb: R
// as BF16 codes (batch layout) or null.  In the BF16 tier the intermediate
// products store their codes from the GEMM epilogue (the same RNE codec as
// the encode pass, so the operands are bit-identical) and rb replaces the
// encode of r32.
int correction_products(msq_runtime* rt, int tier, const Layout& L,
                        const std::vector<uint32_t>& members, const msq::Segment* dsegs,
                        const GemmPlan& plan, const Operand& ov, const Operand& ovt,
                        const float* r32, const uint16_t* rb, const float* roots,
                        const int* active, CorrectionBuffers& cb, double* x_update = nullptr,
                        const GemmPlan* upper = nullptr) {
  const uint32_t count = static_cast<uint32_t>(L.order.size());
  const bool fused = x_update && upper;
  const msq::GemmType type = gemm_type(tier);
  const bool direct = tier == MSQ_BF16;
  const float* alpha = nullptr;
  Operand opr, opa;
  // Output of an intermediate product: codes of `enc` (BF16) or float `t`.
  auto out_to = [&](msq::GemmArgs& g, TierBuffers& enc, float* t) {
    if (direct) {
      g.c = nullptr;
      g.c_bf16 = static_cast<uint16_t*>(enc.codes);
    } else {
      g.c = t;
    }
  };
  auto operand_of = [&](TierBuffers& enc, const float* t, Operand& op) -> int {
    if (direct) {
      op = Operand{};
      op.data = enc.codes;
      return MSQ_OK;
    }
    return encode_operand(rt, tier, L, members, dsegs, t, enc, op);
  };
  // W^T <- (V^T R)^T
  if (direct && rb) {
    opr = Operand{};
    opr.data = rb;
  } else {
    MSQ_TRY(encode_operand(rt, tier, L, members, dsegs, r32, cb.enc_b, opr));
  }
  MSQ_TRY(fp8_alpha(rt, tier, ov, opr, cb.alpha, count, &alpha));
  msq::GemmArgs g = product_args(ov, opr, cb.t1, alpha, active);
  g.transpose_c = 1;
  out_to(g, cb.enc_a, cb.t1);
  MSQ_TRY(run_gemm(rt, plan, type, msq::Epilogue::Store, g));
  // Y <- (W V) o H
  MSQ_TRY(operand_of(cb.enc_a, cb.t1, opa));
  MSQ_TRY(fp8_alpha(rt, tier, opa, ov, cb.alpha, count, &alpha));
  g = product_args(opa, ov, cb.t2, alpha, active);
  g.aux = roots;
  g.mirror = fused ? 1 : 0;
  out_to(g, cb.enc_b, cb.t2);
  MSQ_TRY(run_gemm(rt, fused ? *upper : plan, type, msq::Epilogue::Hadamard, g));
  // W2^T storage <- V Y, row-major
  MSQ_TRY(operand_of(cb.enc_b, cb.t2, opr));
  MSQ_TRY(fp8_alpha(rt, tier, ovt, opr, cb.alpha, count, &alpha));
  g = product_args(ovt, opr, cb.t1, alpha, active);
  g.transpose_c = 1;
  out_to(g, cb.enc_a, cb.t1);
  MSQ_TRY(run_gemm(rt, plan, type, msq::Epilogue::Store, g));
  // E <- W2 V^T
  MSQ_TRY(operand_of(cb.enc_a, cb.t1, opa));
  MSQ_TRY(fp8_alpha(rt, tier, opa, ovt, cb.alpha, count, &alpha));
  g = product_args(opa, ovt, fused ? static_cast<void*>(x_update) : cb.t2, alpha, active);
  MSQ_TRY(run_gemm(rt, fused ? *upper : plan, type,
                   fused ? msq::Epilogue::AddSym : msq::Epilogue::Store, g));
  return MSQ_OK;
}

// X0 = G G^T, G = V diag(sqrt(s)) in the tier (into cb.t2).
//
// two_term (MXFP4): G is carried as a two-term expansion hi + lo, hi the
// MXFP4 rounding of G and lo the MXFP4 rounding of G - hi, and the product is
// hi hi^T + 2 hi lo^T (lo lo^T dropped): the symmetrisation that follows,
// (C + C^T) / 2, turns it into hi hi^T + hi lo^T + lo hi^T.  One E2M1
// mantissa bit leaves X0 about 0.17 off the root at n = 4096, outside the
// range where three corrections reach 1e-6; the expansion brings it to about
// 0.03 with one more scaled MFMA product.
int reconstruction_product(msq_runtime* rt, int tier, const Layout& L,
                           const std::vector<uint32_t>& members, const msq::Segment* dsegs,
                           const GemmPlan& plan, const float* V, const float* roots,
                           const int* active, CorrectionBuffers& cb, bool two_term = false) {
  const uint32_t count = static_cast<uint32_t>(L.order.size());
  // Only the matrices the product computes (active: the solve's running
  // matrices, or those whose low-bit basis was rejected).  BF16: G goes
  // straight to its codes.
  Operand og, olo;
  if (tier == MSQ_BF16) {
    HIP_TRY(msq::launch_scale_rows_transpose_bf16(dsegs, static_cast<uint32_t>(members.size()),
                                                  pair_count(L, members), V, roots, 1,
                                                  static_cast<uint16_t*>(cb.enc_a.codes), active,
                                                  rt->stream));
    og.data = cb.enc_a.codes;
  } else {
    HIP_TRY(msq::launch_scale_rows_transpose(dsegs, static_cast<uint32_t>(members.size()),
                                             pair_count(L, members),
                                             V, roots, 1, cb.t1, active, rt->stream));
    MSQ_TRY(encode_operand(rt, tier, L, members, dsegs, cb.t1, cb.enc_a, og));
  }
  two_term = two_term && tier == MSQ_MXFP4_E2M1 && !members.empty();
  if (two_term) {
    // lo = mxfp4(G - dec(hi)) over the members' span of the batch (the span
    // the encode covers); cb.t2 holds dec(hi) until the products overwrite it.
    uint64_t begin = UINT64_MAX, end = 0;
    for (uint32_t i : members) {
      begin = std::min<uint64_t>(begin, L.off[i]);
      end = std::max<uint64_t>(end, L.off[i] + static_cast<uint64_t>(L.ld[i]) * L.ld[i]);
    }
    HIP_TRY(msq::launch_decode(msq::Format::Mxfp4E2M1,
                               static_cast<const uint8_t*>(cb.enc_a.codes) + begin / 2,
                               cb.enc_a.scales + begin / 32, cb.t2 + begin, end - begin,
                               rt->stream));
    HIP_TRY(msq::launch_subtract(cb.t1 + begin, cb.t2 + begin, cb.t1 + begin, end - begin,
                                 rt->stream));
    MSQ_TRY(encode_operand(rt, tier, L, members, dsegs, cb.t1, cb.enc_b, olo));
  }
  const float* alpha = nullptr;
  MSQ_TRY(fp8_alpha(rt, tier, og, og, cb.alpha, count, &alpha));
  msq::GemmArgs g = product_args(og, og, cb.t2, alpha, active);
  MSQ_TRY(run_gemm(rt, plan, gemm_type(tier), msq::Epilogue::Store, g));
  if (two_term) {
    g = product_args(og, olo, cb.t2, nullptr, active);
    g.c_scale = 2.0f;
    MSQ_TRY(run_gemm(rt, plan, gemm_type(tier), msq::Epilogue::Accumulate, g));
  }
  return MSQ_OK;
}

// R = A - X X in binary64 with ||R||_F^2 per matrix; R (binary64), R32
// (float) and Rb (BF16 codes) are each stored when non-null.  symmetric_x:
// X is exactly symmetric (the solve's iterates are), so (X X)^T = X X; the
// plan then lists the tiles on and above the diagonal (build_plan upper) and
// the epilogue stores each off-diagonal tile's mirror, A_ji - (X X)_ij (half
// the products).
// Basis acceptance tolerance of a tier (basisOrthogonalityTolerance,
// src/configuration.cpp:109-118): BF16 5e-3, 4-bit formats 0.5, the other
// 8-bit formats 0.15.
double basis_tolerance(int tier) {
  switch (tier) {
    case MSQ_FP32: return 1.0e-3;
    case MSQ_BF16: return 5.0e-3;
    case MSQ_FP4_E2M1:
    case MSQ_MXFP4_E2M1:
    case MSQ_FP4_E2M1_LOG_S: return 5.0e-1;
    default: return 1.5e-1;
  }
}

int residual_product(msq_runtime* rt, const GemmPlan& plan, const double* A, const double* X,
                     double* R, float* R32, uint16_t* Rb, double* norms, const int* active,
                     uint32_t count, bool symmetric_x = true) {
  HIP_TRY(msq::launch_zero_f64(norms, count, rt->stream));
  msq::GemmArgs g{};
  g.a = X;
  g.b = X;
  g.c = R;
  g.aux = A;
  g.c_float = R32;
  g.c_bf16 = Rb;
  g.mirror = symmetric_x ? 1 : 0;
  g.norms = norms;
  g.active = active;
  MSQ_TRY(run_gemm(rt, plan, msq::GemmType::F64, msq::Epilogue::Residual, g));
  return MSQ_OK;
}

void want_correction(Scratch& S, CorrectionBuffers& cb, const Layout& L, int max_tier) {
  S.want(cb.t1, L.total);
  S.want(cb.t2, L.total);
  tier_want(S, cb.enc_a, max_tier, L);
  tier_want(S, cb.enc_b, max_tier, L);
  S.want(cb.alpha, L.order.size());
}

// The widest tier needs the most code bytes; FP8 needs sigma arrays, MXFP4
// scale planes: reserve all three shapes when several tiers share buffers.
void tier_want_all(Scratch& S, TierBuffers& T, const Layout& L) {
  const size_t count = L.order.size();
  S.want(reinterpret_cast<uint8_t*&>(T.codes), operand_bytes(MSQ_BF16, L.total));
  S.want(T.scales, L.total / 32);
  S.want(T.sigma, count);
  S.want(T.amax, count);
}

int validate_tier(int p, const char* what) {
  if (!product_supported(p))
    return set_error(MSQ_ERROR_UNSUPPORTED,
                     "%s precision %d has no MFMA product (use FP32, BF16, FP8_E4M3FN_SCALED "
                     "or MXFP4_E2M1)",
                     what, p);
  return MSQ_OK;
}

// ---------------------------------------------------------- native stages
// One set of matrices corrected in a native storage format (the reference's
// refineNativeStaged, src/metal_runtime.cpp:312-411): every operand and
// product stored in the format on the compact n x n layout, products formed
// in FP32 on the decoded values.
struct NativeSet {
  int format = 0;
  uint32_t count = 0;
  uint32_t max_order = 0;
  const msq::Segment* dsegs = nullptr;
  const uint64_t* dcoff = nullptr;
  const GemmPlan* plan = nullptr;  // F32 tiles of the set's matrices
  uint8_t* codes = nullptr;        // compact scratch of one stage
  uint8_t* scales = nullptr;
};

// `in` rounded to the set's format (read transposed when in_t), decoded into
// `out` (stored transposed when out_t): one matmul + quantize pair's output.
int native_round(msq_runtime* rt, const Layout& L, const NativeSet& g, const float* in, bool in_t,
                 float* out, bool out_t, const int* active) {
  HIP_TRY(msq::launch_compact_quantize(g.dsegs, g.dcoff, g.count, g.max_order, g.format, in, in_t,
                                       g.codes, g.scales, active, rt->stream));
  HIP_TRY(msq::launch_compact_dequantize(g.dsegs, g.dcoff, g.count, L.max_ld, g.format, g.codes,
                                         g.scales, out, out_t, active, rt->stream));
  return MSQ_OK;
}

// C = A B on decoded float operands: `at` holds A^T column-major (the GEMM's
// K-contiguous left operand), b holds B.
int native_product(msq_runtime* rt, const NativeSet& g, const float* at, const float* b, float* c,
                   const int* active) {
  Operand oa, ob;
  oa.data = at;
  ob.data = b;
  return run_gemm(rt, *g.plan, msq::GemmType::F32, msq::Epilogue::Store,
                  product_args(oa, ob, c, nullptr, active));
}

// The stored basis: Q = quant(V) and Q_t = quant(V^T) (the reference stores
// the transposed eigenvectors separately, src/metal_solver.cpp:707-711),
// each decoded as is and transposed.
int native_basis(msq_runtime* rt, const Layout& L, const NativeSet& g, const float* V, float* q,
                 float* qT, float* qt, float* qtT, const int* active) {
  MSQ_TRY(native_round(rt, L, g, V, false, q, false, active));
  MSQ_TRY(native_round(rt, L, g, V, false, qT, true, active));
  MSQ_TRY(native_round(rt, L, g, V, true, qt, false, active));
  MSQ_TRY(native_round(rt, L, g, V, true, qtT, true, active));
  return MSQ_OK;
}

// One staged correction from R (float): U = Q_t R, C = U Q, Y = C / (s_i +
// s_j), V2 = Q Y, E = V2 Q_t, each stored in the format; E decoded into t2.
int native_correction(msq_runtime* rt, const Layout& L, const NativeSet& g, const float* r32,
                      const float* q, const float* qT, const float* qt, const float* qtT,
                      const float* s, float* t1, float* t2, const int* active) {
  MSQ_TRY(native_round(rt, L, g, r32, false, t2, false, active));  // R
  MSQ_TRY(native_product(rt, g, qtT, t2, t1, active));             // U = Q_t R
  MSQ_TRY(native_round(rt, L, g, t1, false, t2, true, active));    // U, transposed
  MSQ_TRY(native_product(rt, g, t2, q, t1, active));               // C = U Q
  MSQ_TRY(native_round(rt, L, g, t1, false, t2, false, active));
  HIP_TRY(msq::launch_sylvester_divide(g.dsegs, g.count, L.max_ld, t2, s, t1, active, rt->stream));
  MSQ_TRY(native_round(rt, L, g, t1, false, t2, false, active));   // Y
  MSQ_TRY(native_product(rt, g, qT, t2, t1, active));              // V2 = Q Y
  MSQ_TRY(native_round(rt, L, g, t1, false, t2, true, active));    // V2, transposed
  MSQ_TRY(native_product(rt, g, t2, qt, t1, active));              // E = V2 Q_t
  MSQ_TRY(native_round(rt, L, g, t1, false, t2, false, active));   // E
  return MSQ_OK;
}

// Compact element offsets of the members (n^2 from a 64-element boundary)
// and the largest order; returns the compact total.
uint64_t compact_offsets(const Layout& L, const std::vector<uint32_t>& members,
                         std::vector<uint64_t>& coff, uint32_t& max_order) {
  uint64_t c = 0;
  coff.clear();
  max_order = 0;
  for (uint32_t i : members) {
    coff.push_back(c);
    c += (static_cast<uint64_t>(L.order[i]) * L.order[i] + 63) / 64 * 64;
    max_order = std::max(max_order, L.order[i]);
  }
  return c;
}

bool native_format(int p) { return p >= MSQ_FP8_E4M3FN && p <= MSQ_MXFP4_E2M1; }

// Host-side facts of a solve that its reports need (per matrix).
struct SolveMeta {
  std::vector<uint8_t> stagnation;  // MSQ_STOPPING_STAGNATION: the best iterate is the result
  std::vector<uint8_t> native;      // native-quantized corrections (no certificate)
  std::vector<uint8_t> low_basis;   // a low-bit basis tier was tried
  std::vector<uint8_t> refined;     // the second Jacobi pass ran
  std::vector<double> residual_tolerance;
};

// The reference's checks of a solver configuration (runMetalCase,
// src/metal_solver.cpp:359-370; validateNativeStoppingPolicy,
// validateNativeDispatchPolicy and validateCorrectionMode,
// src/configuration.cpp:290-346) and this backend's limits.
int validate_solve_options(const msq_solve_options& o) {
  auto bad = [](const char* what) { return set_error(MSQ_ERROR_INVALID_ARGUMENT, "%s", what); };
  if (o.kernel_path != MSQ_KERNEL_PATH_LEGACY && o.kernel_path != MSQ_KERNEL_PATH_TILED)
    return bad("kernel_path must be MSQ_KERNEL_PATH_LEGACY or MSQ_KERNEL_PATH_TILED");
  if (o.correction_mode != MSQ_CORRECTION_FP32_REORTHOGONALIZED &&
      o.correction_mode != MSQ_CORRECTION_NATIVE_QUANTIZED)
    return bad("correction_mode must be fp32_reorthogonalized or native_quantized");
  if (o.stopping_policy != MSQ_STOPPING_FIXED_BUDGET && o.stopping_policy != MSQ_STOPPING_STAGNATION)
    return bad("stopping_policy must be fixed_budget or stagnation");
  if (o.dispatch_policy != MSQ_DISPATCH_PER_STAGE && o.dispatch_policy != MSQ_DISPATCH_FUSED)
    return bad("dispatch_policy must be per_stage or fused");
  const bool native = o.correction_mode == MSQ_CORRECTION_NATIVE_QUANTIZED;
  if (o.stopping_policy == MSQ_STOPPING_STAGNATION && !native)
    return bad("stagnation stopping requires native_quantized correction");
  if (o.dispatch_policy == MSQ_DISPATCH_PER_STAGE && !native)
    return bad("per_stage dispatch requires native_quantized correction");
  if (o.stagnation_window == 0) return bad("stagnation_window must be a positive integer");
  if (!std::isfinite(o.minimum_relative_improvement) || o.minimum_relative_improvement < 0.0)
    return bad("minimum_relative_improvement must be finite and nonnegative");
  if (o.max_corrections >= msq::kSolveHistory)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "max_corrections must be < %u",
                     msq::kSolveHistory);
  if (!(o.resolution >= 0.0f) || !std::isfinite(o.resolution))
    return bad("resolution must be finite and nonnegative");
  if (!(o.basis_tolerance >= 0.0)) return bad("basis_tolerance must be nonnegative");
  if (o.spectrum_precision < MSQ_FP32 || o.spectrum_precision > MSQ_FP4_E2M1_LOG_S)
    return set_error(MSQ_ERROR_UNSUPPORTED, "spectrum precision %d", o.spectrum_precision);
  if (is_log_spectrum(o.basis_precision))
    return bad("a log-spectrum format cannot store the frozen basis");
  if (native) {
    if (is_log_spectrum(o.spectrum_precision))
      return bad("native quantized correction does not support log-spectrum storage");
    if (o.basis_precision != o.spectrum_precision)
      return bad("native quantized correction requires matching basis and spectrum precisions");
    if (o.kernel_path != MSQ_KERNEL_PATH_LEGACY)
      return bad("native quantized correction requires the legacy kernel path");
    if (o.basis_precision < MSQ_FP8_E4M3FN || o.basis_precision > MSQ_MXFP4_E2M1)
      return bad("native quantized correction supports FP8 E4M3FN, FP8 E5M2, MXFP8 E4M3FN, "
                 "MXFP8 E5M2, FP4 E2M1, or MXFP4 E2M1 storage");
    if (o.correction_precision != o.basis_precision)
      return bad("native quantized correction runs its products in the basis storage format: "
                 "correction_precision must equal basis_precision");
    return MSQ_OK;
  }
  MSQ_TRY(validate_tier(o.basis_precision, "basis"));
  return validate_tier(o.correction_precision, "correction");
}

// Second Jacobi pass of the matrices whose smallest eigenvalue the FP32
// Jacobi does not resolve (solve state `unresolved`, options with
// refine_unresolved).  B = V^T A V is formed in binary64 (two FP64 MFMA
// products on V widened exactly), so its entries carry the small eigenvalues
// to full float precision, where the first pass only reached eps32 ||A||;
// the batched Jacobi then diagonalises B (stopping once the frozen-factor
// correction is predicted to contract by 1e-2), V <- V W, and the solve
// state is rebuilt from the new diagonal.  Host-driven (it reads the flags
// back), so such solves run eagerly.
int refine_unresolved(msq_runtime* rt, const Layout& L, const std::vector<msq::SolveLimits>& limits,
                      const double* matrices, float* V, float* diag, float* sroot,
                      msq::JacobiStatus* jst, msq::SolveState* st, const msq::Segment* dsegs_all,
                      const msq::SolveLimits* dlim, int* active, SolveMeta& meta) {
  const uint32_t count = static_cast<uint32_t>(L.order.size());
  std::vector<msq::SolveState> hs(count);
  HIP_TRY(hipMemcpyAsync(hs.data(), st, count * sizeof(msq::SolveState), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  std::vector<uint32_t> F;
  for (uint32_t i = 0; i < count; ++i)
    if (limits[i].refine && hs[i].unresolved && hs[i].status == msq::kSolveRunning) F.push_back(i);
  if (F.empty()) return MSQ_OK;
  const uint32_t nf = static_cast<uint32_t>(F.size());
  std::vector<uint32_t> orders_f(nf);
  for (uint32_t k = 0; k < nf; ++k) orders_f[k] = L.order[F[k]];
  Layout LF;
  MSQ_TRY(make_layout(nf, orders_f.data(), LF));
  const std::vector<uint32_t> sub = iota_members(nf);
  const std::vector<msq::Segment> segs_from = segments_of(L, F), segs_to = segments_of(LF, sub);
  GemmPlan p64, p32;
  build_plan(LF, sub, false, p64, rt, msq::GemmType::F64);
  build_plan(LF, sub, false, p32, rt, msq::GemmType::F32);
  Scratch S{rt, &rt->refine};
  double *a64 = nullptr, *v64 = nullptr, *c64 = nullptr, *b64 = nullptr;
  float *vs = nullptr, *vst = nullptr, *w = nullptr, *vn = nullptr, *dg = nullptr, *rs = nullptr;
  msq::JacobiStatus *js = nullptr, *jcomb = nullptr;
  msq::JacobiJob* jobs = nullptr;
  msq::Segment *dfrom = nullptr, *dto = nullptr;
  uint32_t* dindex = nullptr;
  S.want(a64, LF.total);
  S.want(v64, LF.total);
  S.want(c64, LF.total);
  S.want(b64, LF.total);
  S.want(vs, LF.total);
  S.want(vst, LF.total);
  S.want(w, LF.total);
  S.want(vn, LF.total);
  S.want(dg, LF.vtotal);
  S.want(rs, LF.vtotal);
  S.want(js, nf);
  S.want(jcomb, nf);
  S.want(jobs, nf);
  S.want(dfrom, nf);
  S.want(dto, nf);
  S.want(dindex, nf);
  plan_want(S, p64);
  plan_want(S, p32);
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dfrom, segs_from));
  MSQ_TRY(upload(rt, dto, segs_to));
  MSQ_TRY(upload(rt, dindex, F));
  MSQ_TRY(plan_upload(rt, p64));
  MSQ_TRY(plan_upload(rt, p32));
  const uint32_t max_ld = L.max_ld;
  HIP_TRY(msq::launch_move_tiles(dfrom, dto, nf, max_ld, matrices, a64, sizeof(double), rt->stream));
  HIP_TRY(msq::launch_move_tiles(dfrom, dto, nf, max_ld, V, vs, sizeof(float), rt->stream));
  HIP_TRY(msq::launch_widen_f64(vs, v64, LF.total, rt->stream));
  Operand oa, ob;
  oa.data = a64;  // A symmetric: its K-contiguous storage is itself
  ob.data = v64;
  MSQ_TRY(run_gemm(rt, p64, msq::GemmType::F64, msq::Epilogue::Store,
                   product_args(oa, ob, c64, nullptr, nullptr)));  // C = A V
  oa.data = v64;  // V^T's K-contiguous storage is V column-major
  ob.data = c64;
  MSQ_TRY(run_gemm(rt, p64, msq::GemmType::F64, msq::Epilogue::Store,
                   product_args(oa, ob, b64, nullptr, nullptr)));  // B = V^T C
  msq_jacobi_options jo{};
  jo.max_sweeps = 30;
  jo.off_tol = 0.0f;      // off(B) is at the FP32 floor of A already
  jo.stagnation = 0.5f;
  jo.floor_tol = 1e-30f;  // no floor: the small block converges below it
  jo.correction_rate = 1e-2f;
  MSQ_TRY(run_jacobi(rt, LF, b64, w, dg, rs, &jo, 1, js, jobs));
  HIP_TRY(msq::launch_transpose(dto, nf, pair_count(segs_to), vs, vst, rt->stream));
  oa.data = vst;
  ob.data = w;
  MSQ_TRY(run_gemm(rt, p32, msq::GemmType::F32, msq::Epilogue::Store,
                   product_args(oa, ob, vn, nullptr, nullptr)));  // V W
  HIP_TRY(msq::launch_move_tiles(dto, dfrom, nf, max_ld, vn, V, sizeof(float), rt->stream));
  HIP_TRY(msq::launch_move_vectors(dto, dfrom, nf, max_ld, dg, diag, rt->stream));
  HIP_TRY(msq::launch_move_vectors(dto, dfrom, nf, max_ld, rs, sroot, rt->stream));
  // Jacobi reports: the two passes' sweeps and histories in sequence, the
  // second pass' final measures.
  std::vector<msq::JacobiStatus> first(count), second(nf);
  HIP_TRY(hipMemcpyAsync(first.data(), jst, count * sizeof(msq::JacobiStatus),
                         hipMemcpyDeviceToHost, rt->stream));
  HIP_TRY(hipMemcpyAsync(second.data(), js, nf * sizeof(msq::JacobiStatus), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  std::vector<msq::JacobiStatus> comb(nf);
  for (uint32_t k = 0; k < nf; ++k) {
    msq::JacobiStatus c = first[F[k]];
    const msq::JacobiStatus& b = second[k];
    for (uint32_t q = 0; q < b.sweeps && c.sweeps + q < msq::kJacobiHistory; ++q)
      c.off_history[c.sweeps + q] = b.off_history[q];
    c.sweeps += b.sweeps;
    c.flags = b.flags;  // the second pass decides the spectrum
    c.off_norm = b.off_norm;
    c.min_diagonal = b.min_diagonal;
    c.correction_rate = b.correction_rate;
    comb[k] = c;
    meta.refined[F[k]] = 1;
  }
  MSQ_TRY(upload(rt, jcomb, comb));
  HIP_TRY(msq::launch_scatter_status(jcomb, dindex, nf, jst, rt->stream));
  HIP_TRY(msq::launch_solve_state(dsegs_all, jst, diag, dlim, st, active, count, false,
                                  rt->stream));
  // comb is a host temporary read by the upload above.
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

}  // namespace

// ============================================================== C entry points
// ----------------------------------------------------------- solve graphs
// A captured solve: the device work of one batch signature (orders, options,
// input and output pointers) as a HIP graph, with the buffers it reads in an
// arena of its own.
struct GraphEntry {
  std::vector<uint32_t> orders;
  std::vector<unsigned char> options;
  const void* matrices = nullptr;
  msq_solve_outputs outputs{};
  Arena arena;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Correction steps after `split` in a graph of their own, launched only
  // while some matrix is still running after the first graph; `split` is
  // the last step that the eager call needed.
  hipGraph_t tail_graph = nullptr;
  hipGraphExec_t tail_exec = nullptr;
  uint32_t split = UINT32_MAX;
  int calls = 0;  // eager calls so far (the second one is captured)
  Layout layout;
  msq::SolveState* st = nullptr;
  msq::JacobiStatus* jst = nullptr;
  float* V = nullptr;
  float* sroot = nullptr;
  int* active = nullptr;
  std::vector<int> flags;  // host copy of `active` at the split
};

namespace {

void release_graph(msq_runtime* rt, GraphEntry& g) {
  // msq_last_factors' pointers into this arena die with it.
  if (rt->last_vectors && rt->last_vectors == g.V) {
    rt->last_vectors = nullptr;
    rt->last_roots = nullptr;
  }
  if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (g.graph) (void)hipGraphDestroy(g.graph);
  if (g.tail_exec) (void)hipGraphExecDestroy(g.tail_exec);
  if (g.tail_graph) (void)hipGraphDestroy(g.tail_graph);
  if (g.arena.base) (void)hipFree(g.arena.base);
  g = GraphEntry{};
}

bool same_outputs(const msq_solve_outputs& a, const msq_solve_outputs& b) {
  return a.roots == b.roots && a.initial_roots == b.initial_roots && a.vectors == b.vectors &&
         a.root_eigenvalues == b.root_eigenvalues;
}

// The entry of this signature (created on first use; at most kMaxGraphs,
// the oldest evicted).
GraphEntry* find_graph(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                       const double* matrices, const msq_solve_outputs& outputs,
                       const msq_solve_options* options, uint32_t options_count) {
  const unsigned char* ob = reinterpret_cast<const unsigned char*>(options);
  const size_t obytes = static_cast<size_t>(options_count) * sizeof(msq_solve_options);
  for (auto& g : rt->graphs)
    if (g->matrices == matrices && same_outputs(g->outputs, outputs) && g->orders.size() == count &&
        std::equal(g->orders.begin(), g->orders.end(), orders) && g->options.size() == obytes &&
        std::equal(g->options.begin(), g->options.end(), ob))
      return g.get();
  if (rt->graphs.size() >= kMaxGraphs) {
    (void)hipStreamSynchronize(rt->stream);
    release_graph(rt, *rt->graphs.front());
    rt->graphs.erase(rt->graphs.begin());
  }
  auto g = std::make_unique<GraphEntry>();
  g->orders.assign(orders, orders + count);
  g->options.assign(ob, ob + obytes);
  g->matrices = matrices;
  g->outputs = outputs;
  rt->graphs.push_back(std::move(g));
  return rt->graphs.back().get();
}

// Graphs are captured and launched on a stream of the runtime's own (the
// null stream cannot be captured), ordered after and before the caller's
// stream with events.
template <typename F>
int capture_graph(msq_runtime* rt, hipGraph_t& out_graph, hipGraphExec_t& out_exec, F&& enqueue) {
  hipStream_t caller = rt->stream;
  HIP_TRY(hipEventRecord(rt->graph_in, caller));
  HIP_TRY(hipStreamWaitEvent(rt->own_stream, rt->graph_in, 0));
  rt->stream = rt->own_stream;
  hipError_t e = hipStreamBeginCapture(rt->stream, hipStreamCaptureModeRelaxed);
  int status = e == hipSuccess ? enqueue() : MSQ_OK;
  hipGraph_t graph = nullptr;
  const hipError_t e2 = e == hipSuccess ? hipStreamEndCapture(rt->stream, &graph) : e;
  rt->stream = caller;
  if (e != hipSuccess || e2 != hipSuccess || status != MSQ_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    if (status != MSQ_OK) return status;
    return set_error(MSQ_ERROR_HIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(graph);
    return set_error(MSQ_ERROR_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  out_graph = graph;
  out_exec = exec;
  return MSQ_OK;
}

int launch_exec(msq_runtime* rt, hipGraphExec_t exec) {
  HIP_TRY(hipEventRecord(rt->graph_in, rt->stream));
  HIP_TRY(hipStreamWaitEvent(rt->own_stream, rt->graph_in, 0));
  HIP_TRY(hipGraphLaunch(exec, rt->own_stream));
  HIP_TRY(hipEventRecord(rt->graph_out, rt->own_stream));
  HIP_TRY(hipStreamWaitEvent(rt->stream, rt->graph_out, 0));
  return MSQ_OK;
}

// The entry's first graph, then its tail graph if some matrix is still
// running (the host reads the per-matrix running flags in between).
int launch_graph(msq_runtime* rt, GraphEntry& g) {
  MSQ_TRY(launch_exec(rt, g.exec));
  if (!g.tail_exec) return MSQ_OK;
  const uint32_t count = static_cast<uint32_t>(g.layout.order.size());
  g.flags.resize(count);
  HIP_TRY(hipMemcpyAsync(g.flags.data(), g.active, count * sizeof(int), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  for (uint32_t i = 0; i < count; ++i)
    if (g.flags[i]) return launch_exec(rt, g.tail_exec);
  return MSQ_OK;
}


// Reports of a solve whose device work is enqueued: copies the per-matrix
// states back and waits.  Status mapping: a running matrix has exhausted its
// budget; under the stagnation policy the best iterate decides convergence
// and the reference's termination reasons (src/metal_solver.cpp:904-935); a
// matrix that stops converged while unresolved reports MSQ_SOLVE_UNRESOLVED.
int finish_solve(msq_runtime* rt, const Layout& L, const msq::SolveState* st,
                 const msq::JacobiStatus* jst, float* V, float* sroot, msq_solve_report* reports,
                 const SolveMeta& meta, uint32_t* max_steps = nullptr) {
  const uint32_t count = static_cast<uint32_t>(L.order.size());
  std::vector<msq::SolveState> hs(count);
  std::vector<msq::JacobiStatus> hj(count);
  HIP_TRY(hipMemcpyAsync(hs.data(), st, count * sizeof(msq::SolveState), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipMemcpyAsync(hj.data(), jst, count * sizeof(msq::JacobiStatus), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  if (max_steps) {
    *max_steps = 0;
    for (uint32_t i = 0; i < count; ++i) *max_steps = std::max(*max_steps, hs[i].steps);
  }
  if (reports) {
    for (uint32_t i = 0; i < count; ++i) {
      msq_solve_report& r = reports[i];
      std::memset(&r, 0, sizeof(r));
      const msq::SolveState& s = hs[i];
      r.order = L.order[i];
      r.sweeps = hj[i].sweeps;
      r.corrections = s.corrections;
      uint32_t status = s.status == msq::kSolveRunning ? msq::kSolveBudgetExhausted : s.status;
      r.final_residual = s.final_residual;
      r.final_forward_estimate = s.final_forward_estimate;
      if (meta.stagnation[i] && status != msq::kSolveNonfinite &&
          status != msq::kSolveInvalidSpectrum) {
        r.final_residual = s.best_residual;
        const uint32_t last = s.steps > 0 ? s.steps - 1 : 0;
        if (s.best_residual <= meta.residual_tolerance[i])
          status = msq::kSolveConverged;
        else if (status != msq::kSolveStagnated)
          status = last < msq::kSolveHistory && s.residuals[last] > s.residuals[0]
                       ? msq::kSolveDiverged
                       : msq::kSolveBudgetExhausted;
      }
      const bool unresolved = s.unresolved && !meta.refined[i];
      if (status == msq::kSolveConverged && unresolved) status = msq::kSolveUnresolved;
      r.status = status;
      r.matrix_norm = s.matrix_norm;
      r.min_root = s.min_root;
      r.off_norm = hj[i].off_norm;
      r.correction_rate = hj[i].correction_rate;
      r.basis_orthogonality = static_cast<float>(s.basis_orthogonality);
      r.basis_rejected = s.escalated;
      r.spectrum_rejected = s.spectrum_rejected;
      r.precision_escalations = s.escalated + s.spectrum_rejected;
      r.unresolved = unresolved ? 1u : 0u;
      r.refined = meta.refined[i];
      r.best_iteration = s.best_iteration;
      r.stagnation_steps = s.stagnation_steps;
      r.best_residual = s.best_residual;
      // Certificate (src/metal_solver.cpp:511-524): the accepted tier's basis
      // orthogonality, else the FP32 basis' when measured.
      const bool tier_basis = meta.low_basis[i] && !s.escalated;
      const double orth = tier_basis ? s.basis_orthogonality : s.fp32_orthogonality;
      r.frozen_basis_orthogonality = orth;
      r.frozen_sylvester_inverse_norm =
          meta.native[i] ? NAN : 1.0 / (2.0 * s.frozen_min_root);
      r.certificate_usable = !meta.native[i] && std::isfinite(orth) && orth < 1.0e-3 ? 1u : 0u;
      for (uint32_t k = 0; k < MSQ_SOLVE_HISTORY && k < msq::kSolveHistory; ++k) {
        r.residuals[k] = k < s.steps ? s.residuals[k] : 0.0;
        r.forward_estimates[k] = k < s.steps ? s.forward_estimates[k] : 0.0;
      }
      for (uint32_t k = 0; k < MSQ_JACOBI_HISTORY; ++k)
        r.off_history[k] = k < hj[i].sweeps && k < msq::kJacobiHistory ? hj[i].off_history[k] : 0.0f;
    }
  }
  rt->last_vectors = V;
  rt->last_roots = sroot;
  return MSQ_OK;
}

}  // namespace

extern "C" {

const char* msq_last_error(void) { return g_last_error.c_str(); }
int msq_abi_version(void) { return MSQ_ABI_VERSION; }

int msq_runtime_create(int device, msq_runtime** out) {
  if (!out) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "out is null");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "device %d not in [0, %d)", device, n);
  HIP_TRY(hipSetDevice(device));
  msq_runtime* rt = new msq_runtime();
  rt->device = device;
  hipError_t e = hipStreamCreateWithFlags(&rt->own_stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete rt;
    return set_error(MSQ_ERROR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  rt->stream = rt->own_stream;
  for (int l = 0; l < kLanes && e == hipSuccess; ++l) {
    e = hipStreamCreateWithFlags(&rt->lanes[l], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&rt->joins[l], hipEventDisableTiming);
  }
  if (e == hipSuccess) e = hipEventCreateWithFlags(&rt->fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&rt->graph_in, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&rt->graph_out, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipDeviceGetAttribute(&rt->cu_count, hipDeviceAttributeMultiprocessorCount, device);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&rt->jacobi_counter), 256);
  // The allocation-free entry point's scratch address on this device (once
  // per device; the device is current here).
  if (e == hipSuccess) e = msq::noalloc_register(device);
  if (e != hipSuccess) {
    msq_runtime_destroy(rt);
    return set_error(MSQ_ERROR_HIP, "runtime streams: %s", hipGetErrorString(e));
  }
  *out = rt;
  return MSQ_OK;
}

int msq_runtime_destroy(msq_runtime* rt) {
  if (!rt) return MSQ_OK;
  // Teardown is best effort: report the first failure but release everything.
  int status = MSQ_OK;
  hipError_t e = hipSetDevice(rt->device);
  if (e == hipSuccess) e = hipStreamSynchronize(rt->stream);
  if (e != hipSuccess) status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  if (rt->own_stream) (void)hipStreamSynchronize(rt->own_stream);
  for (auto& g : rt->graphs) release_graph(rt, *g);
  rt->graphs.clear();
  if (rt->arena.base && (e = hipFree(rt->arena.base)) != hipSuccess && status == MSQ_OK)
    status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  if (rt->large.base && (e = hipFree(rt->large.base)) != hipSuccess && status == MSQ_OK)
    status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  if (rt->refine.base && (e = hipFree(rt->refine.base)) != hipSuccess && status == MSQ_OK)
    status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  if (rt->jacobi_counter && (e = hipFree(rt->jacobi_counter)) != hipSuccess && status == MSQ_OK)
    status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  for (int l = 0; l < kLanes; ++l) {
    if (rt->lanes[l] && (e = hipStreamSynchronize(rt->lanes[l])) == hipSuccess)
      e = hipStreamDestroy(rt->lanes[l]);
    if (e != hipSuccess && status == MSQ_OK)
      status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
    if (rt->joins[l] && (e = hipEventDestroy(rt->joins[l])) != hipSuccess && status == MSQ_OK)
      status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  }
  if (rt->fork && (e = hipEventDestroy(rt->fork)) != hipSuccess && status == MSQ_OK)
    status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  for (hipEvent_t ev : {rt->graph_in, rt->graph_out})
    if (ev && (e = hipEventDestroy(ev)) != hipSuccess && status == MSQ_OK)
      status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  if (rt->own_stream && (e = hipStreamDestroy(rt->own_stream)) != hipSuccess && status == MSQ_OK)
    status = set_error(MSQ_ERROR_HIP, "destroy: %s", hipGetErrorString(e));
  delete rt;
  return status;
}

int msq_runtime_set_stream(msq_runtime* rt, void* stream) {
  MSQ_TRY(activate(rt));
  rt->stream = stream == MSQ_OWN_STREAM ? rt->own_stream : static_cast<hipStream_t>(stream);
  return MSQ_OK;
}

int msq_runtime_set_graphs(msq_runtime* rt, int enable) {
  MSQ_TRY(activate(rt));
  if (!enable && !rt->graphs.empty()) {
    HIP_TRY(hipStreamSynchronize(rt->stream));
    HIP_TRY(hipStreamSynchronize(rt->own_stream));
    for (auto& g : rt->graphs) release_graph(rt, *g);
    rt->graphs.clear();
  }
  rt->graphs_enabled = enable != 0;
  return MSQ_OK;
}

int msq_runtime_set_rounding(msq_runtime* rt, int mode, uint32_t seed) {
  MSQ_TRY(activate(rt));
  if (mode != MSQ_ROUNDING_NEAREST && mode != MSQ_ROUNDING_STOCHASTIC)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "rounding mode %d", mode);
  rt->rounding_mode = mode;
  rt->rounding_seed = seed;
  return MSQ_OK;
}

int msq_runtime_set_jacobi_method(msq_runtime* rt, int method) {
  MSQ_TRY(activate(rt));
  if (method != MSQ_JACOBI_SCALAR && method != MSQ_JACOBI_BLOCKED_MFMA)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "Jacobi method %d", method);
  if (method != rt->jacobi_method && !rt->graphs.empty()) {
    HIP_TRY(hipStreamSynchronize(rt->stream));
    HIP_TRY(hipStreamSynchronize(rt->own_stream));
    for (auto& g : rt->graphs) release_graph(rt, *g);
    rt->graphs.clear();
  }
  rt->jacobi_method = method;
  return MSQ_OK;
}

int msq_synchronize(msq_runtime* rt) {
  MSQ_TRY(activate(rt));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_alloc(msq_runtime* rt, size_t bytes, void** ptr) {
  MSQ_TRY(activate(rt));
  if (!ptr) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "ptr is null");
  hipError_t e = hipMalloc(ptr, bytes ? bytes : 1);
  if (e != hipSuccess)
    return set_error(MSQ_ERROR_OUT_OF_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
  return MSQ_OK;
}

int msq_free(msq_runtime* rt, void* ptr) {
  MSQ_TRY(activate(rt));
  if (ptr) HIP_TRY(hipFree(ptr));
  return MSQ_OK;
}

int msq_copy_to_device(msq_runtime* rt, void* dst, const void* src, size_t bytes) {
  MSQ_TRY(activate(rt));
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_copy_to_host(msq_runtime* rt, void* dst, const void* src, size_t bytes) {
  MSQ_TRY(activate(rt));
  HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_memset(msq_runtime* rt, void* dst, int value, size_t bytes) {
  MSQ_TRY(activate(rt));
  HIP_TRY(hipMemsetAsync(dst, value, bytes, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

size_t msq_storage_bytes(int precision, size_t elements) {
  size_t bits = 32;
  switch (precision) {
    case MSQ_FP32: bits = 32; break;
    case MSQ_BF16: bits = 16; break;
    case MSQ_FP4_E2M1:
    case MSQ_MXFP4_E2M1:
    case MSQ_FP4_E2M1_LOG_S: bits = 4; break;
    default: bits = 8; break;
  }
  return (elements * bits + 7) / 8;
}

size_t msq_scale_bytes(int precision, size_t elements) {
  return is_mx(precision) ? (elements + 31) / 32 : 0;
}

size_t msq_spectrum_metadata_bytes(int precision, size_t elements) {
  if (is_log_spectrum(precision)) return 2 * sizeof(float) * ((elements + 7) / 8);
  return msq_scale_bytes(precision, elements);
}

int msq_batch_layout(uint32_t count, const uint32_t* orders, uint32_t* lds, uint64_t* offsets,
                     uint64_t* vector_offsets, uint64_t* total, uint64_t* vtotal) {
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  for (uint32_t i = 0; i < count; ++i) {
    if (lds) lds[i] = L.ld[i];
    if (offsets) offsets[i] = L.off[i];
    if (vector_offsets) vector_offsets[i] = L.voff[i];
  }
  if (total) *total = L.total;
  if (vtotal) *vtotal = L.vtotal;
  return MSQ_OK;
}

int msq_quantize_stream(msq_runtime* rt, int precision, const float* src, void* values,
                        void* scales, size_t elements, uint32_t rounding_stream) {
  MSQ_TRY(activate(rt));
  const bool stochastic = rt->rounding_mode == MSQ_ROUNDING_STOCHASTIC;
  if (elements && (!src || !values)) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (reinterpret_cast<uintptr_t>(src) % 16 || reinterpret_cast<uintptr_t>(values) % 16)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "buffers must be 16-byte aligned");
  if (is_log_spectrum(precision)) {
    if (!scales) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "log-spectrum storage requires metadata");
    if (elements > UINT32_MAX)
      return set_error(MSQ_ERROR_INVALID_ARGUMENT, "more than 2^32 spectrum elements");
    HIP_TRY(msq::launch_encode_log(precision == MSQ_FP8_E4M3FN_LOG_S, src,
                                   static_cast<uint8_t*>(values), static_cast<float*>(scales),
                                   elements, stochastic, rt->rounding_seed, rounding_stream,
                                   rt->stream));
    HIP_TRY(hipStreamSynchronize(rt->stream));
    return MSQ_OK;
  }
  bool ok = false;
  const msq::Format f = format_of(precision, &ok);
  if (!ok) return set_error(MSQ_ERROR_UNSUPPORTED, "precision %d has no linear codec", precision);
  if (is_mx(precision) && !scales)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "MX storage requires a scale plane");
  if (stochastic) {
    // The hash takes 32-bit element indices, as the reference's kernels do.
    if (elements > UINT32_MAX)
      return set_error(MSQ_ERROR_INVALID_ARGUMENT, "more than 2^32 elements to round stochastically");
    HIP_TRY(msq::launch_encode_stochastic(f, src, values, static_cast<uint8_t*>(scales), elements,
                                          rt->rounding_seed, rounding_stream, rt->stream));
  } else {
    HIP_TRY(msq::launch_encode(f, src, values, static_cast<uint8_t*>(scales), elements,
                               rt->stream));
  }
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_quantize(msq_runtime* rt, int precision, const float* src, void* values, void* scales,
                 size_t elements) {
  return msq_quantize_stream(rt, precision, src, values, scales, elements, 0u);
}

int msq_decode(msq_runtime* rt, int precision, const void* values, const void* scales,
               float* dst, size_t elements) {
  MSQ_TRY(activate(rt));
  if (is_log_spectrum(precision)) {
    if (!scales) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "log-spectrum storage requires metadata");
    if (elements && (!dst || !values)) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
    HIP_TRY(msq::launch_decode_log(precision == MSQ_FP8_E4M3FN_LOG_S,
                                   static_cast<const uint8_t*>(values),
                                   static_cast<const float*>(scales), dst, elements, rt->stream));
    HIP_TRY(hipStreamSynchronize(rt->stream));
    return MSQ_OK;
  }
  bool ok = false;
  const msq::Format f = format_of(precision, &ok);
  if (!ok) return set_error(MSQ_ERROR_UNSUPPORTED, "precision %d has no linear codec", precision);
  if (is_mx(precision) && !scales)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "MX storage requires a scale plane");
  if (elements && (!dst || !values)) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (reinterpret_cast<uintptr_t>(dst) % 16 || reinterpret_cast<uintptr_t>(values) % 16)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "buffers must be 16-byte aligned");
  HIP_TRY(msq::launch_decode(f, values, static_cast<const uint8_t*>(scales), dst, elements,
                             rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_quantize_and_decode(msq_runtime* rt, int precision, const float* src, void* values,
                            void* scales, float* decoded, size_t elements) {
  MSQ_TRY(msq_quantize(rt, precision, src, values, scales, elements));
  return msq_decode(rt, precision, values, scales, decoded, elements);
}

int msq_quantize_fp8_scaled(msq_runtime* rt, const float* src, void* values, float* sigma,
                            size_t elements) {
  MSQ_TRY(activate(rt));
  if (!sigma || (elements && (!src || !values)))
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  uint32_t* amax = nullptr;
  Scratch S{rt};
  S.want(amax, 1);
  MSQ_TRY(S.commit());
  HIP_TRY(msq::launch_fp8_amax_sigma(src, elements, amax, sigma, rt->stream));
  HIP_TRY(msq::launch_encode_fp8_scaled(src, static_cast<uint8_t*>(values), sigma, elements,
                                        rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_decode_fp8_scaled(msq_runtime* rt, const void* values, const float* sigma, float* dst,
                          size_t elements) {
  MSQ_TRY(activate(rt));
  if (!sigma || (elements && (!dst || !values)))
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  HIP_TRY(msq::launch_decode_fp8_scaled(static_cast<const uint8_t*>(values), sigma, dst, elements,
                                        rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_jacobi_batched(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                       const double* matrices, float* vectors, float* diagonal, float* roots,
                       const msq_jacobi_options* options, uint32_t options_count,
                       msq_jacobi_report* reports) {
  MSQ_TRY(activate(rt));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!matrices || !vectors || !diagonal || !roots || !options)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (options_count != 1 && options_count != count)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "options_count must be 1 or count");
  Scratch S{rt};
  msq::JacobiStatus* jst = nullptr;
  msq::JacobiJob* djobs = nullptr;
  S.want(jst, count);
  S.want(djobs, count);
  MSQ_TRY(S.commit());
  MSQ_TRY(run_jacobi(rt, L, matrices, vectors, diagonal, roots, options, options_count, jst, djobs));
  std::vector<msq::JacobiStatus> hs(count);
  HIP_TRY(hipMemcpyAsync(hs.data(), jst, count * sizeof(msq::JacobiStatus), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  if (reports)
    for (uint32_t i = 0; i < count; ++i) fill_jacobi_report(hs[i], reports[i]);
  return MSQ_OK;
}

// msq_jacobi_batched without allocation: plan, status and reports in the
// per-device static scratch (noalloc.hip); no hipMalloc, hipFree or host heap
// use.  The plan kernel lays the jobs out as plan_jacobi does (classes of
// padded order, largest first, matrices in index order within a class), so
// the launches are the same as the allocating entry point's.
int msq_jacobi_batched_noalloc(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                               const double* matrices, float* vectors, float* diagonal,
                               float* roots, const msq_jacobi_options* options,
                               uint32_t options_count, msq_jacobi_report* reports) {
  MSQ_TRY(activate(rt));
  if (count == 0 || count > msq::kNoallocMaxMatrices)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT,
                     "count %u outside [1, %u] (the allocation-free scratch holds %u matrices)",
                     count, msq::kNoallocMaxMatrices, msq::kNoallocMaxMatrices);
  if (!orders || !matrices || !vectors || !diagonal || !roots || !options)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (options_count != 1 && options_count != count)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "options_count must be 1 or count");
  uint32_t class_count[msq::kNoallocClasses] = {};
  double class_cost[msq::kNoallocClasses] = {};
  for (uint32_t i = 0; i < count; ++i) {
    const uint32_t o = orders[i];
    if (o < 1 || o > msq::kBatchedJacobiMaxOrder)
      return set_error(MSQ_ERROR_INVALID_ARGUMENT,
                       "order %u of matrix %u is outside [1, %u] (the allocation-free entry "
                       "point runs the LDS-resident Jacobi only)",
                       o, i, msq::kBatchedJacobiMaxOrder);
    const int c = msq::noalloc_class(o);
    ++class_count[c];
    class_cost[c] += static_cast<double>(o) * o * o;
  }
  const msq::NoallocScratch s = msq::noalloc_scratch(rt->device);
  if (!s.jobs)
    return set_error(MSQ_ERROR_HIP, "device %d has no allocation-free scratch", rt->device);
  // Classes largest first, as plan_jacobi lays them out.
  msq::NoallocClasses starts{};
  JacobiClass cls[kMaxJacobiClasses];
  int classes = 0;
  uint32_t pos = 0;
  for (int c = msq::kNoallocClasses - 1; c >= 0; --c) {
    starts.start[c] = pos;
    if (class_count[c]) {
      cls[classes].first = pos;
      cls[classes].count = class_count[c];
      cls[classes].order = static_cast<uint32_t>(16 * (c + 1));
      cls[classes].cost = class_cost[c];
      ++classes;
    }
    pos += class_count[c];
  }
  HIP_TRY(hipMemcpyAsync(s.orders, orders, count * sizeof(uint32_t), hipMemcpyHostToDevice,
                         rt->stream));
  HIP_TRY(hipMemcpyAsync(s.options, options, options_count * sizeof(msq_jacobi_options),
                         hipMemcpyHostToDevice, rt->stream));
  HIP_TRY(msq::launch_noalloc_plan(s, options_count, count, starts, rt->stream));
  MSQ_TRY(enqueue_jacobi_classes(rt, cls, classes, s.jobs, matrices, vectors, diagonal, roots,
                                 s.status));
  if (reports) {
    HIP_TRY(msq::launch_noalloc_reports(s, count, rt->stream));
    HIP_TRY(hipMemcpyAsync(reports, s.reports, count * sizeof(msq_jacobi_report),
                           hipMemcpyDeviceToHost, rt->stream));
  }
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_reconstruct(msq_runtime* rt, int precision, uint32_t count, const uint32_t* orders,
                    const float* vectors, const float* roots, float* root) {
  MSQ_TRY(activate(rt));
  MSQ_TRY(validate_tier(precision, "reconstruction"));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!vectors || !roots || !root) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  const std::vector<uint32_t> members = iota_members(count);
  const std::vector<msq::Segment> segs = segments_of(L, members);
  GemmPlan plan;
  build_plan(L, members, false, plan, rt, gemm_type(precision));
  Scratch S{rt};
  msq::Segment* dsegs = nullptr;
  CorrectionBuffers cb;
  S.want(dsegs, count);
  want_correction(S, cb, L, precision);
  plan_want(S, plan);
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dsegs, segs));
  MSQ_TRY(plan_upload(rt, plan));
  MSQ_TRY(reconstruction_product(rt, precision, L, members, dsegs, plan, vectors, roots, nullptr,
                                 cb));
  HIP_TRY(hipMemcpyAsync(root, cb.t2, L.total * sizeof(float), hipMemcpyDeviceToDevice, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_residual_fp64(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                      const double* matrices, const double* root, double* residual,
                      float* residual_f32, double* relative_norms) {
  MSQ_TRY(activate(rt));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!matrices || !root || !residual || !residual_f32)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  const std::vector<uint32_t> members = iota_members(count);
  const std::vector<msq::Segment> segs = segments_of(L, members);
  GemmPlan plan;
  build_plan(L, members, true, plan, rt, msq::GemmType::F64);
  Scratch S{rt};
  msq::Segment* dsegs = nullptr;
  double* norms = nullptr;
  double* anorms = nullptr;
  S.want(dsegs, count);
  S.want(norms, count);
  S.want(anorms, count);
  plan_want(S, plan);
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dsegs, segs));
  MSQ_TRY(plan_upload(rt, plan));
  MSQ_TRY(residual_product(rt, plan, matrices, root, residual, residual_f32, nullptr, norms, nullptr,
                           count, false));
  HIP_TRY(msq::launch_zero_f64(anorms, count, rt->stream));
  HIP_TRY(msq::launch_sumsq_f64(dsegs, count, L.max_ld, matrices, anorms, rt->stream));
  std::vector<double> hn(count), ha(count);
  HIP_TRY(hipMemcpyAsync(hn.data(), norms, count * sizeof(double), hipMemcpyDeviceToHost, rt->stream));
  HIP_TRY(hipMemcpyAsync(ha.data(), anorms, count * sizeof(double), hipMemcpyDeviceToHost, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  if (relative_norms)
    for (uint32_t i = 0; i < count; ++i) relative_norms[i] = std::sqrt(hn[i] / ha[i]);
  return MSQ_OK;
}

int msq_correction(msq_runtime* rt, int precision, uint32_t count, const uint32_t* orders,
                   const float* vectors, const float* roots, const float* residual_f32,
                   float* correction) {
  MSQ_TRY(activate(rt));
  MSQ_TRY(validate_tier(precision, "correction"));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!vectors || !roots || !residual_f32 || !correction)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  const std::vector<uint32_t> members = iota_members(count);
  const std::vector<msq::Segment> segs = segments_of(L, members);
  GemmPlan plan;
  build_plan(L, members, false, plan, rt, gemm_type(precision));
  Scratch S{rt};
  msq::Segment* dsegs = nullptr;
  float* vt = nullptr;
  CorrectionBuffers cb;
  TierBuffers ev, evt;
  S.want(dsegs, count);
  S.want(vt, L.total);
  want_correction(S, cb, L, precision);
  tier_want(S, ev, precision, L);
  tier_want(S, evt, precision, L);
  plan_want(S, plan);
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dsegs, segs));
  MSQ_TRY(plan_upload(rt, plan));
  HIP_TRY(msq::launch_transpose(dsegs, count, pair_count(segs), vectors, vt, rt->stream));
  Operand ov, ovt;
  MSQ_TRY(encode_operand(rt, precision, L, members, dsegs, vectors, ev, ov));
  MSQ_TRY(encode_operand(rt, precision, L, members, dsegs, vt, evt, ovt));
  MSQ_TRY(correction_products(rt, precision, L, members, dsegs, plan, ov, ovt, residual_f32, nullptr,
                              roots, nullptr, cb));
  HIP_TRY(hipMemcpyAsync(correction, cb.t2, L.total * sizeof(float), hipMemcpyDeviceToDevice,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_refine_native_staged(msq_runtime* rt, int precision, uint32_t count,
                             const uint32_t* orders, const float* vectors, const float* roots,
                             const float* residual_f32, float* correction) {
  MSQ_TRY(activate(rt));
  if (!native_format(precision))
    return set_error(MSQ_ERROR_INVALID_ARGUMENT,
                     "native staged correction supports FP8 E4M3FN, FP8 E5M2, MXFP8 E4M3FN, MXFP8 "
                     "E5M2, FP4 E2M1, or MXFP4 E2M1 storage (precision %d)",
                     precision);
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!vectors || !roots || !residual_f32 || !correction)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  const std::vector<uint32_t> members = iota_members(count);
  const std::vector<msq::Segment> segs = segments_of(L, members);
  std::vector<uint64_t> coff;
  uint32_t max_order = 0;
  const uint64_t compact = compact_offsets(L, members, coff, max_order);
  GemmPlan plan;
  build_plan(L, members, false, plan, rt, msq::GemmType::F32);
  Scratch S{rt};
  msq::Segment* dsegs = nullptr;
  uint64_t* dcoff = nullptr;
  float *q = nullptr, *qT = nullptr, *qt = nullptr, *qtT = nullptr, *qs = nullptr, *t1 = nullptr,
        *t2 = nullptr;
  int* rejected = nullptr;
  uint8_t *codes = nullptr, *scales = nullptr;
  S.want(dsegs, count);
  S.want(dcoff, count);
  S.want(q, L.total);
  S.want(qT, L.total);
  S.want(qt, L.total);
  S.want(qtT, L.total);
  S.want(t1, L.total);
  S.want(t2, L.total);
  S.want(qs, L.vtotal);
  S.want(rejected, count);
  S.want(codes, compact);
  S.want(scales, compact / 32 + 1);
  plan_want(S, plan);
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dsegs, segs));
  MSQ_TRY(upload(rt, dcoff, coff));
  MSQ_TRY(plan_upload(rt, plan));
  NativeSet g;
  g.format = precision;
  g.count = count;
  g.max_order = max_order;
  g.dsegs = dsegs;
  g.dcoff = dcoff;
  g.plan = &plan;
  g.codes = codes;
  g.scales = scales;
  // Roots stored in the format (copySpectrumToStorage); the reference fails on
  // a stored root that is not positive and finite.
  HIP_TRY(msq::launch_segment_spectrum(dsegs, count, precision, roots, qs, rejected, nullptr,
                                       false, 0u, 1u, rt->stream));
  std::vector<int> hrej(count);
  HIP_TRY(hipMemcpyAsync(hrej.data(), rejected, count * sizeof(int), hipMemcpyDeviceToHost,
                         rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  for (uint32_t i = 0; i < count; ++i)
    if (hrej[i])
      return set_error(MSQ_ERROR_INVALID_ARGUMENT,
                       "the stored roots of matrix %u contain a value that is not positive and "
                       "finite",
                       i);
  MSQ_TRY(native_basis(rt, L, g, vectors, q, qT, qt, qtT, nullptr));
  MSQ_TRY(native_correction(rt, L, g, residual_f32, q, qT, qt, qtT, qs, t1, t2, nullptr));
  HIP_TRY(msq::launch_copy_words(t2, correction, L.total, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  return MSQ_OK;
}

int msq_orthogonality(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                      const float* vectors, double* results) {
  MSQ_TRY(activate(rt));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!vectors || !results) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  const std::vector<uint32_t> members = iota_members(count);
  const std::vector<msq::Segment> segs = segments_of(L, members);
  GemmPlan plan;
  build_plan(L, members, false, plan, rt, msq::GemmType::F32);
  Scratch S{rt};
  msq::Segment* dsegs = nullptr;
  float* gram = nullptr;
  double* out = nullptr;
  S.want(dsegs, count);
  S.want(gram, L.total);
  S.want(out, count);
  plan_want(S, plan);
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dsegs, segs));
  MSQ_TRY(plan_upload(rt, plan));
  Operand ov;
  ov.data = vectors;
  msq::GemmArgs g = product_args(ov, ov, gram, nullptr, nullptr);
  MSQ_TRY(run_gemm(rt, plan, msq::GemmType::F32, msq::Epilogue::Store, g));
  HIP_TRY(msq::launch_zero_f64(out, count, rt->stream));
  HIP_TRY(msq::launch_orthogonality(dsegs, count, L.max_ld, gram, out, rt->stream));
  std::vector<double> h(count);
  HIP_TRY(hipMemcpyAsync(h.data(), out, count * sizeof(double), hipMemcpyDeviceToHost, rt->stream));
  HIP_TRY(hipStreamSynchronize(rt->stream));
  for (uint32_t i = 0; i < count; ++i) results[i] = std::sqrt(h[i] / L.order[i]);
  return MSQ_OK;
}

void msq_solve_options_init(msq_solve_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->basis_precision = MSQ_FP32;
  o->correction_precision = MSQ_FP32;
  o->jacobi.max_sweeps = 30;
  o->jacobi.off_tol = 1e-7f;
  o->jacobi.stagnation = 0.5f;
  o->jacobi.floor_tol = 1e-5f;
  o->min_corrections = 0;
  o->max_corrections = 12;
  o->residual_tolerance = 1e-12;
  o->forward_tolerance = 0.0;
  o->spectrum_precision = MSQ_FP32;
  o->kernel_path = MSQ_KERNEL_PATH_TILED;
  o->correction_mode = MSQ_CORRECTION_FP32_REORTHOGONALIZED;
  o->stopping_policy = MSQ_STOPPING_FIXED_BUDGET;
  o->dispatch_policy = MSQ_DISPATCH_FUSED;
  o->stagnation_window = 4;
  o->minimum_relative_improvement = 1e-3;
}

int msq_solve_batched(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                      const double* matrices, double* roots_out, const msq_solve_options* options,
                      uint32_t options_count, msq_solve_report* reports) {
  msq_solve_outputs out{};
  out.roots = roots_out;
  return msq_solve_batched_ex(rt, count, orders, matrices, &out, options, options_count, reports);
}

int msq_solve_batched_ex(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                         const double* matrices, const msq_solve_outputs* outputs,
                         const msq_solve_options* options, uint32_t options_count,
                         msq_solve_report* reports) {
  MSQ_TRY(activate(rt));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!matrices || !outputs || !outputs->roots || !options)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (options_count != 1 && options_count != count)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "options_count must be 1 or count");
  double* roots_out = outputs->roots;
  auto opt = [&](uint32_t i) -> const msq_solve_options& {
    return options[options_count == 1 ? 0 : i];
  };
  // Validation: the reference's checks of a configuration (runMetalCase,
  // src/metal_solver.cpp:359-370; validateCorrectionMode and the policy
  // validators, src/configuration.cpp:290-346) and this backend's limits.
  for (uint32_t k = 0; k < options_count; ++k) MSQ_TRY(validate_solve_options(options[k]));
  SolveMeta meta;
  meta.stagnation.assign(count, 0);
  meta.native.assign(count, 0);
  meta.low_basis.assign(count, 0);
  meta.refined.assign(count, 0);
  meta.residual_tolerance.assign(count, 0.0);
  bool refine_any = false, stagnation_any = false, certificate_any = false;
  for (uint32_t i = 0; i < count; ++i) {
    const msq_solve_options& o = opt(i);
    const bool native = o.correction_mode == MSQ_CORRECTION_NATIVE_QUANTIZED;
    meta.native[i] = native;
    meta.stagnation[i] = o.stopping_policy == MSQ_STOPPING_STAGNATION;
    meta.low_basis[i] = !native && o.basis_precision != MSQ_FP32;
    meta.residual_tolerance[i] = o.residual_tolerance;
    refine_any = refine_any || (o.refine_unresolved && o.resolution > 0.0f);
    stagnation_any = stagnation_any || meta.stagnation[i];
    certificate_any = certificate_any || (o.certificate && !native);
  }
  GraphEntry* graph = nullptr;
  if (rt->graphs_enabled) {
    graph = find_graph(rt, count, orders, matrices, *outputs, options, options_count);
    if (graph && graph->exec) {
      // Replay: the plans and uploads of this signature live in the entry's
      // arena; only the device work runs again.
      MSQ_TRY(launch_graph(rt, *graph));
      return finish_solve(rt, graph->layout, graph->st, graph->jst, graph->V, graph->sroot,
                          reports, meta);
    }
  }
  // Groups: matrices sharing (basis, correction) tiers of the reorthogonalized
  // path, and native-quantized matrices by storage format.
  std::map<std::pair<int, int>, std::vector<uint32_t>> groups;
  std::map<int, std::vector<uint32_t>> by_corr, by_spectrum, by_native;
  std::vector<uint32_t> cert_members, stag_members;
  uint32_t max_corrections = 0;
  std::vector<msq_jacobi_options> jopts(options_count);
  std::vector<msq::SolveLimits> limits(count);
  for (uint32_t i = 0; i < count; ++i) {
    const msq_solve_options& o = opt(i);
    if (meta.native[i]) {
      by_native[o.basis_precision].push_back(i);
    } else {
      groups[{o.basis_precision, o.correction_precision}].push_back(i);
      by_corr[o.correction_precision].push_back(i);
      if (o.spectrum_precision != MSQ_FP32) by_spectrum[o.spectrum_precision].push_back(i);
      if (o.certificate) cert_members.push_back(i);
    }
    if (meta.stagnation[i]) stag_members.push_back(i);
    max_corrections = std::max(max_corrections, o.max_corrections);
    msq::SolveLimits& lim = limits[i];
    lim = msq::SolveLimits{};
    lim.residual_tolerance = o.residual_tolerance;
    lim.forward_tolerance = o.forward_tolerance;
    lim.minimum_relative_improvement = o.minimum_relative_improvement;
    lim.basis_tolerance = o.basis_tolerance;
    lim.min_corrections = std::min(o.min_corrections, o.max_corrections);
    lim.max_corrections = o.max_corrections;
    lim.stagnation_window = meta.stagnation[i] ? o.stagnation_window : 0u;
    lim.resolution = o.resolution;
    lim.refine = o.refine_unresolved;
  }
  for (uint32_t k = 0; k < options_count; ++k) jopts[k] = options[k].jacobi;

  // Host plans.
  const std::vector<uint32_t> all = iota_members(count);
  const std::vector<msq::Segment> segs_all = segments_of(L, all);
  GemmPlan plan_res;
  build_plan(L, all, true, plan_res, rt, msq::GemmType::F64, true);
  struct Group {
    int basis, corr;
    std::vector<uint32_t> members;
    std::vector<msq::Segment> segs;
    GemmPlan plan;
    GemmPlan upper;  // tiles on and above the diagonal (symmetric products)
    msq::Segment* dsegs = nullptr;
  };
  std::vector<Group> gl;
  for (auto& kv : groups) {
    Group g;
    g.basis = kv.first.first;
    g.corr = kv.first.second;
    g.members = kv.second;
    g.segs = segments_of(L, g.members);
    build_plan(L, g.members, false, g.plan, rt, gemm_type(g.corr));
    build_plan(L, g.members, false, g.upper, rt, gemm_type(g.corr), true);
    gl.push_back(std::move(g));
  }
  struct CorrTier {
    int tier;
    std::vector<uint32_t> members;
    std::vector<msq::Segment> segs;
    msq::Segment* dsegs = nullptr;
    TierBuffers ev, evt;
    Operand ov, ovt;
  };
  std::vector<CorrTier> tiers;
  for (auto& kv : by_corr) {
    CorrTier t;
    t.tier = kv.first;
    t.members = kv.second;
    t.segs = segments_of(L, t.members);
    tiers.push_back(std::move(t));
  }
  struct SpectrumTier {
    int precision;
    std::vector<msq::Segment> segs;
    msq::Segment* dsegs = nullptr;
  };
  std::vector<SpectrumTier> stiers;
  for (auto& kv : by_spectrum) {
    SpectrumTier t;
    t.precision = kv.first;
    t.segs = segments_of(L, kv.second);
    stiers.push_back(std::move(t));
  }
  // Native groups: compact storage of every member (n^2 elements from a
  // 64-element boundary), F32 products on decoded operands.
  struct NativeGroup {
    int format;
    std::vector<uint32_t> members;
    std::vector<msq::Segment> segs;
    std::vector<uint64_t> coff;
    GemmPlan plan;
    msq::Segment* dsegs = nullptr;
    uint64_t* dcoff = nullptr;
    uint32_t max_order = 0;
    NativeSet set() const {
      NativeSet n;
      n.format = format;
      n.count = static_cast<uint32_t>(members.size());
      n.max_order = max_order;
      n.dsegs = dsegs;
      n.dcoff = dcoff;
      n.plan = &plan;
      return n;
    }
  };
  std::vector<NativeGroup> nl;
  uint64_t compact_total = 0;
  for (auto& kv : by_native) {
    NativeGroup g;
    g.format = kv.first;
    g.members = kv.second;
    g.segs = segments_of(L, g.members);
    compact_total = std::max(compact_total, compact_offsets(L, g.members, g.coff, g.max_order));
    build_plan(L, g.members, false, g.plan, rt, msq::GemmType::F32);
    nl.push_back(std::move(g));
  }
  GemmPlan plan_cert;
  std::vector<msq::Segment> segs_cert = segments_of(L, cert_members);
  if (!cert_members.empty())
    build_plan(L, cert_members, false, plan_cert, rt, msq::GemmType::F32, true);
  std::vector<msq::Segment> segs_stag = segments_of(L, stag_members);
  std::vector<int> stag_flags(count, 0);
  for (uint32_t i : stag_members) stag_flags[i] = 1;

  // Device buffers.
  Scratch S{rt};
  float *V = nullptr, *Vt = nullptr, *diag = nullptr, *sroot = nullptr, *qroot = nullptr,
        *R32 = nullptr;
  uint16_t* Rb = nullptr;
  double* norms = nullptr;
  double* orth2 = nullptr;  // per matrix: ||V_t^T V_t - I||_F^2 (basis acceptance)
  double* cert2 = nullptr;  // per matrix: ||V^T V - I||_F^2 of the FP32 basis (certificate)
  int* esc = nullptr;       // per matrix: basis tier rejected
  int* srej = nullptr;      // per matrix: spectrum tier rejected
  int* active = nullptr;
  int* save = nullptr;       // per matrix: this step's iterate is the best (stagnation policy)
  int* dstag = nullptr;      // per matrix: stagnation policy
  double* xbest = nullptr;   // best iterates (stagnation policy)
  msq::JacobiStatus* jst = nullptr;
  msq::JacobiJob* djobs = nullptr;
  msq::SolveState* st = nullptr;
  msq::SolveLimits* dlim = nullptr;
  msq::Segment* dsegs_all = nullptr;
  msq::Segment* dsegs_cert = nullptr;
  msq::Segment* dsegs_stag = nullptr;
  float *nq = nullptr, *nqT = nullptr, *nqt = nullptr, *nqtT = nullptr;  // native basis operands
  uint8_t *ncodes = nullptr, *nscales = nullptr;
  CorrectionBuffers cb;
  S.want(V, L.total);
  S.want(Vt, L.total);
  S.want(diag, L.vtotal);
  S.want(sroot, L.vtotal);
  S.want(qroot, L.vtotal);
  // R feeds the corrections as float (FP32 / FP8 / MXFP4 tiers, native) and
  // as BF16 codes (BF16 tier); the binary64 R is not kept.
  bool any_bf16 = false, any_float = !nl.empty();
  for (const Group& g : gl) (g.corr == MSQ_BF16 ? any_bf16 : any_float) = true;
  if (any_float) S.want(R32, L.total);
  if (any_bf16) S.want(Rb, L.total);
  S.want(norms, count);
  S.want(orth2, count);
  S.want(cert2, count);
  S.want(esc, count);
  S.want(srej, count);
  S.want(active, count);
  S.want(save, count);
  S.want(dstag, count);
  if (stagnation_any) S.want(xbest, L.total);
  S.want(jst, count);
  S.want(djobs, count);
  S.want(st, count);
  S.want(dlim, count);
  S.want(dsegs_all, count);
  S.want(dsegs_cert, std::max<size_t>(1, cert_members.size()));
  S.want(dsegs_stag, std::max<size_t>(1, stag_members.size()));
  S.want(cb.t1, L.total);
  S.want(cb.t2, L.total);
  S.want(cb.alpha, count);
  tier_want_all(S, cb.enc_a, L);
  tier_want_all(S, cb.enc_b, L);
  plan_want(S, plan_res);
  if (!cert_members.empty()) plan_want(S, plan_cert);
  for (Group& g : gl) {
    S.want(g.dsegs, g.members.size());
    plan_want(S, g.plan);
    plan_want(S, g.upper);
  }
  for (CorrTier& t : tiers) {
    S.want(t.dsegs, t.members.size());
    tier_want(S, t.ev, t.tier, L);
    tier_want(S, t.evt, t.tier, L);
  }
  for (SpectrumTier& t : stiers) S.want(t.dsegs, t.segs.size());
  if (!nl.empty()) {
    S.want(nq, L.total);
    S.want(nqT, L.total);
    S.want(nqt, L.total);
    S.want(nqtT, L.total);
    S.want(ncodes, compact_total);
    S.want(nscales, compact_total / 32 + 1);
    for (NativeGroup& g : nl) {
      S.want(g.dsegs, g.members.size());
      S.want(g.dcoff, g.members.size());
      plan_want(S, g.plan);
    }
  }
  // Graph mode: the call's buffers come from the signature's own arena, so
  // that a captured launch sequence keeps valid pointers.
  if (graph) std::swap(rt->arena, graph->arena);
  struct ArenaRestore {
    msq_runtime* rt;
    GraphEntry* g;
    ~ArenaRestore() {
      if (g) std::swap(rt->arena, g->arena);
    }
  } restore{rt, graph};
  MSQ_TRY(S.commit());
  JacobiPlan jplan;
  plan_jacobi(L, jopts.data(), options_count, jplan);
  MSQ_TRY(upload(rt, djobs, jplan.jobs));
  MSQ_TRY(upload(rt, dsegs_all, segs_all));
  MSQ_TRY(upload(rt, dlim, limits));
  MSQ_TRY(upload(rt, dstag, stag_flags));
  MSQ_TRY(upload(rt, dsegs_cert, segs_cert));
  MSQ_TRY(upload(rt, dsegs_stag, segs_stag));
  MSQ_TRY(plan_upload(rt, plan_res));
  if (!cert_members.empty()) MSQ_TRY(plan_upload(rt, plan_cert));
  for (Group& g : gl) {
    MSQ_TRY(upload(rt, g.dsegs, g.segs));
    MSQ_TRY(plan_upload(rt, g.plan));
    MSQ_TRY(plan_upload(rt, g.upper));
  }
  for (CorrTier& t : tiers) MSQ_TRY(upload(rt, t.dsegs, t.segs));
  for (SpectrumTier& t : stiers) MSQ_TRY(upload(rt, t.dsegs, t.segs));
  for (NativeGroup& g : nl) {
    MSQ_TRY(upload(rt, g.dsegs, g.segs));
    MSQ_TRY(upload(rt, g.dcoff, g.coff));
    MSQ_TRY(plan_upload(rt, g.plan));
  }
  const bool stochastic = rt->rounding_mode == MSQ_ROUNDING_STOCHASTIC;

  auto native_set = [&](const NativeGroup& g) {
    NativeSet n = g.set();
    n.codes = ncodes;
    n.scales = nscales;
    return n;
  };

  // Device work: Jacobi, reconstruction, frozen factor, residual and
  // corrections (no host synchronisation below, except for orders above 256
  // and the second Jacobi pass of unresolved matrices).
  // 5. Residual of X0, then corrections until each matrix stops.
  auto residual_step = [&](uint32_t step) -> int {
    MSQ_TRY(residual_product(rt, plan_res, matrices, roots_out, nullptr, R32, Rb, norms, active,
                             count));
    HIP_TRY(msq::launch_convergence(st, norms, active, stagnation_any ? save : nullptr, count,
                                    step, rt->stream));
    if (stagnation_any)
      HIP_TRY(msq::launch_copy_tiles_f64(dsegs_stag, static_cast<uint32_t>(stag_members.size()),
                                         L.max_ld, roots_out, xbest, save, rt->stream));
    return MSQ_OK;
  };
  // Correction steps first..last (each ends with the residual of the new
  // iterate).
  auto correction_steps = [&](uint32_t first, uint32_t last) -> int {
    for (uint32_t step = first; step <= last; ++step) {
      for (Group& g : gl) {
        const CorrTier* t = nullptr;
        for (const CorrTier& c : tiers)
          if (c.tier == g.corr) t = &c;
        MSQ_TRY(correction_products(rt, g.corr, L, g.members, g.dsegs, g.plan, t->ov, t->ovt, R32,
                                    Rb, sroot, active, cb, roots_out, &g.upper));
      }
      // Native: U = Q_t R, C = U Q, Y = C / (s_i + s_j), V2 = Q Y, E = V2 Q_t,
      // each rounded to the format; X = sym(X + E).
      for (NativeGroup& g : nl) {
        const uint32_t gm = static_cast<uint32_t>(g.members.size());
        MSQ_TRY(native_correction(rt, L, native_set(g), R32, nq, nqT, nqt, nqtT, qroot, cb.t1,
                                  cb.t2, active));
        HIP_TRY(msq::launch_add_symmetrize_f64(g.dsegs, gm, L.max_ld, roots_out, cb.t2, active,
                                               rt->stream));
      }
      MSQ_TRY(residual_step(step));
    }
    return MSQ_OK;
  };
  auto enqueue = [&](uint32_t last) -> int {
    // 1. Jacobi (symmetricJacobi).
    MSQ_TRY(enqueue_jacobi(rt, jplan, djobs, matrices, V, diag, sroot, jst));
    if (jplan.has_large)
      MSQ_TRY(run_large_jacobi(rt, L, matrices, V, diag, sroot, jopts.data(), options_count, jst));
    HIP_TRY(msq::launch_solve_state(dsegs_all, jst, diag, dlim, st, active, count, refine_any,
                                    rt->stream));
    // 1b. Matrices whose smallest eigenvalue the FP32 Jacobi does not
    //     resolve: a second pass on V^T A V formed in binary64.
    if (refine_any) MSQ_TRY(refine_unresolved(rt, L, limits, matrices, V, diag, sroot, jst, st,
                                              dsegs_all, dlim, active, meta));
    // 2. Spectrum tier: the roots rounded to the spectrum precision, accepted
    //    per matrix when every decoded root is positive and finite
    //    (src/metal_solver.cpp:240-263); the others keep the FP32 roots.
    HIP_TRY(msq::launch_copy_words(sroot, qroot, L.vtotal, rt->stream));
    HIP_TRY(msq::launch_fill_u32(reinterpret_cast<uint32_t*>(srej), count, 0u, rt->stream));
    for (SpectrumTier& t : stiers)
      HIP_TRY(msq::launch_segment_spectrum(t.dsegs, static_cast<uint32_t>(t.segs.size()),
                                           t.precision, sroot, qroot, srej, active, stochastic,
                                           rt->rounding_seed, 1u, rt->stream));
    // 3. Reconstruction X0 = G G^T in the basis tier (with the frozen
    //    roots), X = sym(X0) in binary64.  A low-bit basis is accepted per
    //    matrix as the reference's quantizeAndDecode + measureF32Orthogonality
    //    do (src/metal_solver.cpp:217-236): ||V_t^T V_t - I||_F / sqrt(n) of V
    //    rounded to the tier within the tolerance (src/configuration.cpp:
    //    109-118, or the options' override); the matrices that miss it get X0
    //    from the FP32 basis instead.
    HIP_TRY(msq::launch_fill_u32(reinterpret_cast<uint32_t*>(esc), count, 0u, rt->stream));
    for (Group& g : gl) {
      const uint32_t gm = static_cast<uint32_t>(g.members.size());
      if (g.basis != MSQ_FP32) {
        Operand ob;
        const float* alpha = nullptr;
        MSQ_TRY(encode_operand(rt, g.basis, L, g.members, g.dsegs, V, cb.enc_a, ob));
        MSQ_TRY(fp8_alpha(rt, g.basis, ob, ob, cb.alpha, count, &alpha));
        // Gram product V_t^T V_t reduced to ||V_t^T V_t - I||_F^2 in its
        // epilogue (symmetric: the upper tiles, off-diagonal ones twice).
        HIP_TRY(msq::launch_zero_f64(orth2, count, rt->stream));
        msq::GemmArgs gg = product_args(ob, ob, nullptr, alpha, active);
        gg.norms = orth2;
        gg.mirror = 1;
        MSQ_TRY(run_gemm(rt, g.upper, gemm_type(g.basis), msq::Epilogue::Gram, gg));
        HIP_TRY(msq::launch_basis_escalation(g.dsegs, gm, orth2, basis_tolerance(g.basis), dlim,
                                             active, esc, st, rt->stream));
      }
      MSQ_TRY(reconstruction_product(rt, g.basis, L, g.members, g.dsegs, g.plan, V, qroot, active,
                                     cb, /*two_term=*/true));
      if (g.basis != MSQ_FP32)  // the rejected ones again, from the FP32 basis
        MSQ_TRY(reconstruction_product(rt, MSQ_FP32, L, g.members, g.dsegs, g.plan, V, qroot,
                                       esc, cb));
      HIP_TRY(msq::launch_symmetrize_to_f64(g.dsegs, gm, pair_count(g.segs), cb.t2, roots_out,
                                            active, rt->stream));
      HIP_TRY(msq::launch_frozen_roots(g.dsegs, gm, qroot, srej, st, rt->stream));
    }
    // Certificate: ||V^T V - I||_F of the FP32 basis.
    if (!cert_members.empty()) {
      HIP_TRY(msq::launch_zero_f64(cert2, count, rt->stream));
      Operand ov;
      ov.data = V;
      msq::GemmArgs gg = product_args(ov, ov, nullptr, nullptr, active);
      gg.norms = cert2;
      gg.mirror = 1;
      MSQ_TRY(run_gemm(rt, plan_cert, msq::GemmType::F32, msq::Epilogue::Gram, gg));
      HIP_TRY(msq::launch_record_orthogonality(dsegs_cert,
                                               static_cast<uint32_t>(cert_members.size()), cert2,
                                               st, rt->stream));
    }
    // 3b. Native groups: basis, transposed basis and roots in the storage
    //     format (copyToStorage / copySpectrumToStorage); X0 = quant(quant(Q
    //     diag(s)) Q_t) (scaleColumns + matmul, src/metal_solver.cpp:
    //     767-787), the spectrum rejected as an invalid spectrum.
    for (NativeGroup& g : nl) {
      const uint32_t gm = static_cast<uint32_t>(g.members.size());
      const NativeSet ns = native_set(g);
      HIP_TRY(msq::launch_segment_spectrum(g.dsegs, gm, g.format, sroot, qroot, srej, active,
                                           false, 0u, 1u, rt->stream));
      HIP_TRY(msq::launch_reject_spectrum(g.dsegs, gm, srej, st, active, rt->stream));
      HIP_TRY(msq::launch_frozen_roots(g.dsegs, gm, qroot, srej, st, rt->stream));
      MSQ_TRY(native_basis(rt, L, ns, V, nq, nqT, nqt, nqtT, active));
      // Basis orthogonality of Q for the report.
      HIP_TRY(msq::launch_zero_f64(orth2, count, rt->stream));
      {
        Operand oq;
        oq.data = nq;
        msq::GemmArgs gg = product_args(oq, oq, nullptr, nullptr, active);
        gg.norms = orth2;
        MSQ_TRY(run_gemm(rt, g.plan, msq::GemmType::F32, msq::Epilogue::Gram, gg));
        // Recorded, never rejected (the native path has no basis acceptance).
        HIP_TRY(msq::launch_basis_escalation(g.dsegs, gm, orth2, INFINITY, nullptr, active, esc,
                                             st, rt->stream));
      }
      HIP_TRY(msq::launch_scale_columns(g.dsegs, gm, L.max_ld, nq, qroot, cb.t1, active,
                                        rt->stream));
      MSQ_TRY(native_round(rt, L, ns, cb.t1, false, cb.t2, true, active));
      MSQ_TRY(native_product(rt, ns, cb.t2, nqt, cb.t1, active));
      MSQ_TRY(native_round(rt, L, ns, cb.t1, false, cb.t2, false, active));
      HIP_TRY(msq::launch_symmetrize_to_f64(g.dsegs, gm, pair_count(g.segs), cb.t2, roots_out,
                                            active, rt->stream));
    }
    if (outputs->initial_roots)
      HIP_TRY(msq::launch_copy_words(roots_out, outputs->initial_roots, 2 * L.total, rt->stream));
    // The best iterate starts as X0 (stagnation policy).
    if (stagnation_any)
      HIP_TRY(msq::launch_copy_tiles_f64(dsegs_stag, static_cast<uint32_t>(stag_members.size()),
                                         L.max_ld, roots_out, xbest, nullptr, rt->stream));
    // 4. Frozen factor in each correction tier: V and V^T.  A BF16 tier
    //    takes both codes from one pass over V; the float V^T is formed
    //    only for the other tiers.
    bool float_vt = false;
    for (const CorrTier& t : tiers) float_vt = float_vt || t.tier != MSQ_BF16;
    if (float_vt)
      HIP_TRY(msq::launch_transpose(dsegs_all, count, pair_count(segs_all), V, Vt, rt->stream));
    for (CorrTier& t : tiers) {
      if (t.tier == MSQ_BF16) {
        HIP_TRY(msq::launch_transpose_encode_bf16(
            t.dsegs, static_cast<uint32_t>(t.members.size()), pair_count(t.segs), V,
            static_cast<uint16_t*>(t.ev.codes), static_cast<uint16_t*>(t.evt.codes), rt->stream));
        t.ov.data = t.ev.codes;
        t.ovt.data = t.evt.codes;
        continue;
      }
      MSQ_TRY(encode_operand(rt, t.tier, L, t.members, t.dsegs, V, t.ev, t.ov));
      MSQ_TRY(encode_operand(rt, t.tier, L, t.members, t.dsegs, Vt, t.evt, t.ovt));
    }
    MSQ_TRY(residual_step(0));
    MSQ_TRY(correction_steps(1, last));
    // The stagnation policy returns the best iterate.
    if (stagnation_any)
      HIP_TRY(msq::launch_copy_tiles_f64(dsegs_stag, static_cast<uint32_t>(stag_members.size()),
                                         L.max_ld, xbest, roots_out, dstag, rt->stream));
    if (outputs->vectors) HIP_TRY(msq::launch_copy_words(V, outputs->vectors, L.total, rt->stream));
    if (outputs->root_eigenvalues)
      HIP_TRY(msq::launch_copy_words(sroot, outputs->root_eigenvalues, L.vtotal, rt->stream));
    return MSQ_OK;
  };
  if (graph && !jplan.has_large && !refine_any && graph->calls >= 1) {
    // The steps after the last one that the eager call needed go to a
    // second graph (not under the stagnation policy, whose best iterate is
    // copied after the last step).
    const uint32_t split =
        stagnation_any ? max_corrections : std::min(graph->split, max_corrections);
    hipGraph_t hg = nullptr, tg = nullptr;
    hipGraphExec_t he = nullptr, te = nullptr;
    MSQ_TRY(capture_graph(rt, hg, he, [&] { return enqueue(split); }));
    if (split < max_corrections) {
      const int st2 = capture_graph(
          rt, tg, te, [&] { return correction_steps(split + 1, max_corrections); });
      if (st2 != MSQ_OK) {
        (void)hipGraphExecDestroy(he);
        (void)hipGraphDestroy(hg);
        return st2;
      }
    }
    graph->graph = hg;
    graph->exec = he;
    graph->tail_graph = tg;
    graph->tail_exec = te;
    graph->layout = L;
    graph->st = st;
    graph->jst = jst;
    graph->V = V;
    graph->sroot = sroot;
    graph->active = active;
    MSQ_TRY(launch_graph(rt, *graph));
  } else {
    MSQ_TRY(enqueue(max_corrections));
    if (graph) ++graph->calls;
  }
  uint32_t steps = 0;
  const int status = finish_solve(rt, L, st, jst, V, sroot, reports, meta, &steps);
  if (status == MSQ_OK && graph && !graph->exec) graph->split = steps > 0 ? steps - 1 : 0;
  return status;
}

int msq_last_factors(msq_runtime* rt, const float** vectors, const float** roots) {
  MSQ_TRY(activate(rt));
  if (!rt->last_vectors) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "no solve has run");
  if (vectors) *vectors = rt->last_vectors;
  if (roots) *roots = rt->last_roots;
  return MSQ_OK;
}


// ------------------------------------------------------------ tier selector
static_assert(msq::kMlpParams == MSQ_MLP_PARAMS && msq::kMlpInputs == MSQ_MLP_INPUTS &&
                  msq::kMlpHidden == MSQ_MLP_HIDDEN && msq::kMlpOutputs == MSQ_MLP_OUTPUTS &&
                  msq::kMaxFeatureOrder == MSQ_FEATURE_MAX_ORDER,
              "selector constants of the C ABI");

namespace {

// Features of the listed members of a batch (rows of `features` indexed by
// batch position; other rows untouched).
int matrix_features_of(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                       const double* matrices, uint32_t member_count, const uint32_t* member_list,
                       float* features) {
  MSQ_TRY(activate(rt));
  Layout L;
  MSQ_TRY(make_layout(count, orders, L));
  if (!matrices || !features) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  for (uint32_t q = 0; q < member_count; ++q) {
    const uint32_t i = member_list ? member_list[q] : q;
    if (i >= count) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "member %u of %u", i, count);
    if (L.order[i] > static_cast<uint32_t>(msq::kMaxFeatureOrder))
      return set_error(MSQ_ERROR_INVALID_ARGUMENT, "order %u above %d", L.order[i],
                       msq::kMaxFeatureOrder);
  }
  if (member_count == 0) return MSQ_OK;
  // Segments grouped by order range, one launch per range: the ranges run
  // concurrently on the runtime's lanes (the fifth on the first lane).
  constexpr int kKinds = 5;
  const int kinds[kKinds] = {256, 192, 128, 64, 0};
  std::vector<uint32_t> members;
  uint32_t bounds[kKinds + 1] = {};
  for (int k = 0; k < kKinds; ++k) {
    for (uint32_t q = 0; q < member_count; ++q) {
      const uint32_t i = member_list ? member_list[q] : q;
      if (msq::feature_rows_for_order(L.order[i]) == kinds[k]) members.push_back(i);
    }
    bounds[k + 1] = static_cast<uint32_t>(members.size());
  }
  const std::vector<msq::Segment> segs = segments_of(L, members);
  Scratch S{rt};
  msq::Segment* dsegs = nullptr;
  S.want(dsegs, members.size());
  MSQ_TRY(S.commit());
  MSQ_TRY(upload(rt, dsegs, segs));
  HIP_TRY(hipEventRecord(rt->fork, rt->stream));
  for (int l = 0; l < kLanes; ++l) HIP_TRY(hipStreamWaitEvent(rt->lanes[l], rt->fork, 0));
  for (int k = 0; k < kKinds; ++k)
    HIP_TRY(msq::launch_features(kinds[k], dsegs + bounds[k], bounds[k + 1] - bounds[k], matrices,
                                 features, rt->lanes[k % kLanes]));
  for (int l = 0; l < kLanes; ++l) {
    HIP_TRY(hipEventRecord(rt->joins[l], rt->lanes[l]));
    HIP_TRY(hipStreamWaitEvent(rt->stream, rt->joins[l], 0));
  }
  return MSQ_OK;
}

}  // namespace

int msq_matrix_features(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                        const double* matrices, float* features) {
  return matrix_features_of(rt, count, orders, matrices, count, nullptr, features);
}

int msq_matrix_features_subset(msq_runtime* rt, uint32_t count, const uint32_t* orders,
                               const double* matrices, uint32_t member_count,
                               const uint32_t* members, float* features) {
  if (member_count && !members) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null member list");
  return matrix_features_of(rt, count, orders, matrices, member_count, members, features);
}

int msq_mlp_forward(msq_runtime* rt, const float* params, const float* x, uint32_t batch,
                    float* logits, float* h1, float* h2) {
  MSQ_TRY(activate(rt));
  if (!params || !x || !logits) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (batch > (1u << 24)) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "batch too large");
  HIP_TRY(msq::launch_mlp_forward(params, x, static_cast<int>(batch), logits, h1, h2,
                                  rt->stream));
  return MSQ_OK;
}

int msq_mlp_forward_bf16(msq_runtime* rt, const uint16_t* params, const float* x, uint32_t batch,
                         float* logits) {
  MSQ_TRY(activate(rt));
  if (!params || !x || !logits) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (batch > (1u << 24)) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "batch too large");
  HIP_TRY(msq::launch_mlp_forward_bf16(params, x, static_cast<int>(batch), logits, rt->stream));
  return MSQ_OK;
}

int msq_mlp_backward(msq_runtime* rt, const float* params, const float* x, const float* h1,
                     const float* h2, const float* logits, const float* targets,
                     const float* mask, uint32_t batch, float mask_total, float* grads,
                     float* losses) {
  MSQ_TRY(activate(rt));
  if (!params || !x || !h1 || !h2 || !logits || !targets || !mask || !grads)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (batch == 0 || batch > (1u << 24))
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "batch must be in [1, 2^24]");
  Scratch S{rt};
  float *dz = nullptr, *dh1 = nullptr, *dh2 = nullptr, *loss = nullptr;
  S.want(dz, static_cast<size_t>(batch) * msq::kMlpOutputs);
  S.want(dh1, static_cast<size_t>(batch) * msq::kMlpHidden);
  S.want(dh2, static_cast<size_t>(batch) * msq::kMlpHidden);
  if (!losses) S.want(loss, batch);
  MSQ_TRY(S.commit());
  HIP_TRY(msq::launch_mlp_backward(params, x, h1, h2, logits, targets, mask,
                                   static_cast<int>(batch), mask_total, dz, dh1, dh2,
                                   losses ? losses : loss, grads, rt->stream));
  return MSQ_OK;
}

int msq_mlp_adam(msq_runtime* rt, float* params, const float* grads, float* m1, float* m2,
                 uint32_t count, float lr, float beta1, float beta2, float eps, float decay,
                 uint32_t step) {
  MSQ_TRY(activate(rt));
  if (!params || !grads || !m1 || !m2) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "null buffer");
  if (step == 0) return set_error(MSQ_ERROR_INVALID_ARGUMENT, "step counts from 1");
  HIP_TRY(msq::launch_mlp_adam(params, grads, m1, m2, static_cast<int>(count), lr, beta1, beta2,
                               eps, decay, static_cast<int>(step), rt->stream));
  return MSQ_OK;
}

// ---------------------------------------------------------------- autotuner
int msq_autotune(msq_runtime* rt, uint32_t max_order, uint8_t* table) {
  MSQ_TRY(activate(rt));
  if (max_order == 0 || max_order > 65535)
    return set_error(MSQ_ERROR_INVALID_ARGUMENT, "max_order must be in [1, 65535]");
  const msq::GemmType types[kGemmTypes] = {msq::GemmType::F64, msq::GemmType::F32,
                                           msq::GemmType::Bf16, msq::GemmType::Fp8,
                                           msq::GemmType::Mxfp4};
  // Synthetic batches of about 2^25 elements per class (operand contents do
  // not change MFMA timing; zero codes are valid in every format).
  const size_t target = size_t{1} << 25;
  std::vector<uint32_t> lds;
  for (uint32_t ld = 32; ld <= 256 && ld <= ((max_order + 31) / 32) * 32; ld += 32) lds.push_back(ld);
  if (max_order > 256) lds.push_back(512);
  const size_t max_elems = target + 512 * 512;
  char* buf = nullptr;
  const size_t bytes = max_elems * (8 * 4 + 4) + max_elems / 32 * 2 + 4096 * 16 + 1024;
  HIP_TRY(hipMalloc(&buf, bytes));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int status = MSQ_OK;
  do {
    hipError_t e = hipMemsetAsync(buf, 0, bytes, rt->stream);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) { status = set_error(MSQ_ERROR_HIP, "autotune: %s", hipGetErrorString(e)); break; }
    char* p = buf;
    auto take = [&](size_t b) { char* q = p; p += (b + 255) / 256 * 256; return q; };
    void* a = take(max_elems * 8);
    void* bo = take(max_elems * 8);
    void* c = take(max_elems * 8);
    void* aux = take(max_elems * 8);
    float* cf = reinterpret_cast<float*>(take(max_elems * 4));
    uint8_t* sa = reinterpret_cast<uint8_t*>(take(max_elems / 32));
    uint8_t* sb = reinterpret_cast<uint8_t*>(take(max_elems / 32));
    double* norms = reinterpret_cast<double*>(take(4096 * sizeof(double)));
    float* alpha = reinterpret_cast<float*>(take(4096 * sizeof(float)));
    for (uint32_t ld : lds) {
      const uint32_t count = static_cast<uint32_t>(std::min<size_t>(4096, std::max<size_t>(1, target / (size_t{ld} * ld))));
      std::vector<uint32_t> orders(count, ld);
      Layout L;
      MSQ_TRY(make_layout(count, orders.data(), L));
      const std::vector<uint32_t> members = iota_members(count);
      for (int ti = 0; ti < kGemmTypes && status == MSQ_OK; ++ti) {
        const msq::GemmType type = types[ti];
        const msq::Epilogue epi = type == msq::GemmType::F64 ? msq::Epilogue::Residual
                                                             : msq::Epilogue::Store;
        float best = INFINITY;
        int best_edge = 0;
        for (int k = 0; k < 3 && status == MSQ_OK; ++k) {
          const uint8_t saved = rt->tiles[ti][tile_class(ld)];
          rt->tiles[ti][tile_class(ld)] = static_cast<uint8_t>(kTileEdges[k]);
          GemmPlan plan;
          build_plan(L, members, true, plan, rt, type, type == msq::GemmType::F64);
          rt->tiles[ti][tile_class(ld)] = saved;
          Scratch S{rt};
          plan_want(S, plan);
          status = S.commit();
          if (status == MSQ_OK) status = plan_upload(rt, plan);
          if (status != MSQ_OK) break;
          msq::GemmArgs g{};
          g.a = a;
          g.b = bo;
          g.c = c;
          g.aux = aux;
          g.c_float = cf;
          g.norms = norms;
          g.a_scales = sa;
          g.b_scales = sb;
          g.alpha = type == msq::GemmType::Fp8 ? alpha : nullptr;
          g.mirror = type == msq::GemmType::F64 ? 1 : 0;
          status = run_gemm(rt, plan, type, epi, g);  // warm
          float ms = INFINITY;
          for (int r = 0; r < 3 && status == MSQ_OK; ++r) {
            e = hipEventRecord(e0, rt->stream);
            if (e == hipSuccess) status = run_gemm(rt, plan, type, epi, g);
            if (e == hipSuccess) e = hipEventRecord(e1, rt->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float t = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
            if (e != hipSuccess) { status = set_error(MSQ_ERROR_HIP, "autotune: %s", hipGetErrorString(e)); break; }
            ms = std::min(ms, t);
          }
          if (ms < best) {
            best = ms;
            best_edge = kTileEdges[k];
          }
        }
        if (status == MSQ_OK) rt->tiles[ti][tile_class(ld)] = static_cast<uint8_t>(best_edge);
      }
      if (status != MSQ_OK) break;
    }
  } while (false);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(rt->stream);
  hipError_t fe = hipFree(buf);
  if (status == MSQ_OK && fe != hipSuccess)
    status = set_error(MSQ_ERROR_HIP, "hipFree: %s", hipGetErrorString(fe));
  if (status == MSQ_OK && table) std::memcpy(table, rt->tiles, sizeof(rt->tiles));
  return status;
}

int msq_set_tile_table(msq_runtime* rt, const uint8_t* table) {
  MSQ_TRY(activate(rt));
  if (!table) {
    std::memset(rt->tiles, 0, sizeof(rt->tiles));
    return MSQ_OK;
  }
  for (size_t k = 0; k < sizeof(rt->tiles); ++k)
    if (table[k] != 0 && table[k] != 32 && table[k] != 64 && table[k] != 128)
      return set_error(MSQ_ERROR_INVALID_ARGUMENT, "tile edges are 0, 32, 64 or 128");
  std::memcpy(rt->tiles, table, sizeof(rt->tiles));
  return MSQ_OK;
}

}  // extern "C"
