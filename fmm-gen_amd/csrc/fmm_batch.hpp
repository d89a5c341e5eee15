// fmm_batch.hpp -- many independent C += A*B problems in one persistent launch.
//
// A batch holds a problem list as a device job table: every problem
// contributes the jobs of one FMM call (fringe GEMMs, then the FMM core with
// its product order counters) or one plain-GEMM job, each with its own
// algorithm tables and C-tile shape.  The table, its header {njobs,
// total_units} and the workspace (ticket, error flag, order counters) live in
// device memory owned by the batch and grow only.  Running a batch is two
// launches on the caller's stream -- a fill of the whole workspace capacity
// and the persistent kernel over all CUs -- whose arguments depend only on
// the batch's buffers, so a graph captured once replays any later contents
// set with batch_set (same kernel variant, same capacity).  Units are claimed
// from one ticket in job order; a unit waits only for the previous product of
// its own tile position, which holds a lower ticket, so waits end as in a
// single call.
#pragma once
#include "fmm_host.hpp"

namespace fmm {

struct BatchSpec {
  int m, n, k;
  const void* A;
  long long lda;
  const void* B;
  long long ldb;
  void* C;
  long long ldc;
  Plan* plan;   // nullptr: plain GEMM
  int tile;     // 0: default C-tile shape, 1: transposed (low-precision types)
};

struct Batch {
  int dtype = 0;
  int dev = -1;
  JobList JL;
  int max_terms = 1;
  bool has_t = false, has_d = false;
  // device buffers (grow-only) and the pinned staging of the table upload
  Job* djobs = nullptr;
  int* dhdr = nullptr;
  int* dws = nullptr;
  size_t jobs_cap = 0, ws_cap = 0;
  char* staging = nullptr;
  size_t staging_cap = 0;
  // host-mapped pinned word: every run ends by copying its error flags here
  // (err_copy_kernel), so a caller that saw the run complete (an event) reads
  // them without synchronising the device
  int* status_h = nullptr;
  int* status_d = nullptr;
  int generation = 0;   // bumped whenever a device buffer is reallocated
  bool valid = false;   // false after a failed batch_set: runs are refused
  void release() {
    if (djobs) (void)hipFree(djobs);
    if (dhdr) (void)hipFree(dhdr);
    if (dws) (void)hipFree(dws);
    if (staging) (void)hipHostFree(staging);
    if (status_h) (void)hipHostFree(status_h);
    djobs = nullptr;
    dhdr = nullptr;
    dws = nullptr;
    staging = nullptr;
    status_h = status_d = nullptr;
    jobs_cap = ws_cap = staging_cap = 0;
    valid = false;
  }
};

inline int batch_reserve(Batch* b, size_t nj, size_t wsn);

// Kernel variant a batch needs (0..3): FP64 by term count (2: single-term
// products only, 0: up to 2 terms, 1: more), low precision by the C-tile
// shapes present.  A captured launch is valid for one variant.
inline int batch_variant(const Batch& b) {
  if (b.dtype == kF64) return b.max_terms <= 1 ? 2 : b.max_terms <= 2 ? 0 : 1;
  if (b.has_t && b.has_d) return 3;
  return b.has_t ? 2 : 1;
}

// (Re)build batch `b` from `specs` and upload its table with a copy on
// `stream`.  The batch's previous run must have completed (the staging
// buffer and the device table are reused).
inline int batch_set(Batch* b, int dtype, int nspec, const BatchSpec* specs, hipStream_t stream) {
  if (!b || nspec < 0 || (nspec > 0 && !specs)) return kEInval;
  KGeo G0;
  if (!dtype_geo(dtype, &G0)) return kEUnsupported;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  if (b->dev >= 0 && b->dev != dev) return kEInval;
  // The new contents are built in locals and committed only once the upload
  // is enqueued; on any failure the batch is marked invalid, so no run can
  // pair the old device table with the new kernel variant or vice versa.
  b->valid = false;
  JobList JL;
  int max_terms = 1;
  bool has_t = false, has_d = false;
  Prod* ip;
  Term* it;
  int rc = device_tables(identity_plan(), &ip, &it);
  if (rc != kOK) return rc;
  for (int i = 0; i < nspec; ++i) {
    const BatchSpec& s = specs[i];
    if (s.m < 0 || s.n < 0 || s.k < 0 || s.lda < std::max(1, s.m) || s.ldb < std::max(1, s.k) ||
        s.ldc < std::max(1, s.m) || (s.tile != 0 && s.tile != 1))
      return kEInval;
  }
  for (int i = 0; i < nspec; ++i) {
    const BatchSpec& s = specs[i];
    if (s.m == 0 || s.n == 0 || s.k == 0) continue;
    KGeo G;
    int tflag = 0;
    dtype_geo(dtype, &G, s.tile, &tflag);
    const size_t first = JL.jobs.size();
    if (s.plan) {
      Prod* dp;
      Term* dt;
      if ((rc = device_tables(s.plan, &dp, &dt)) != kOK) return rc;
      add_fmm_jobs(JL, G, *s.plan, dp, dt, ip, it, s.m, s.n, s.k, s.A, s.lda, s.B, s.ldb, s.C, s.ldc);
      max_terms = std::max(max_terms, s.plan->max_nnz);
    } else {
      add_gemm_job(JL, G, ip, it, s.m, s.n, s.k, s.A, s.lda, s.B, s.ldb, s.C, s.ldc);
    }
    for (size_t j = first; j < JL.jobs.size(); ++j) {
      JL.jobs[j].flags |= tflag;
      (tflag ? has_t : has_d) = true;
    }
  }
  group_products(JL, 1);
  const size_t nj = JL.jobs.size();
  const size_t wsn = (size_t)kWsCounters + (size_t)JL.counters;
  b->dev = dev;
  int rc2 = batch_reserve(b, nj, wsn);
  if (rc2 != kOK) return rc2;
  const size_t bytes = sizeof(Job) * nj + 2 * sizeof(int);
  if (bytes > b->staging_cap) {
    if (b->staging) (void)hipHostFree(b->staging);
    b->staging_cap = std::max(bytes, 2 * b->staging_cap);
    if (hipHostMalloc(&b->staging, b->staging_cap, hipHostMallocDefault) != hipSuccess) {
      b->staging = nullptr;
      b->staging_cap = 0;
      return kENoMem;
    }
  }
  int hdr[2] = {(int)nj, JL.total_units};
  if (nj) memcpy(b->staging, JL.jobs.data(), sizeof(Job) * nj);
  memcpy(b->staging + sizeof(Job) * nj, hdr, sizeof(hdr));
  if (nj && hipMemcpyAsync(b->djobs, b->staging, sizeof(Job) * nj, hipMemcpyHostToDevice, stream) != hipSuccess)
    return kEHip;
  if (hipMemcpyAsync(b->dhdr, b->staging + sizeof(Job) * nj, sizeof(hdr), hipMemcpyHostToDevice, stream) != hipSuccess)
    return kEHip;
  b->dtype = dtype;
  b->JL = std::move(JL);
  b->max_terms = max_terms;
  b->has_t = has_t;
  b->has_d = has_d;
  b->valid = true;
  return kOK;
}

// Grow the device buffers to hold `nj` jobs and `wsn` workspace ints.
inline int batch_reserve(Batch* b, size_t nj, size_t wsn) {
  if (!b->status_h) {
    if (hipHostMalloc(&b->status_h, 64, hipHostMallocMapped) != hipSuccess) {
      b->status_h = nullptr;
      return kENoMem;
    }
    *b->status_h = 0;
    if (hipHostGetDevicePointer(reinterpret_cast<void**>(&b->status_d), b->status_h, 0) != hipSuccess) {
      (void)hipHostFree(b->status_h);
      b->status_h = b->status_d = nullptr;
      return kEHip;
    }
  }
  if (!b->dhdr) {
    if (hipMalloc(&b->dhdr, 2 * sizeof(int)) != hipSuccess) return kENoMem;
    ++b->generation;
  }
  if (nj > b->jobs_cap || !b->djobs) {
    if (b->djobs) (void)hipFree(b->djobs);
    b->jobs_cap = std::max<size_t>(std::max(nj, 2 * b->jobs_cap), 16);
    if (hipMalloc(&b->djobs, sizeof(Job) * b->jobs_cap) != hipSuccess) {
      b->djobs = nullptr;
      b->jobs_cap = 0;
      return kENoMem;
    }
    ++b->generation;
  }
  if (wsn > b->ws_cap || !b->dws) {
    if (b->dws) (void)hipFree(b->dws);
    b->ws_cap = std::max<size_t>(std::max(wsn, 2 * b->ws_cap), 1024);
    if (hipMalloc(&b->dws, sizeof(int) * b->ws_cap) != hipSuccess) {
      b->dws = nullptr;
      b->ws_cap = 0;
      return kENoMem;
    }
    ++b->generation;
  }
  return kOK;
}

// One persistent launch over the batch's device table (all CUs; the kernel
// reads the job count and unit total from the header).
inline int launch_table(void (*kern)(Launch), int nt, const Batch& b, hipStream_t stream) {
  DevInfo di;
  int rc = dev_info(reinterpret_cast<const void*>(kern), nt, &di);
  if (rc != kOK) return rc;
  Launch L{};
  L.use_ticket = 1;
  L.nq = 1;
  L.local_handover = 0;
  L.jobs = b.djobs;
  L.hdr = b.dhdr;
  L.ws = b.dws;
  L.spin_limit = spin_ticks(di);
  const int n = (int)b.ws_cap;
  hipLaunchKernelGGL(ws_clear_kernel, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, stream, b.dws, n);
  if (hipGetLastError() != hipSuccess) return kEHip;
  hipLaunchKernelGGL(kern, dim3(std::max(1, di.cus * di.blocks_per_cu)), dim3(nt), 0, stream, L);
  if (hipGetLastError() != hipSuccess) return kEHip;
  hipLaunchKernelGGL(err_copy_kernel, dim3(1), dim3(64), 0, stream, b.status_d, b.dws);
  if (hipGetLastError() != hipSuccess) return kEHip;
  return kOK;
}

template <class C1, class C2>
inline int batch_launch_lowp(const Batch& b, hipStream_t s) {
  switch (batch_variant(b)) {
    case 3: return launch_table(abc_lowp_kernel<C1, C2>, C1::NT, b, s);
    case 2: return launch_table(abc_lowp_kernel<C2, C2>, C1::NT, b, s);
    default: return launch_table(abc_lowp_kernel<C1, C1>, C1::NT, b, s);
  }
}

// Query the launch properties of every kernel variant of the batch's dtype
// now (they are cached per device; the first query launches a probe and
// synchronises, which must not happen inside a graph capture of a run).
inline int batch_prepare(const Batch& b) {
  DevInfo di;
  int rc = kOK;
  auto q = [&](void (*k)(Launch), int nt) {
    if (rc == kOK) rc = dev_info(reinterpret_cast<const void*>(k), nt, &di);
  };
  switch (b.dtype) {
    case kF64:
      q(abc_f64_kernel<32, 1>, NT);
      q(abc_f64_kernel<32, 2>, NT);
      q(abc_f64_kernel<16, 4>, NT);
      break;
    case kF32:
      q(abc_lowp_kernel<CfgF32, CfgF32T>, CfgF32::NT);
      q(abc_lowp_kernel<CfgF32T, CfgF32T>, CfgF32::NT);
      q(abc_lowp_kernel<CfgF32, CfgF32>, CfgF32::NT);
      break;
    case kBF16:
      q(abc_lowp_kernel<CfgBF16, CfgBF16T>, CfgBF16::NT);
      q(abc_lowp_kernel<CfgBF16T, CfgBF16T>, CfgBF16::NT);
      q(abc_lowp_kernel<CfgBF16, CfgBF16>, CfgBF16::NT);
      break;
    case kF16:
      q(abc_lowp_kernel<CfgF16, CfgF16T>, CfgF16::NT);
      q(abc_lowp_kernel<CfgF16T, CfgF16T>, CfgF16::NT);
      q(abc_lowp_kernel<CfgF16, CfgF16>, CfgF16::NT);
      break;
    case kBF16x3: q(abc_lowp_kernel<CfgBF16x3, CfgBF16x3>, CfgBF16x3::NT); break;
    default: return kEUnsupported;
  }
  return rc;
}

inline int batch_run(const Batch& b, hipStream_t s) {
  if (!b.valid || !b.djobs || !b.dws || !b.dhdr || !b.status_d) return kEInval;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  if (dev != b.dev) return kEInval;
  switch (b.dtype) {
    case kF64:
      switch (batch_variant(b)) {
        case 2: return launch_table(abc_f64_kernel<32, 1>, NT, b, s);
        case 0: return launch_table(abc_f64_kernel<32, 2>, NT, b, s);
        default: return launch_table(abc_f64_kernel<16, 4>, NT, b, s);
      }
    case kF32: return batch_launch_lowp<CfgF32, CfgF32T>(b, s);
    case kBF16: return batch_launch_lowp<CfgBF16, CfgBF16T>(b, s);
    case kF16: return batch_launch_lowp<CfgF16, CfgF16T>(b, s);
    case kBF16x3: return launch_table(abc_lowp_kernel<CfgBF16x3, CfgBF16x3>, CfgBF16x3::NT, b, s);
    default: return kEUnsupported;
  }
}

}  // namespace fmm
