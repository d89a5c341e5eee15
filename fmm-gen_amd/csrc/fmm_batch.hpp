// fmm_batch.hpp -- many independent C += A*B problems in one persistent launch.
//
// A batch is built once from a problem list: every problem contributes the
// jobs of one FMM call (fringe GEMMs, then the FMM core with its product
// order counters) or one plain-GEMM job, each with its own algorithm tables
// and C-tile shape.  The job table and the workspace (tickets, error flag,
// order counters) live in device memory, so running the batch is two kernel
// launches on the caller's stream (a workspace fill and the persistent
// kernel), with no allocation or copy: it can be captured in a HIP graph and
// replayed.  Units are claimed from one ticket in job order; a unit waits only
// for the previous product of its own tile position, which holds a lower
// ticket, so waits end as in a single call.
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
  Job* djobs = nullptr;
  int* dws = nullptr;
  size_t ws_n = 0;
  void release() {
    if (djobs) (void)hipFree(djobs);
    if (dws) (void)hipFree(dws);
    djobs = nullptr;
    dws = nullptr;
  }
};

inline int batch_build(Batch* b, int dtype, int nspec, const BatchSpec* specs) {
  if (!b || nspec < 0 || (nspec > 0 && !specs)) return kEInval;
  KGeo G0;
  if (!dtype_geo(dtype, &G0)) return kEUnsupported;
  if (hipGetDevice(&b->dev) != hipSuccess) return kEHip;
  b->dtype = dtype;
  Prod* ip;
  Term* it;
  int rc = device_tables(identity_plan(), &ip, &it);
  if (rc != kOK) return rc;
  for (int i = 0; i < nspec; ++i) {
    const BatchSpec& s = specs[i];
    if (s.m < 0 || s.n < 0 || s.k < 0 || s.lda < std::max(1, s.m) || s.ldb < std::max(1, s.k) ||
        s.ldc < std::max(1, s.m))
      return kEInval;
    if (s.m == 0 || s.n == 0 || s.k == 0) continue;
    KGeo G;
    int tflag = 0;
    dtype_geo(dtype, &G, s.tile, &tflag);
    const size_t first = b->JL.jobs.size();
    if (s.plan) {
      Prod* dp;
      Term* dt;
      if ((rc = device_tables(s.plan, &dp, &dt)) != kOK) return rc;
      add_fmm_jobs(b->JL, G, *s.plan, dp, dt, ip, it, s.m, s.n, s.k, s.A, s.lda, s.B, s.ldb, s.C, s.ldc);
      b->max_terms = std::max(b->max_terms, s.plan->max_nnz);
    } else {
      add_gemm_job(b->JL, G, ip, it, s.m, s.n, s.k, s.A, s.lda, s.B, s.ldb, s.C, s.ldc);
    }
    for (size_t j = first; j < b->JL.jobs.size(); ++j) {
      b->JL.jobs[j].flags |= tflag;
      (tflag ? b->has_t : b->has_d) = true;
    }
  }
  group_products(b->JL, 1);
  b->ws_n = (size_t)kWsCounters + (size_t)b->JL.counters;
  const size_t jb = sizeof(Job) * std::max<size_t>(b->JL.jobs.size(), 1);
  if (hipMalloc(&b->djobs, jb) != hipSuccess) return kENoMem;
  if (hipMalloc(&b->dws, b->ws_n * sizeof(int)) != hipSuccess) return kENoMem;
  if (!b->JL.jobs.empty() &&
      hipMemcpy(b->djobs, b->JL.jobs.data(), sizeof(Job) * b->JL.jobs.size(), hipMemcpyHostToDevice) != hipSuccess)
    return kEHip;
  return kOK;
}

// One persistent launch over a device job table.
inline int launch_table(void (*kern)(Launch), int nt, const Batch& b, hipStream_t stream) {
  DevInfo di;
  int rc = dev_info(reinterpret_cast<const void*>(kern), nt, &di);
  if (rc != kOK) return rc;
  Launch L{};
  L.njobs = (int)b.JL.jobs.size();
  L.total_units = b.JL.total_units;
  L.use_ticket = 1;
  L.nq = 1;
  L.local_handover = 0;
  L.jobs = b.djobs;
  L.ws = b.dws;
  L.spin_limit = di.spin_ticks;
  const int n = (int)b.ws_n;
  hipLaunchKernelGGL(ws_clear_kernel, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, stream, b.dws, n);
  if (hipGetLastError() != hipSuccess) return kEHip;
  const int grid = std::max(1, std::min(b.JL.total_units, di.cus * di.blocks_per_cu));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), 0, stream, L);
  if (hipGetLastError() != hipSuccess) return kEHip;
  return kOK;
}

template <class C1, class C2>
inline int batch_launch_lowp(const Batch& b, hipStream_t s) {
  if (b.has_t && b.has_d) return launch_table(abc_lowp_kernel<C1, C2>, C1::NT, b, s);
  if (b.has_t) return launch_table(abc_lowp_kernel<C2, C2>, C1::NT, b, s);
  return launch_table(abc_lowp_kernel<C1, C1>, C1::NT, b, s);
}

inline int batch_run(const Batch& b, hipStream_t s) {
  if (b.JL.total_units == 0) return kOK;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  if (dev != b.dev) return kEInval;
  switch (b.dtype) {
    case kF64:
      return b.max_terms <= 2 ? launch_table(abc_f64_kernel<32, 2>, NT, b, s)
                              : launch_table(abc_f64_kernel<16, 4>, NT, b, s);
    case kF32: return batch_launch_lowp<CfgF32, CfgF32T>(b, s);
    case kBF16: return batch_launch_lowp<CfgBF16, CfgBF16T>(b, s);
    case kF16: return batch_launch_lowp<CfgF16, CfgF16T>(b, s);
    case kBF16x3: return launch_table(abc_lowp_kernel<CfgBF16x3, CfgBF16x3>, CfgBF16x3::NT, b, s);
    default: return kEUnsupported;
  }
}

}  // namespace fmm
