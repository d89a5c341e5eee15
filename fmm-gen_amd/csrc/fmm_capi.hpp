// fmm_capi.hpp -- the extern "C" entry points of include/fmm_amd.h.
//
// Included by exactly one translation unit per shared library: the runtime
// (fmm_runtime.hip) or a generated per-job source.
#pragma once
#include <cstdlib>

#include "fmm_batch.hpp"
#include "fmm_refine.hpp"
#include "fmm_host.hpp"
#include "fmm_noalloc.hpp"
#include "fmm_amd.h"

struct fmm_plan_s {
  fmm::Plan p;
};

struct fmm_batch_s {
  fmm::Batch b;
};

namespace fmm {

// err_dst: where the flags of this call's ordered launch are OR-ed (a word
// of the caller's status header; nullptr: they stay in ws, or go to the
// device's sticky word when the call allocates ws itself).
inline int run_abc(Plan* P, int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb,
                   void* C, int ldc, void* ws, size_t ws_bytes, hipStream_t s, int* err_dst = nullptr) {
  if (bad_dims(m, n, k, lda, ldb, ldc)) return kEInval;
  KGeo G;
  if (!dtype_geo(dtype, &G, 0, nullptr, P->unit_single)) return kEUnsupported;
  if (m == 0 || n == 0 || k == 0) return kOK;
  Prod* dp;
  Term* dt;
  int rc = device_tables(P, &dp, &dt);
  if (rc != kOK) return rc;
  Prod* ip;
  Term* it;
  rc = device_tables(identity_plan(), &ip, &it);
  if (rc != kOK) return rc;
  JobList JL;
  add_fmm_jobs(JL, G, *P, dp, dt, ip, it, m, n, k, A, lda, B, ldb, C, ldc);
  const size_t need = ws_ints(JL) * sizeof(int);
  int* w = static_cast<int*>(ws);
  bool own = false;
  if (need > 0 && !w) {
    if (hipMallocAsync(reinterpret_cast<void**>(&w), need, s) != hipSuccess) return kENoMem;
    own = true;
  } else if (need > ws_bytes && w) {
    return kEInval;
  }
  rc = launch_jobs_dtype(dtype, JL, P->max_nnz, w, s, P->unit_single);
  if (rc == kOK && need > 0) {
    if (err_dst) rc = fold_error(err_dst, w, s);
    else if (own) rc = fold_owned(w, s);
  }
  if (own) (void)hipFreeAsync(w, s);
  return rc;
}

// One FMM call of pack type `pack`.  AB / Naive need the workspace of
// pack_layout; a shape without an FMM core runs the ABC path (plain GEMMs).
inline int run_fmm(Plan* P, int pack, int dtype, int m, int n, int k, const void* A, int lda, const void* B,
                   int ldb, void* C, int ldc, void* ws, size_t ws_bytes, hipStream_t s) {
  if (pack != kPackABC && pack != kPackAB && pack != kPackNaive) return kEInval;
  if (bad_dims(m, n, k, lda, ldb, ldc)) return kEInval;
  KGeo G;
  if (!dtype_geo(dtype, &G)) return kEUnsupported;
  if (m == 0 || n == 0 || k == 0) return kOK;
  PackLayout Lo;
  const int ks = (k - k % P->Kx) / P->Kx;
  if (pack == kPackABC ||
      !pack_layout(*P, pack, m, n, k, &Lo, G.sin, G.sc, naive_fused(dtype), use_cs(*P, dtype, ks)))
    return run_abc(P, dtype, m, n, k, A, lda, B, ldb, C, ldc, ws, ws_bytes, s);
  void* w = ws;
  bool own = false;
  if (!w) {
    if (hipMallocAsync(&w, Lo.bytes, s) != hipSuccess) return kENoMem;
    own = true;
  } else if (ws_bytes < Lo.bytes) {
    return kEInval;
  }
  int rc = launch_pack(P, pack, dtype, Lo, m, n, k, A, lda, B, ldb, C, ldc, w, s);
  if (own) {
    if (rc == kOK) rc = fold_owned(static_cast<const int*>(w) + Lo.o_off / sizeof(int), s);
    (void)hipFreeAsync(w, s);
  }
  return rc;
}

inline size_t workspace_bytes(Plan* P, int pack, int m, int n, int k) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  PackLayout Lo;
  const int ks = (k - k % P->Kx) / P->Kx;
  if (pack != kPackABC && pack_layout(*P, pack, m, n, k, &Lo, 8, 8, false, use_cs(*P, kF64, ks))) return Lo.bytes;
  JobList JL;
  add_fmm_jobs(JL, *P, nullptr, nullptr, nullptr, nullptr, m, n, k, nullptr, std::max(m, 1), nullptr,
               std::max(k, 1), nullptr, std::max(m, 1));
  return ws_ints(JL) * sizeof(int);
}

inline int run_gemm(int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb, void* C, int ldc,
                    hipStream_t s) {
  if (bad_dims(m, n, k, lda, ldb, ldc)) return kEInval;
  // Split-BF16 with the direct kernel: the identity plan's Naive call splits
  // A and B into hi / lo BF16 planes (its pre-sums) in a workspace of its own
  // and runs the product over them; the same products in the same order as
  // the register-staged split GEMM, so the result is the same bit for bit.
  if (dtype == kBF16x3 && h16_direct())
    return run_fmm(identity_plan(), kPackNaive, dtype, m, n, k, A, lda, B, ldb, C, ldc, nullptr, 0, s);
  KGeo G;
  if (!dtype_geo(dtype, &G, 0, nullptr, true)) return kEUnsupported;
  if (m == 0 || n == 0 || k == 0) return kOK;
  Prod* ip;
  Term* it;
  int rc = device_tables(identity_plan(), &ip, &it);
  if (rc != kOK) return rc;
  JobList JL;
  add_gemm_job(JL, G, ip, it, m, n, k, A, lda, B, ldb, C, ldc);
  return launch_jobs_dtype(dtype, JL, 1, nullptr, s, true);
}

// Staging buffers of the host-pointer drop-ins (grow-only, one set per
// process, serialised by a mutex like the reference's global packing pool).
struct HostStage {
  std::mutex mu;
  double* d[3] = {nullptr, nullptr, nullptr};
  size_t cap[3] = {0, 0, 0};
  int dev = -1;
  void release() {
    for (int i = 0; i < 3; ++i) {
      if (d[i]) (void)hipFree(d[i]);
      d[i] = nullptr;
      cap[i] = 0;
    }
  }
};
inline HostStage& host_stage() {
  static HostStage h;
  return h;
}

inline int stage_buf(HostStage& h, int i, size_t elems) {
  if (h.cap[i] >= elems && h.d[i]) return kOK;
  if (h.d[i]) (void)hipFree(h.d[i]);
  h.d[i] = nullptr;
  h.cap[i] = 0;
  if (hipMalloc(&h.d[i], std::max<size_t>(elems, 1) * sizeof(double)) != hipSuccess) return kENoMem;
  h.cap[i] = elems;
  return kOK;
}

// Copy in, run `body(dA, dB, dC)`, copy C back.
template <class F>
inline int host_call(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C, int ldc,
                     F body) {
  if (bad_dims(m, n, k, lda, ldb, ldc)) return kEInval;
  if (m == 0 || n == 0) return kOK;
  HostStage& h = host_stage();
  std::lock_guard<std::mutex> g(h.mu);
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  if (h.dev != dev) {
    h.release();
    h.dev = dev;
  }
  int rc;
  if ((rc = stage_buf(h, 0, (size_t)m * std::max(k, 1))) != kOK) return rc;
  if ((rc = stage_buf(h, 1, (size_t)std::max(k, 1) * n)) != kOK) return rc;
  if ((rc = stage_buf(h, 2, (size_t)m * n)) != kOK) return rc;
  const size_t es = sizeof(double);
  if (k > 0) {
    if (hipMemcpy2D(h.d[0], m * es, A, (size_t)lda * es, m * es, k, hipMemcpyHostToDevice) != hipSuccess) return kEHip;
    if (hipMemcpy2D(h.d[1], k * es, B, (size_t)ldb * es, k * es, n, hipMemcpyHostToDevice) != hipSuccess) return kEHip;
  }
  if (hipMemcpy2D(h.d[2], m * es, C, (size_t)ldc * es, m * es, n, hipMemcpyHostToDevice) != hipSuccess) return kEHip;
  rc = body(h.d[0], std::max(m, 1), h.d[1], std::max(k, 1), h.d[2], std::max(m, 1));
  if (rc != kOK) return rc;
  if (hipMemcpy2D(C, (size_t)ldc * es, h.d[2], m * es, m * es, n, hipMemcpyDeviceToHost) != hipSuccess) return kEHip;
  return kOK;
}

// End of a host-pointer call: wait for the device and turn the flags its
// library-owned workspaces left into the call's status.
inline int host_finish(int rc) {
  if (rc != kOK) return rc;
  int flags = 0;
  rc = take_sticky(&flags);
  if (rc != kOK) return rc;
  return flags ? kESpin : kOK;
}

inline void report(const char* fn, int rc) {
  if (rc != kOK) fprintf(stderr, "%s(): error %d\n", fn, rc);
}

}  // namespace fmm

extern "C" {

int fmm_plan_create(fmm_plan_t* plan, int nlevels, const int* dims, int rank, const double* U, const double* V,
                    const double* W) {
  if (!plan) return fmm::kEInval;
  *plan = nullptr;
  fmm_plan_s* p = new (std::nothrow) fmm_plan_s();
  if (!p) return fmm::kENoMem;
  int rc = fmm::build_plan(&p->p, nlevels, dims, rank, U, V, W);
  if (rc != fmm::kOK) {
    delete p;
    return rc;
  }
  *plan = p;
  return fmm::kOK;
}

int fmm_plan_destroy(fmm_plan_t plan) {
  if (!plan) return fmm::kEInval;
  fmm::release_plan(&plan->p);
  delete plan;
  return fmm::kOK;
}

int fmm_plan_info(fmm_plan_t plan, int* Mx, int* Kx, int* Nx, int* rank) {
  if (!plan) return fmm::kEInval;
  if (Mx) *Mx = plan->p.Mx;
  if (Kx) *Kx = plan->p.Kx;
  if (Nx) *Nx = plan->p.Nx;
  if (rank) *rank = plan->p.R;
  return fmm::kOK;
}

size_t fmm_workspace_bytes(fmm_plan_t plan, int m, int n, int k) {
  if (!plan) return 0;
  return fmm::workspace_bytes(&plan->p, fmm::kPackABC, m, n, k);
}

size_t fmm_workspace_bytes_pack(fmm_plan_t plan, int pack, int m, int n, int k) {
  if (!plan) return 0;
  return fmm::workspace_bytes(&plan->p, pack, m, n, k);
}

int fmm_run(fmm_plan_t plan, int pack, int dtype, int m, int n, int k, const void* A, int lda, const void* B,
            int ldb, void* C, int ldc, void* ws, size_t ws_bytes, void* stream) {
  if (!plan) return fmm::kEInval;
  return fmm::run_fmm(&plan->p, pack, dtype, m, n, k, A, lda, B, ldb, C, ldc, ws, ws_bytes,
                      static_cast<hipStream_t>(stream));
}

int fmm_abc(fmm_plan_t plan, int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb, void* C,
            int ldc, void* ws, size_t ws_bytes, void* stream) {
  return fmm_run(plan, FMM_PACK_ABC, dtype, m, n, k, A, lda, B, ldb, C, ldc, ws, ws_bytes, stream);
}

int fmm_ab(fmm_plan_t plan, int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb, void* C,
           int ldc, void* ws, size_t ws_bytes, void* stream) {
  return fmm_run(plan, FMM_PACK_AB, dtype, m, n, k, A, lda, B, ldb, C, ldc, ws, ws_bytes, stream);
}

int fmm_naive(fmm_plan_t plan, int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb,
              void* C, int ldc, void* ws, size_t ws_bytes, void* stream) {
  return fmm_run(plan, FMM_PACK_NAIVE, dtype, m, n, k, A, lda, B, ldb, C, ldc, ws, ws_bytes, stream);
}

int fmm_abc_noalloc(fmm_plan_t plan, int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb,
                    void* C, int ldc, void* stream) {
  if (!plan) return fmm::kEInval;
  return fmm::run_abc_noalloc(&plan->p, dtype, m, n, k, A, lda, B, ldb, C, ldc, static_cast<hipStream_t>(stream));
}

int fmm_gemm(int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb, void* C, int ldc,
             void* stream) {
  return fmm::run_gemm(dtype, m, n, k, A, lda, B, ldb, C, ldc, static_cast<hipStream_t>(stream));
}

int fmm_ws_status(const void* ws, int* flags) {
  if (!flags) return fmm::kEInval;
  if (!ws) {
    int rc = fmm::take_sticky(flags);
    int na = 0;
    if (rc == fmm::kOK) rc = fmm::noalloc_take_flags(&na);
    if (rc != fmm::kOK) return rc;
    *flags |= na;
    return *flags ? fmm::kESpin : fmm::kOK;
  }
  if (hipDeviceSynchronize() != hipSuccess) return fmm::kEHip;
  int v = 0;
  if (hipMemcpy(&v, static_cast<const int*>(ws) + fmm::kWsError, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return fmm::kEHip;
  *flags = v;
  return v ? fmm::kESpin : fmm::kOK;
}

int fmm_host_run(fmm_plan_t plan, int pack, int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                 double* C, int ldc) {
  if (!plan) return fmm::kEInval;
  return fmm::host_call(m, n, k, A, lda, B, ldb, C, ldc,
                        [&](const double* a, int la, const double* b, int lb, double* c, int lc) {
                          int rc = fmm::run_fmm(&plan->p, pack, fmm::kF64, m, n, k, a, la, b, lb, c, lc, nullptr, 0, nullptr);
                          return fmm::host_finish(rc);
                        });
}

int fmm_host_abc(fmm_plan_t plan, int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C,
                 int ldc) {
  return fmm_host_run(plan, FMM_PACK_ABC, m, n, k, A, lda, B, ldb, C, ldc);
}

void bl_dgemm(int m, int n, int k, double* XA, int lda, double* XB, int ldb, double* XC, int ldc) {
  int rc = fmm::host_call(m, n, k, XA, lda, XB, ldb, XC, ldc,
                          [&](const double* a, int la, const double* b, int lb, double* c, int lc) {
                            int r = fmm::run_gemm(fmm::kF64, m, n, k, a, la, b, lb, c, lc, nullptr);
                            if (r == fmm::kOK && hipDeviceSynchronize() != hipSuccess) r = fmm::kEHip;
                            return r;
                          });
  fmm::report("bl_dgemm", rc);
}

void bl_dgemm_ref(int m, int n, int k, double* XA, int lda, double* XB, int ldb, double* XC, int ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double c = XC[(size_t)j * ldc + i];
      for (int p = 0; p < k; ++p) c += XA[(size_t)p * lda + i] * XB[(size_t)j * ldb + p];
      XC[(size_t)j * ldc + i] = c;
    }
}

}  // extern "C"

namespace fmm {
inline int batch_specs(int nprob, const fmm_problem_t* probs, int nplans, const fmm_plan_t* plans,
                       std::vector<BatchSpec>* specs) {
  if (nprob < 0 || (nprob > 0 && !probs) || nplans < 0 || (nplans > 0 && !plans)) return kEInval;
  specs->resize((size_t)nprob);
  for (int i = 0; i < nprob; ++i) {
    const fmm_problem_t& p = probs[i];
    if (p.plan < -1 || p.plan >= nplans || (p.plan >= 0 && !plans[p.plan])) return kEInval;
    if (p.tile != 0 && p.tile != 1) return kEInval;
    (*specs)[i] = BatchSpec{p.m, p.n, p.k, p.A, p.lda, p.B, p.ldb, p.C, p.ldc,
                            p.plan >= 0 ? &plans[p.plan]->p : nullptr, p.tile};
  }
  return kOK;
}
}  // namespace fmm

extern "C" {

int fmm_timing(int enable) {
  fmm::LaunchTiming& t = fmm::launch_timing();
  if (enable && !t.ev[0]) {
    if (hipEventCreate(&t.ev[0]) != hipSuccess || hipEventCreate(&t.ev[1]) != hipSuccess) return fmm::kEHip;
  }
  t.on = enable != 0;
  return fmm::kOK;
}

int fmm_timing_last_ms(float* ms) {
  fmm::LaunchTiming& t = fmm::launch_timing();
  if (!ms || !t.ev[0]) return fmm::kEInval;
  if (hipEventSynchronize(t.ev[1]) != hipSuccess) return fmm::kEHip;
  return hipEventElapsedTime(ms, t.ev[0], t.ev[1]) == hipSuccess ? fmm::kOK : fmm::kEHip;
}

int fmm_batch_create(fmm_batch_t* out, int dtype, int nprob, const fmm_problem_t* probs, int nplans,
                     const fmm_plan_t* plans) {
  if (!out) return fmm::kEInval;
  *out = nullptr;
  std::vector<fmm::BatchSpec> specs;
  int rc = fmm::batch_specs(nprob, probs, nplans, plans, &specs);
  if (rc != fmm::kOK) return rc;
  fmm_batch_s* b = new (std::nothrow) fmm_batch_s();
  if (!b) return fmm::kENoMem;
  rc = fmm::batch_set(&b->b, dtype, nprob, specs.data(), nullptr);
  if (rc == fmm::kOK && hipStreamSynchronize(nullptr) != hipSuccess) rc = fmm::kEHip;
  if (rc == fmm::kOK) rc = fmm::batch_prepare(b->b);
  if (rc != fmm::kOK) {
    b->b.release();
    delete b;
    return rc;
  }
  *out = b;
  return fmm::kOK;
}

int fmm_batch_set(fmm_batch_t batch, int nprob, const fmm_problem_t* probs, int nplans, const fmm_plan_t* plans,
                  void* stream) {
  if (!batch) return fmm::kEInval;
  std::vector<fmm::BatchSpec> specs;
  int rc = fmm::batch_specs(nprob, probs, nplans, plans, &specs);
  if (rc != fmm::kOK) return rc;
  return fmm::batch_set(&batch->b, batch->b.dtype, nprob, specs.data(), static_cast<hipStream_t>(stream));
}

int fmm_batch_reserve(fmm_batch_t batch, int jobs, int ws_ints) {
  if (!batch || jobs < 0 || ws_ints < 0) return fmm::kEInval;
  return fmm::batch_reserve(&batch->b, (size_t)jobs, (size_t)ws_ints);
}

int fmm_batch_generation(fmm_batch_t batch) {
  if (!batch) return fmm::kEInval;
  return batch->b.generation;
}

int fmm_batch_variant(fmm_batch_t batch) {
  if (!batch) return fmm::kEInval;
  return fmm::batch_variant(batch->b);
}

int fmm_batch_run(fmm_batch_t batch, void* stream) {
  if (!batch) return fmm::kEInval;
  return fmm::batch_run(batch->b, static_cast<hipStream_t>(stream));
}

int fmm_batch_info(fmm_batch_t batch, int* jobs, int* units) {
  if (!batch) return fmm::kEInval;
  if (jobs) *jobs = (int)batch->b.JL.jobs.size();
  if (units) *units = batch->b.JL.total_units;
  return fmm::kOK;
}

int fmm_batch_status(fmm_batch_t batch, int* flags) {
  if (!batch) return fmm::kEInval;
  return fmm_ws_status(batch->b.dws, flags);
}

int fmm_batch_last_status(fmm_batch_t batch, int* flags) {
  if (!batch || !flags || !batch->b.status_h) return fmm::kEInval;
  *flags = __atomic_load_n(batch->b.status_h, __ATOMIC_ACQUIRE);
  return *flags ? fmm::kESpin : fmm::kOK;
}

int fmm_batch_destroy(fmm_batch_t batch) {
  if (!batch) return fmm::kEInval;
  int cur = 0;
  (void)hipGetDevice(&cur);
  if (batch->b.dev >= 0) (void)hipSetDevice(batch->b.dev);
  batch->b.release();
  (void)hipSetDevice(cur);
  delete batch;
  return fmm::kOK;
}

}  // extern "C"

namespace fmm {
inline bool refine_types_ok(int main_dtype, int corr_dtype) {
  return (main_dtype == kF32 || main_dtype == kBF16 || main_dtype == kF16 || main_dtype == kBF16x3) &&
         (corr_dtype == kF32 || corr_dtype == kBF16x3 || (corr_dtype == kBF16 && main_dtype == kBF16));
}

inline size_t refine_ws(fmm_plan_t mp, int main_dtype, fmm_plan_t cp, int corr_dtype, int m, int n, int k,
                        RefLayout* out) {
  const size_t wm = mp ? workspace_bytes(&mp->p, kPackABC, m, n, k) : 0;
  const size_t wc = cp ? workspace_bytes(&cp->p, kPackABC, m, n, 2 * k) : 0;
  *out = ref_layout(main_dtype, corr_dtype, m, n, k, wm, wc);
  return out->bytes;
}
}  // namespace fmm

extern "C" {

size_t fmm_refine_workspace_bytes(fmm_plan_t main_plan, int main_dtype, fmm_plan_t corr_plan, int corr_dtype, int m,
                                  int n, int k) {
  if (m <= 0 || n <= 0 || k <= 0 || !fmm::refine_types_ok(main_dtype, corr_dtype)) return 0;
  fmm::RefLayout L;
  return fmm::refine_ws(main_plan, main_dtype, corr_plan, corr_dtype, m, n, k, &L);
}

int fmm_refine(fmm_plan_t main_plan, int main_dtype, fmm_plan_t corr_plan, int corr_dtype, int m, int n, int k,
               const double* A, int lda, const double* B, int ldb, double* C, int ldc, void* ws, size_t ws_bytes,
               void* stream) {
  using namespace fmm;
  if (bad_dims(m, n, k, lda, ldb, ldc)) return kEInval;
  if (!refine_types_ok(main_dtype, corr_dtype)) return kEUnsupported;
  if (m == 0 || n == 0 || k == 0) return kOK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  RefLayout Lo;
  const size_t need = refine_ws(main_plan, main_dtype, corr_plan, corr_dtype, m, n, k, &Lo);
  char* w = static_cast<char*>(ws);
  bool own = false;
  if (!w) {
    if (hipMallocAsync(reinterpret_cast<void**>(&w), need, s) != hipSuccess) return kENoMem;
    own = true;
  } else if (ws_bytes < need) {
    return kEInval;
  }
  int rc0 = kOK;
  // status header: the error word the sub-calls' ordered launches fold into
  int* hdr = reinterpret_cast<int*>(w + Lo.hdr);
  hipLaunchKernelGGL(ws_clear_kernel, dim3(1), dim3(256), 0, s, hdr, kWsCounters);
  if (hipGetLastError() != hipSuccess) rc0 = kEHip;
  void* AA = w + Lo.aa;
  void* BB = w + Lo.bb;
  float* Ct = reinterpret_cast<float*>(w + Lo.ct);
  float* Et = reinterpret_cast<float*>(w + Lo.et);
  const long long mk = (long long)m * k, kn = (long long)k * n, mn = (long long)Lo.lt * n;
  int rc = rc0;
  auto chk = [&]() {
    if (rc == kOK && hipGetLastError() != hipSuccess) rc = kEHip;
  };
  u16* a1h = reinterpret_cast<u16*>(w + Lo.a1);
  u16* b1h = reinterpret_cast<u16*>(w + Lo.b1);
  float* a1f = reinterpret_cast<float*>(w + Lo.a1);
  float* b1f = reinterpret_cast<float*>(w + Lo.b1);
  float* aaf = static_cast<float*>(AA);
  float* bbf = static_cast<float*>(BB);
  u16* aah = static_cast<u16*>(AA);
  u16* bbh = static_cast<u16*>(BB);
  const int ga = ew_blocks(mk), gb = ew_blocks(kn);
  if (main_dtype == kBF16 && corr_dtype == kBF16) {
    split_a_kernel<kInBF16, kInBF16><<<ga, kRefThreads, 0, s>>>(m, k, A, lda, a1h, Lo.la1, aah, Lo.laa);
    split_b_kernel<kInBF16, kInBF16><<<gb, kRefThreads, 0, s>>>(k, n, B, ldb, b1h, Lo.lb1, bbh, Lo.lbb);
  } else if (main_dtype == kBF16) {
    split_a_kernel<kInBF16, kInF32><<<ga, kRefThreads, 0, s>>>(m, k, A, lda, a1h, Lo.la1, aaf, Lo.laa);
    split_b_kernel<kInBF16, kInF32><<<gb, kRefThreads, 0, s>>>(k, n, B, ldb, b1h, Lo.lb1, bbf, Lo.lbb);
  } else if (main_dtype == kF16) {
    split_a_kernel<kInF16, kInF32><<<ga, kRefThreads, 0, s>>>(m, k, A, lda, a1h, Lo.la1, aaf, Lo.laa);
    split_b_kernel<kInF16, kInF32><<<gb, kRefThreads, 0, s>>>(k, n, B, ldb, b1h, Lo.lb1, bbf, Lo.lbb);
  } else {   // FP32 inputs (FMM_F32, FMM_BF16X3)
    split_a_kernel<kInF32, kInF32><<<ga, kRefThreads, 0, s>>>(m, k, A, lda, a1f, Lo.la1, aaf, Lo.laa);
    split_b_kernel<kInF32, kInF32><<<gb, kRefThreads, 0, s>>>(k, n, B, ldb, b1f, Lo.lb1, bbf, Lo.lbb);
  }
  chk();
  zero_f32_kernel<<<ew_blocks(mn), kRefThreads, 0, s>>>(Ct, mn);
  zero_f32_kernel<<<ew_blocks(mn), kRefThreads, 0, s>>>(Et, mn);
  chk();
  if (rc == kOK)
    rc = main_plan ? run_abc(&main_plan->p, main_dtype, m, n, k, w + Lo.a1, (int)Lo.la1, w + Lo.b1, (int)Lo.lb1, Ct,
                             (int)Lo.lt, w + Lo.wsm, Lo.wsc - Lo.wsm, s, hdr + kWsError)
                   : run_gemm(main_dtype, m, n, k, w + Lo.a1, (int)Lo.la1, w + Lo.b1, (int)Lo.lb1, Ct, (int)Lo.lt, s);
  if (rc == kOK)
    rc = corr_plan ? run_abc(&corr_plan->p, corr_dtype, m, n, 2 * k, AA, (int)Lo.laa, BB, (int)Lo.lbb, Et,
                             (int)Lo.lt, w + Lo.wsc, Lo.bytes - Lo.wsc, s, hdr + kWsError)
                   : run_gemm(corr_dtype, m, n, 2 * k, AA, (int)Lo.laa, BB, (int)Lo.lbb, Et, (int)Lo.lt, s);
  if (rc == kOK) {
    combine_kernel<<<ew_blocks((long long)m * n), kRefThreads, 0, s>>>(m, n, Ct, Et, Lo.lt, C, ldc);
    chk();
  }
  if (own) {
    if (rc == kOK) rc = fold_owned(hdr, s);
    (void)hipFreeAsync(w, s);
  }
  return rc;
}

void bl_finalize(void) {
  fmm::HostStage& h = fmm::host_stage();
  std::lock_guard<std::mutex> g(h.mu);
  h.release();
  h.dev = -1;
}

}  // extern "C"
