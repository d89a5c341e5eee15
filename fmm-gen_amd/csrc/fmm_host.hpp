// fmm_host.hpp -- host side: plans, partition, job lists and launches.
//
// Shared by the generic runtime library (fmm_runtime.hip) and by the
// per-job libraries the generator emits (their tables are compiled in).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "fmm_core.hpp"
#include "fmm_cs.hpp"
#include "fmm_lowp.hpp"
#include "fmm_h16d.hpp"
#include "fmm_f64d.hpp"
#include "fmm_csd.hpp"

namespace fmm {

// Status codes of the C-ABI (mirrored in include/fmm_amd.h).
enum Status : int {
  kOK = 0,
  kEInval = -1,
  kEHip = -2,
  kENoMem = -3,
  kEUnsupported = -4,
  kESpin = -5,
  kETooBig = -6,
};

inline bool bad_dims(int m, int n, int k, int lda, int ldb, int ldc) {
  if (m < 0 || n < 0 || k < 0) return true;
  if (lda < std::max(1, m) || ldb < std::max(1, k) || ldc < std::max(1, m)) return true;
  return false;
}

// XCD-local C hand-over between ordered products (see release_c); 0 keeps
// the agent-scope release everywhere.  At run time FMM_LOCAL_HANDOVER=0 in
// the environment does the same for the launches that follow.
#ifndef FMM_LOCAL_HANDOVER
#define FMM_LOCAL_HANDOVER 1
#endif
inline bool local_handover_allowed() {
  if (!FMM_LOCAL_HANDOVER) return false;
  const char* e = getenv("FMM_LOCAL_HANDOVER");
  return !(e && *e && atoi(e) == 0);
}

// Partition granularity of the reference (common/include/bl_config.h:16-17):
// the FMM core is md x kd x nd with md % (Mx*MR) == 0, nd % (Nx*NR) == 0.
constexpr int kMR = 8;
constexpr int kNR = 4;

// Host vector with lazily created copies on each device ordinal.
template <class T>
struct DevVec {
  std::vector<T> h;
  std::vector<T*> d;
  int get(int dev, T** out) {
    if ((int)d.size() <= dev) d.resize(dev + 1, nullptr);
    if (!d[dev]) {
      T* p = nullptr;
      const size_t bytes = sizeof(T) * std::max<size_t>(h.size(), 1);
      if (hipMalloc(&p, bytes) != hipSuccess) return kENoMem;
      if (!h.empty() && hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(p);
        return kEHip;
      }
      d[dev] = p;
    }
    *out = d[dev];
    return kOK;
  }
  void release() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    for (size_t i = 0; i < d.size(); ++i)
      if (d[i]) {
        (void)hipSetDevice((int)i);
        (void)hipFree(d[i]);
      }
    (void)hipSetDevice(cur);
    d.clear();
  }
};

// A product table: products with their term ranges.
struct Table {
  DevVec<Prod> prods;
  DevVec<Term> terms;
};

struct Plan {
  unsigned long long uid = 0;   // unique per built plan in this process (never 0)
  int nlev = 0;
  std::vector<int> dims;   // 3*nlev
  int Mx = 1, Kx = 1, Nx = 1, R = 0;
  int max_nnz = 0;         // largest per-product source count (A or B)
  bool unit_single = true; // every product: one A and one B term, coefficient +-1
  // ABC: R products, then Mx*Nx k-fringe products.
  Table main;
  // Naive: product r reads workspace blocks S_r, T_r and writes M_r (FP64)...
  Table diag;
  // ... the same products split by which operands have a single term and are
  // read directly (bit 0: A, bit 1: B) instead of through S_r / T_r, in
  // product order within each table (FP64 M-writing launch).
  Table diag_by[4];
  // ... or updates its C blocks (the W terms of `main`) straight from the
  // accumulators, in product order per C tile, as ABC (other types).
  Table naive;
  // AB: product r reads its A/B terms as in ABC and writes M_r.
  Table ab;
  // Naive, C-stationary launch: product r reads S_r / T_r (term bi = -1,
  // bj = r) or, for an operand with a single term, that block of A / B
  // directly, and carries the W terms of `main`; then the k-fringe products
  // of `main`.
  Table naive_cs;
  // For every C block (I*Nx + J, grid order) the products feeding it, in
  // product order, with their W coefficients (the M_Add lists).
  DevVec<int> cptr, cprod;
  DevVec<double> cw;
  // The same terms product by product: for product r (R + b: the k-fringe
  // product of C block b) its C blocks and W coefficients (one-pass M_Add).
  DevVec<int> rptr, rblk;
  DevVec<double> rw;
  // The same split over the 4 waves of a multi-wave M_Add workgroup: wave w
  // owns the C blocks b with b % 4 == w; its entries wbeg[w] .. wbeg[w+1]
  // list, in product order, the products with a term on its blocks (wprod)
  // and those terms (wptr into wblk / ww).
  DevVec<int> wbeg, wprod, wptr, wblk;
  DevVec<double> ww;
  std::mutex mu;
};

// Decode composed row `row` of table `which` (0: U -> A(i,p), 1: V -> B(p,j),
// 2: W -> C(i,j)) into its block in the composed grid; digits are peeled off
// from the last level (meta/common_mix.py:296-310).
inline void decode_row(const std::vector<int>& dims, int which, int row, int* I, int* J) {
  const int a = which == 1 ? 1 : 0, b = which == 0 ? 1 : 2;
  int rb = 1, cb = 1, ii = 0, jj = 0, rest = row;
  const int nlev = (int)dims.size() / 3;
  for (int l = nlev - 1; l >= 0; --l) {
    const int nr = dims[3 * l + a], nc = dims[3 * l + b];
    const int d = rest % (nr * nc);
    rest /= nr * nc;
    ii += (d / nc) * rb;
    jj += (d % nc) * cb;
    rb *= nr;
    cb *= nc;
  }
  *I = ii;
  *J = jj;
}

// Tables are dense row-major [rows x R] in the reference's composed row order.
inline int build_plan(Plan* P, int nlev, const int* dims, int R, const double* U, const double* V,
                      const double* W) {
  if (!P || nlev <= 0 || !dims || R <= 0 || !U || !V || !W) return kEInval;
  static std::atomic<unsigned long long> next_uid{1};
  P->uid = next_uid.fetch_add(1);
  P->nlev = nlev;
  P->dims.assign(dims, dims + 3 * nlev);
  P->Mx = P->Kx = P->Nx = 1;
  for (int l = 0; l < nlev; ++l) {
    if (dims[3 * l] <= 0 || dims[3 * l + 1] <= 0 || dims[3 * l + 2] <= 0) return kEInval;
    P->Mx *= dims[3 * l];
    P->Kx *= dims[3 * l + 1];
    P->Nx *= dims[3 * l + 2];
  }
  if ((long long)P->Mx * P->Kx > (1 << 20) || (long long)P->Kx * P->Nx > (1 << 20) ||
      (long long)P->Mx * P->Nx > (1 << 20))
    return kEInval;
  P->R = R;
  std::vector<Prod>& prods = P->main.prods.h;
  std::vector<Term>& terms = P->main.terms.h;
  prods.clear();
  terms.clear();
  P->max_nnz = 0;
  P->unit_single = true;
  const int rowsU = P->Mx * P->Kx, rowsV = P->Kx * P->Nx, rowsW = P->Mx * P->Nx;
  for (int r = 0; r < R; ++r) {
    Prod pr{};
    auto collect = [&](const double* T, int rows, int which, int* start, int* cnt) {
      *start = (int)terms.size();
      for (int a = 0; a < rows; ++a) {
        const double v = T[(size_t)a * R + r];
        if (v == 0.0) continue;
        Term t;
        t.coef = v;
        decode_row(P->dims, which, a, &t.bi, &t.bj);
        terms.push_back(t);
      }
      *cnt = (int)terms.size() - *start;
    };
    collect(U, rowsU, 0, &pr.a0, &pr.na);
    collect(V, rowsV, 1, &pr.b0, &pr.nb);
    collect(W, rowsW, 2, &pr.c0, &pr.nc);
    pr.fringe = 0;
    P->max_nnz = std::max(P->max_nnz, std::max(pr.na, pr.nb));
    if (!(pr.na == 1 && pr.nb == 1 && std::fabs(terms[pr.a0].coef) == 1.0 && std::fabs(terms[pr.b0].coef) == 1.0))
      P->unit_single = false;
    prods.push_back(pr);
  }
  // k-fringe: C(I,J) += A(I, Kx) B(Kx, J) over depth kr, as one-term products
  // ordered after all R products (bl_dgemm_util.c:325-333 "kr > 0" branch).
  for (int I = 0; I < P->Mx; ++I)
    for (int J = 0; J < P->Nx; ++J) {
      Prod pr{};
      pr.a0 = (int)terms.size();
      terms.push_back(Term{1.0, I, P->Kx});
      pr.na = 1;
      pr.b0 = (int)terms.size();
      terms.push_back(Term{1.0, P->Kx, J});
      pr.nb = 1;
      pr.c0 = (int)terms.size();
      terms.push_back(Term{1.0, I, J});
      pr.nc = 1;
      pr.fringe = 1;
      prods.push_back(pr);
    }
  // Naive and AB tables.  The workspaces hold one leaf block per product,
  // side by side: S_r, T_r and M_r are block (0, r) of S (ld >= ms), T
  // (ld >= ks) and M (ld >= ms).  In the M-writing tables the Mx*Nx k-fringe
  // products follow as products R + I*Nx + J, reading A and B and writing M
  // block R + I*Nx + J; the fused Naive launch runs the k-fringe products of
  // `main` in a launch of their own.
  P->diag.prods.h.clear();
  P->diag.terms.h.clear();
  for (Table& t : P->diag_by) {
    t.prods.h.clear();
    t.terms.h.clear();
  }
  P->naive.prods.h.clear();
  P->naive.terms.h.clear();
  P->ab.prods.h.clear();
  P->ab.terms.h.clear();
  P->naive_cs.prods.h.clear();
  P->naive_cs.terms.h.clear();
  for (int r = 0; r < R + P->Mx * P->Nx; ++r) {
    const Prod& o = prods[r];
    std::vector<Term>& ct = P->naive_cs.terms.h;
    Prod c{};
    c.a0 = (int)ct.size();
    ct.push_back(r >= R || o.na == 1 ? terms[o.a0] : Term{1.0, -1, r});
    c.na = 1;
    c.b0 = (int)ct.size();
    ct.push_back(r >= R || o.nb == 1 ? terms[o.b0] : Term{1.0, -1, r});
    c.nb = 1;
    c.c0 = (int)ct.size();
    for (int t = 0; t < o.nc; ++t) ct.push_back(terms[o.c0 + t]);
    c.nc = o.nc;
    c.fringe = o.fringe;
    P->naive_cs.prods.h.push_back(c);
  }
  for (int r = 0; r < R; ++r) {
    std::vector<Term>& dt = P->diag.terms.h;
    Prod d{};
    d.a0 = (int)dt.size();
    dt.push_back(Term{1.0, 0, r});
    d.na = 1;
    d.b0 = (int)dt.size();
    dt.push_back(Term{1.0, 0, r});
    d.nb = 1;
    d.c0 = (int)dt.size();
    dt.push_back(Term{1.0, 0, r});
    d.nc = 1;
    P->diag.prods.h.push_back(d);

    {
      const Prod& o = prods[r];
      const int combo = (o.na == 1 ? 1 : 0) | (o.nb == 1 ? 2 : 0);
      std::vector<Term>& ct = P->diag_by[combo].terms.h;
      Prod c{};
      c.a0 = (int)ct.size();
      ct.push_back(o.na == 1 ? terms[o.a0] : Term{1.0, 0, r});
      c.na = 1;
      c.b0 = (int)ct.size();
      ct.push_back(o.nb == 1 ? terms[o.b0] : Term{1.0, 0, r});
      c.nb = 1;
      c.c0 = (int)ct.size();
      ct.push_back(Term{1.0, 0, r});
      c.nc = 1;
      P->diag_by[combo].prods.h.push_back(c);
    }

    std::vector<Term>& ft = P->naive.terms.h;
    Prod f{};
    f.a0 = (int)ft.size();
    ft.push_back(Term{1.0, 0, r});
    f.na = 1;
    f.b0 = (int)ft.size();
    ft.push_back(Term{1.0, 0, r});
    f.nb = 1;
    f.c0 = (int)ft.size();
    for (int t = 0; t < prods[r].nc; ++t) ft.push_back(terms[prods[r].c0 + t]);
    f.nc = prods[r].nc;
    P->naive.prods.h.push_back(f);

    std::vector<Term>& at = P->ab.terms.h;
    const Prod& o = prods[r];
    Prod a{};
    a.a0 = (int)at.size();
    for (int t = 0; t < o.na; ++t) at.push_back(terms[o.a0 + t]);
    a.na = o.na;
    a.b0 = (int)at.size();
    for (int t = 0; t < o.nb; ++t) at.push_back(terms[o.b0 + t]);
    a.nb = o.nb;
    a.c0 = (int)at.size();
    at.push_back(Term{1.0, 0, r});
    a.nc = 1;
    P->ab.prods.h.push_back(a);
  }
  for (Table* T : {&P->diag, &P->ab})
    for (int cb = 0; cb < P->Mx * P->Nx; ++cb) {
      std::vector<Term>& t = T->terms.h;
      Prod pr{};
      pr.a0 = (int)t.size();
      t.push_back(Term{1.0, cb / P->Nx, P->Kx});
      pr.na = 1;
      pr.b0 = (int)t.size();
      t.push_back(Term{1.0, P->Kx, cb % P->Nx});
      pr.nb = 1;
      pr.c0 = (int)t.size();
      t.push_back(Term{1.0, 0, R + cb});
      pr.nc = 1;
      pr.fringe = 1;
      T->prods.h.push_back(pr);
    }
  // M_Add lists per C block.
  P->cptr.h.assign(1, 0);
  P->cprod.h.clear();
  P->cw.h.clear();
  for (int cb = 0; cb < P->Mx * P->Nx; ++cb) {
    const int I = cb / P->Nx, J = cb % P->Nx;
    for (int r = 0; r < R; ++r) {
      const Prod& o = prods[r];
      for (int t = 0; t < o.nc; ++t) {
        const Term& c = terms[o.c0 + t];
        if (c.bi == I && c.bj == J) {
          P->cprod.h.push_back(r);
          P->cw.h.push_back(c.coef);
        }
      }
    }
    P->cprod.h.push_back(R + cb);   // its k-fringe product, last (as in ABC)
    P->cw.h.push_back(1.0);
    P->cptr.h.push_back((int)P->cprod.h.size());
  }
  P->rptr.h.assign(1, 0);
  P->rblk.h.clear();
  P->rw.h.clear();
  for (int r = 0; r < R; ++r) {
    const Prod& o = prods[r];
    for (int t = 0; t < o.nc; ++t) {
      const Term& c = terms[o.c0 + t];
      P->rblk.h.push_back(c.bi * P->Nx + c.bj);
      P->rw.h.push_back(c.coef);
    }
    P->rptr.h.push_back((int)P->rblk.h.size());
  }
  for (int cb = 0; cb < P->Mx * P->Nx; ++cb) {
    P->rblk.h.push_back(cb);
    P->rw.h.push_back(1.0);
    P->rptr.h.push_back((int)P->rblk.h.size());
  }
  P->wbeg.h.assign(1, 0);
  P->wprod.h.clear();
  P->wptr.h.assign(1, 0);
  P->wblk.h.clear();
  P->ww.h.clear();
  const int nr = (int)P->rptr.h.size() - 1;   // R + Mx*Nx products
  for (int w = 0; w < kMa4Waves; ++w) {
    for (int r = 0; r < nr; ++r) {
      bool any = false;
      for (int l = P->rptr.h[r]; l < P->rptr.h[r + 1]; ++l) {
        if (P->rblk.h[l] % kMa4Waves != w) continue;
        P->wblk.h.push_back(P->rblk.h[l]);
        P->ww.h.push_back(P->rw.h[l]);
        any = true;
      }
      if (any) {
        P->wprod.h.push_back(r);
        P->wptr.h.push_back((int)P->wblk.h.size());
      }
    }
    P->wbeg.h.push_back((int)P->wprod.h.size());
  }
  return kOK;
}

inline int device_tables(Plan* P, Table& T, Prod** dp, Term** dt) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  std::lock_guard<std::mutex> g(P->mu);
  int rc = T.prods.get(dev, dp);
  if (rc == kOK) rc = T.terms.get(dev, dt);
  return rc;
}

inline int device_tables(Plan* P, Prod** dp, Term** dt) { return device_tables(P, P->main, dp, dt); }

inline int device_clists(Plan* P, int** ptr, int** prod, double** w) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  std::lock_guard<std::mutex> g(P->mu);
  int rc = P->cptr.get(dev, ptr);
  if (rc == kOK) rc = P->cprod.get(dev, prod);
  if (rc == kOK) rc = P->cw.get(dev, w);
  return rc;
}

inline int device_wlists(Plan* P, int** beg, int** prod, int** ptr, int** blk, double** w) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  std::lock_guard<std::mutex> g(P->mu);
  int rc = P->wbeg.get(dev, beg);
  if (rc == kOK) rc = P->wprod.get(dev, prod);
  if (rc == kOK) rc = P->wptr.get(dev, ptr);
  if (rc == kOK) rc = P->wblk.get(dev, blk);
  if (rc == kOK) rc = P->ww.get(dev, w);
  return rc;
}

inline int device_rlists(Plan* P, int** ptr, int** blk, double** w) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  std::lock_guard<std::mutex> g(P->mu);
  int rc = P->rptr.get(dev, ptr);
  if (rc == kOK) rc = P->rblk.get(dev, blk);
  if (rc == kOK) rc = P->rw.get(dev, w);
  return rc;
}

#ifndef FMM_MADD_PF_LO
#define FMM_MADD_PF_LO 16
#endif
#ifndef FMM_MADD_PF_HI
#define FMM_MADD_PF_HI 32
#endif
#ifndef FMM_MADD_PF_SPLIT
#define FMM_MADD_PF_SPLIT 16
#endif

// FMM_MADD_ALL4=0 keeps algorithms with many C blocks on the one-wave
// M_Add kernel (read per call).
inline bool madd_all4_enabled() {
  const char* e = getenv("FMM_MADD_ALL4");
  return !(e && e[0] == '0');
}

inline bool madd_all_enabled() {
  static const bool on = [] {
    const char* e = getenv("FMM_MADD_ALL");
    return !(e && e[0] == '0');
  }();
  return on;
}

inline void release_plan(Plan* P) {
  for (Table* T : {&P->main, &P->diag, &P->diag_by[0], &P->diag_by[1], &P->diag_by[2], &P->diag_by[3], &P->naive,
                   &P->ab, &P->naive_cs}) {
    T->prods.release();
    T->terms.release();
  }
  P->cptr.release();
  P->cprod.release();
  P->cw.release();
  P->rptr.release();
  P->rblk.release();
  P->rw.release();
  P->wbeg.release();
  P->wprod.release();
  P->wptr.release();
  P->wblk.release();
  P->ww.release();
}

// The plain-GEMM plan: one product, A(0,0) B(0,0) -> C(0,0).
inline Plan* identity_plan() {
  static Plan* P = [] {
    Plan* p = new Plan();
    const int dims[3] = {1, 1, 1};
    const double one = 1.0;
    build_plan(p, 1, dims, 1, &one, &one, &one);
    return p;
  }();
  return P;
}

// --- per-device launch properties ------------------------------------------
struct DevInfo {
  int cus = 0;
  int blocks_per_cu = 0;
  long long spin_ticks = 0;       // order-wait bound (FMM_SPIN_LIMIT_MS, default 5 s)
  long long ticks_per_ms = 0;
  bool xcd_round_robin = false;   // workgroup i runs on XCC i % 8 (probed)
};

// Bound of one product-order wait, in wall_clock64() ticks: FMM_SPIN_LIMIT_MS
// milliseconds when that variable is set (0 flags every wait that is not
// satisfied at once; used to exercise the error path), else 5 s.
inline long long spin_ticks(const DevInfo& d) {
  const char* e = getenv("FMM_SPIN_LIMIT_MS");
  if (e && *e) {
    const long long ms = atoll(e);
    if (ms >= 0) return ms * d.ticks_per_ms;
  }
  return d.spin_ticks;
}

// One-time probe: launch 256 workgroups and check that workgroup i reports
// XCC id i % 8 (the dispatch pattern XCD-local hand-over relies on).  `buf`:
// kXcdProbeInts device ints to use, or nullptr (allocated and freed here).
constexpr int kXcdProbeInts = 256;
inline bool probe_xcd_round_robin(int* buf) {
  const int n = kXcdProbeInts;
  int* d = buf;
  if (!d && hipMalloc(&d, n * sizeof(int)) != hipSuccess) return false;
  bool ok = false;
  xcc_probe_kernel<<<dim3(n), dim3(64), 0, 0>>>(d);
  std::vector<int> h(n, -1);
  if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
      hipMemcpy(h.data(), d, n * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess) {
    ok = true;
    for (int i = 0; i < n; ++i) ok = ok && h[i] == i % kMaxQueues;
  }
  if (!buf) (void)hipFree(d);
  return ok;
}

// Per-device result of the probe: -1 not yet probed, else 0 / 1.
struct XcdProbe {
  std::mutex mu;
  std::vector<int> rr;
};
inline XcdProbe& xcd_probe() {
  static XcdProbe p;
  return p;
}

// Probe the current device once (with `buf` as in probe_xcd_round_robin);
// returns whether workgroups are placed round robin over the XCDs.
inline bool ensure_xcd_probe(int dev, int* buf) {
  XcdProbe& p = xcd_probe();
  std::lock_guard<std::mutex> g(p.mu);
  if ((int)p.rr.size() <= dev) p.rr.resize(dev + 1, -1);
  if (p.rr[dev] < 0) p.rr[dev] = probe_xcd_round_robin(buf) ? 1 : 0;
  return p.rr[dev] == 1;
}

// Launch properties of `kernel` (nt threads per workgroup) on the current
// device, cached per (device, kernel).
inline int dev_info(const void* kernel, int nt, DevInfo* out) {
  struct Entry {
    int dev;
    const void* kernel;
    DevInfo d;
  };
  static std::mutex mu;
  static std::vector<Entry> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  std::lock_guard<std::mutex> g(mu);
  for (const Entry& e : cache)
    if (e.dev == dev && e.kernel == kernel) {
      *out = e.d;
      return kOK;
    }
  int cus = 0, khz = 0, nb = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return kEHip;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, nt, 0) != hipSuccess) return kEHip;
  const bool rr = ensure_xcd_probe(dev, nullptr);
  Entry e;
  e.dev = dev;
  e.kernel = kernel;
  e.d.cus = cus;
  e.d.blocks_per_cu = std::max(nb, 1);
  e.d.xcd_round_robin = rr;
  e.d.ticks_per_ms = (long long)khz;
  e.d.spin_ticks = (long long)khz * 1000LL * 5LL;   // 5 s
  cache.push_back(e);
  *out = e.d;
  return kOK;
}

// Per-device error word of the calls whose workspace the library allocated
// itself (ws == NULL): before such a workspace is released, the flags its
// launches left are OR-ed into this word (err_fold_kernel, stream ordered),
// and fmm_ws_status(NULL, ...) reads and clears it.
struct StickyErrors {
  std::mutex mu;
  std::vector<int*> word;
};
inline StickyErrors& sticky_errors() {
  static StickyErrors s;
  return s;
}

inline int sticky_word(int** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  StickyErrors& s = sticky_errors();
  std::lock_guard<std::mutex> g(s.mu);
  if ((int)s.word.size() <= dev) s.word.resize(dev + 1, nullptr);
  if (!s.word[dev]) {
    int* p = nullptr;
    if (hipMalloc(&p, 256) != hipSuccess) return kENoMem;
    if (hipMemset(p, 0, 256) != hipSuccess) {
      (void)hipFree(p);
      return kEHip;
    }
    s.word[dev] = p;
  }
  *out = s.word[dev];
  return kOK;
}

// dst |= the error word of workspace ws, on `stream`.
inline int fold_error(int* dst, const int* ws, hipStream_t stream) {
  hipLaunchKernelGGL(err_fold_kernel, dim3(1), dim3(64), 0, stream, dst, ws);
  return hipGetLastError() == hipSuccess ? kOK : kEHip;
}

// Fold the flags of a library-owned workspace into the device's sticky word.
inline int fold_owned(const int* ws, hipStream_t stream) {
  int* w = nullptr;
  int rc = sticky_word(&w);
  return rc == kOK ? fold_error(w, ws, stream) : rc;
}

// Read and clear the sticky word of the current device (synchronises it).
inline int take_sticky(int* flags) {
  *flags = 0;
  int* w = nullptr;
  int rc = sticky_word(&w);
  if (rc != kOK) return rc;
  if (hipDeviceSynchronize() != hipSuccess) return kEHip;
  if (hipMemcpy(flags, w, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return kEHip;
  if (*flags != 0 && hipMemset(w, 0, sizeof(int)) != hipSuccess) return kEHip;
  return kOK;
}

template <int BK, int CH>
inline int dev_info(DevInfo* out) {
  return dev_info(reinterpret_cast<const void*>(&abc_f64_kernel<BK, CH>), NT, out);
}

// --- job construction -------------------------------------------------------
struct JobList {
  std::vector<Job> jobs;
  int total_units = 0;
  int counters = 0;       // order counters needed
  bool ordered = false;
};

inline int tiles(int n, int t) { return (n + t - 1) / t; }

// Geometry of the kernel family that runs a job list: its C tile, the bytes
// of an A/B and of a C element, and the bytes of its A vector access (B and C
// vector accesses are 16 bytes in every family).
struct KGeo {
  int bm, bn;
  int sin, sc;
  int va;
};
constexpr KGeo kGeoF64{BM, BN, 8, 8, 16};

// Vector-access flags of a job (bits 0/1/2: A/B/C).  `kblk` is the row
// offset unit of B blocks (the leaf depth for FMM jobs, 0 for a plain GEMM,
// whose only block starts at B itself).
inline int vec_flags(const KGeo& G, const void* A, long long lda, const void* B, long long ldb, long long kblk,
                     const void* C, long long ldc) {
  int f = 0;
  if (((uintptr_t)A % G.va) == 0 && (lda * G.sin) % G.va == 0) f |= 1;
  if (((uintptr_t)B & 15) == 0 && (ldb * G.sin) % 16 == 0 && (kblk * G.sin) % 16 == 0) f |= 2;
  if (((uintptr_t)C & 15) == 0 && (ldc * G.sc) % 16 == 0) f |= 4;
  return f;
}

inline int vec_flags(const void* A, long long lda, const void* B, long long ldb, int ks_even, const void* C,
                     long long ldc) {
  return vec_flags(kGeoF64, A, lda, B, ldb, ks_even ? 0 : 1, C, ldc);
}

// Element offset of a matrix pointer.
inline const void* eoff(const void* p, long long elems, int size) {
  return p ? static_cast<const char*>(p) + elems * size : nullptr;
}
inline void* eoff(void* p, long long elems, int size) {
  return p ? static_cast<char*>(p) + elems * size : nullptr;
}

inline void add_gemm_job(JobList& L, const KGeo& G, const Prod* dp, const Term* dt, int m, int n, int k,
                         const void* A, long long lda, const void* B, long long ldb, void* C, long long ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  Job j{};
  j.A = A;
  j.B = B;
  j.C = C;
  j.lda = lda;
  j.ldb = ldb;
  j.ldc = ldc;
  j.ms = m;
  j.ns = n;
  j.ks = k;
  j.kr = 0;
  j.tiles_m = tiles(m, G.bm);
  j.tiles_n = tiles(n, G.bn);
  j.nprod = 1;
  j.gsize = 1;
  j.ngroups = 1;
  j.unit_base = L.total_units;
  j.order_base = -1;
  j.flags = vec_flags(G, A, lda, B, ldb, 0, C, ldc);
  j.prods = dp;
  j.terms = dt;
  L.total_units += j.tiles_m * j.tiles_n;
  L.jobs.push_back(j);
}

inline void add_gemm_job(JobList& L, const Prod* dp, const Term* dt, int m, int n, int k, const double* A,
                         long long lda, const double* B, long long ldb, double* C, long long ldc) {
  add_gemm_job(L, kGeoF64, dp, dt, m, n, k, A, lda, B, ldb, C, ldc);
}

// Jobs for one FMM call C += A*B (column-major): the two independent fringe
// GEMMs first (they are thin and long, so they start early), then the FMM
// core with the k-fringe products appended (bl_dynamic_peeling,
// common/dgemm/bl_dgemm_util.c:316-353, reorganised as jobs of one launch).
inline void add_fmm_jobs(JobList& L, const KGeo& G, const Plan& P, const Prod* dp, const Term* dt, const Prod* idp,
                         const Term* idt, int m, int n, int k, const void* A, long long lda, const void* B,
                         long long ldb, void* C, long long ldc) {
  const int md = m - m % (P.Mx * kMR), kd = k - k % P.Kx, nd = n - n % (P.Nx * kNR);
  const int ms = md / P.Mx, ks = kd / P.Kx, ns = nd / P.Nx;
  const int kr = k - kd;
  // (ii) right columns C[:, nd:n] += A * B[:, nd:n]
  add_gemm_job(L, G, idp, idt, m, n - nd, k, A, lda, eoff(B, (long long)nd * ldb, G.sin), ldb,
               eoff(C, (long long)nd * ldc, G.sc), ldc);
  // (iii) bottom rows C[md:m, 0:nd] += A[md:m, :] * B[:, 0:nd]
  if (nd > 0) add_gemm_job(L, G, idp, idt, m - md, nd, k, eoff(A, md, G.sin), lda, B, ldb, eoff(C, md, G.sc), ldc);
  if (ms <= 0 || ns <= 0) return;
  if (ks <= 0) {
    // no FMM core depth: the k-fringe alone is a plain GEMM on the core block
    add_gemm_job(L, G, idp, idt, md, nd, k, A, lda, B, ldb, C, ldc);
    return;
  }
  Job j{};
  j.A = A;
  j.B = B;
  j.C = C;
  j.lda = lda;
  j.ldb = ldb;
  j.ldc = ldc;
  j.ms = ms;
  j.ns = ns;
  j.ks = ks;
  j.kr = kr;
  j.tiles_m = tiles(ms, G.bm);
  j.tiles_n = tiles(ns, G.bn);
  j.nprod = P.R + (kr > 0 ? P.Mx * P.Nx : 0);
  j.gsize = 1;
  j.ngroups = j.nprod;
  j.unit_base = L.total_units;
  j.order_base = j.nprod > 1 ? L.counters : -1;
  j.flags = vec_flags(G, A, lda, B, ldb, ks, C, ldc);
  j.prods = dp;
  j.terms = dt;
  if (j.nprod > 1) {
    L.counters += j.tiles_m * j.tiles_n;
    L.ordered = true;
  }
  L.total_units += j.nprod * j.tiles_m * j.tiles_n;
  L.jobs.push_back(j);
}

inline void add_fmm_jobs(JobList& L, const Plan& P, const Prod* dp, const Term* dt, const Prod* idp,
                         const Term* idt, int m, int n, int k, const double* A, long long lda, const double* B,
                         long long ldb, double* C, long long ldc) {
  add_fmm_jobs(L, kGeoF64, P, dp, dt, idp, idt, m, n, k, A, lda, B, ldb, C, ldc);
}

// Group consecutive products of each FMM job into units so that the launch
// has about kUnitsPerSlot units per resident workgroup: fewer, longer units
// mean fewer cross-workgroup hand-overs (each costs an agent-scope release and
// acquire of C), while enough units remain to balance the grid.  A job whose
// products all fit one unit needs no ordering at all.  Unit bases are
// recomputed; counters stay reserved (workspace size does not depend on the
// device).
#ifndef FMM_UNITS_PER_SLOT
#define FMM_UNITS_PER_SLOT 0
#endif
constexpr int kUnitsPerSlot = FMM_UNITS_PER_SLOT;
inline void group_products(JobList& L, int slots) {
  const long long target = kUnitsPerSlot > 0 ? (long long)kUnitsPerSlot * std::max(slots, 1) : (1LL << 40);
  int base = 0;
  bool ordered = false;
  for (Job& j : L.jobs) {
    const int ntile = j.tiles_m * j.tiles_n;
    if (j.nprod > 1) {
      long long ng = (target + ntile - 1) / ntile;
      ng = std::max(1LL, std::min<long long>(ng, j.nprod));
      j.gsize = (int)((j.nprod + ng - 1) / ng);
      j.ngroups = (j.nprod + j.gsize - 1) / j.gsize;
      if (j.ngroups == 1) j.order_base = -1;
    }
    if (j.order_base >= 0) ordered = true;
    j.unit_base = base;
    base += j.ngroups * ntile;
  }
  L.total_units = base;
  L.ordered = ordered;
}

inline size_t ws_ints(const JobList& L) { return L.ordered ? (size_t)(kWsCounters + L.counters) : 0; }

// Launch a job list (<= kInlineJobs jobs, passed by value) on `stream` with
// the persistent kernel `kern` (nt threads per workgroup).  `ws` must hold
// ws_ints(L) ints when L.ordered.
// Optional timing of product launches (for a benchmark's roofline of the
// kernel itself): when enabled, the launch of each product kernel on the
// calling thread is bracketed by two HIP events; the last pair is kept.
struct LaunchTiming {
  bool on = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
};
inline LaunchTiming& launch_timing() {
  static thread_local LaunchTiming t;
  return t;
}

inline int launch_jobs_with(void (*kern)(Launch), int nt, JobList JL, int* ws, hipStream_t stream) {
  if (JL.total_units == 0) return kOK;
  if ((int)JL.jobs.size() > kInlineJobs) return kEInval;
  DevInfo di;
  int rc = dev_info(reinterpret_cast<const void*>(kern), nt, &di);
  if (rc != kOK) return rc;
  const int slots = di.cus * di.blocks_per_cu;
  const size_t ws_n = ws_ints(JL);   // counters as reserved by the caller
  group_products(JL, slots);
  Launch L{};
  L.njobs = (int)JL.jobs.size();
  L.total_units = JL.total_units;
  L.use_ticket = JL.ordered ? 1 : 0;
  L.jobs = nullptr;
  L.ws = ws;
  L.spin_limit = spin_ticks(di);
  for (size_t i = 0; i < JL.jobs.size(); ++i) L.inl[i] = JL.jobs[i];
  if (JL.ordered) {
    if (!ws) return kEInval;
    const int n = (int)ws_n;
    hipLaunchKernelGGL(ws_clear_kernel, dim3(std::min((n + 255) / 256, 1024)), dim3(256), 0, stream, ws, n);
    if (hipGetLastError() != hipSuccess) return kEHip;
  }
  // One resident workgroup per slot; 8 queues (one per XCD) when the grid
  // fills every XCD with a whole number of workgroups.
  int grid = std::max(1, std::min(JL.total_units, slots));
  L.nq = 1;
  L.local_handover = 0;
  if (grid >= 8 * kMaxQueues) {
    grid -= grid % kMaxQueues;
    L.nq = kMaxQueues;
    L.local_handover = JL.ordered && di.xcd_round_robin && local_handover_allowed() ? 1 : 0;
  }
  LaunchTiming& tm = launch_timing();
  if (tm.on && tm.ev[0]) (void)hipEventRecord(tm.ev[0], stream);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(nt), 0, stream, L);
  if (hipGetLastError() != hipSuccess) return kEHip;
  if (tm.on && tm.ev[1]) (void)hipEventRecord(tm.ev[1], stream);
  return kOK;
}

template <int BK, int CH>
inline int launch_jobs_k(JobList JL, int* ws, hipStream_t stream) {
  return launch_jobs_with(abc_f64_kernel<BK, CH>, NT, std::move(JL), ws, stream);
}

// Kernel variant by the largest operand-sum term count of the launch: the
// single-term kernel (BK = 32, no operand-sum code: plain GEMMs, Naive
// products; 231 instead of 255 VGPRs, +7% at 9600^3), BK = 32 with 2-term
// chunks when every product has at most 2 terms per operand (Strassen,
// <2,3,2>), else BK = 16 with 4-term chunks.
// (Running the > 2-term products on the 32-deep 2-term kernel too, their
// further terms loaded without prefetch, measured 2-5% slower at 9600^3 for
// <3,3,3>, <3,2,3> and <2,2,2>x<3,2,3>.)
// The staggered-round kernel (32-deep slices, operand sums built one term at
// a time during the MFMAs, see contract_rounds) runs a launch when every
// product has at most kRoundTerms (8) terms per operand and every FMM core job
// takes its compile-time-shaped path (16-byte operand access, even leaf row
// count, leaf depth at least 32); the chunked kernels serve the rest.
// FMM_ABC_ROUNDS=0 selects the chunked kernels always (the two agree bit for
// bit).  Measured ABC launches at 9600^3: 333 L1 32.5 -> 26.8 ms, 222 x 323
// 51.5 -> 32.9 ms.
inline bool rounds_enabled() {
  const char* e = getenv("FMM_ABC_ROUNDS");
  return !(e && e[0] == '0');
}

inline bool rounds_fit(const JobList& JL) {
  for (const Job& j : JL.jobs)
    if (j.nprod > 1 && !((j.flags & 3) == 3 && j.ms % 2 == 0 && j.ks >= 32)) return false;
  return true;
}

// FP64 launches of single-term products run the direct-to-LDS kernel
// (fmm_f64d.hpp); FMM_F64_DIRECT=0 in the environment selects the
// register-staged single-term kernel (same results bit for bit).
inline bool f64_direct() {
  static const bool on = [] {
    const char* e = getenv("FMM_F64_DIRECT");
    return !(e && e[0] == '0');
  }();
  return on;
}

inline int launch_jobs(const JobList& JL, int max_terms, int* ws, hipStream_t stream) {
  if (max_terms <= 1 && f64_direct()) return launch_jobs_with(f64d_kernel<CfgF64D>, NT, JL, ws, stream);
  if (max_terms <= 1) return launch_jobs_k<32, 1>(JL, ws, stream);
  if (max_terms <= kRoundTerms && rounds_enabled() && rounds_fit(JL))
    return launch_jobs_k<32, kRounds>(JL, ws, stream);
  if (max_terms <= 2) return launch_jobs_k<32, 2>(JL, ws, stream);
  return launch_jobs_k<16, 4>(JL, ws, stream);
}

// --- element types ------------------------------------------------------------
//
// FP64: A, B, C FP64 (abc_f64_kernel).  FP32 / BF16 / FP16: A and B of that
// type, C FP32, products on FP32 / BF16 / FP16 MFMA with FP32 accumulation.
// BF16x3: A, B, C FP32, split-BF16 (compensated) products.
enum DType : int { kF64 = 0, kF32 = 1, kBF16 = 2, kF16 = 3, kBF16x3 = 4 };

template <class C>
constexpr KGeo lgeo() {
  return KGeo{C::BM, C::BN, C::IN == kInF32 ? 4 : 2, 4, C::AR * (C::IN == kInF32 ? 4 : 2)};
}

template <class C>
constexpr KGeo dgeo() {
  return KGeo{C::BM, C::BN, 2, 4, 16};
}

// The single-term 16-bit launches run the direct-to-LDS kernel
// (fmm_h16d.hpp); FMM_H16_DIRECT=0 in the environment selects the
// register-staged single-term kernel instead (same results bit for bit).
inline bool h16_direct() {
  static const bool on = [] {
    const char* e = getenv("FMM_H16_DIRECT");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Geometry of a dtype's kernel for C-tile shape `tile` (0: the default, 1:
// the transposed 128x256 shape; types without it fall back to the default).
// *flag receives the job flag that selects the shape.  `one`: the jobs run
// in a single-term launch (launch_jobs_dtype with one = true), whose
// low-precision kernels have C tiles of their own.
inline bool dtype_geo(int dtype, KGeo* g, int tile = 0, int* flag = nullptr, bool one = false) {
  int f = 0;
  const bool o = one && tile == 0;
  switch (dtype) {
    case kF64: *g = kGeoF64; break;
    case kF32: *g = o ? lgeo<CfgF32One>() : tile ? lgeo<CfgF32T>() : lgeo<CfgF32>(); f = tile ? kJobTileT : 0; break;
    case kBF16:
      *g = o ? (h16_direct() ? dgeo<CfgBF16D>() : lgeo<CfgBF16One>()) : tile ? lgeo<CfgBF16T>() : lgeo<CfgBF16>();
      f = tile ? kJobTileT : 0;
      break;
    case kF16:
      *g = o ? (h16_direct() ? dgeo<CfgF16D>() : lgeo<CfgF16One>()) : tile ? lgeo<CfgF16T>() : lgeo<CfgF16>();
      f = tile ? kJobTileT : 0;
      break;
    case kBF16x3: *g = lgeo<CfgBF16x3>(); break;
    default: return false;
  }
  if (flag) *flag = f;
  return true;
}

template <class C1, class C2>
inline int launch_lowp(const JobList& JL, int* ws, hipStream_t stream) {
  return launch_jobs_with(abc_lowp_kernel<C1, C2>, C1::NT, JL, ws, stream);
}

// Kernel of a dtype (FP64: by the largest operand-sum term count).  A launch
// with jobs of both C-tile shapes runs the two-shape kernel; otherwise the
// one-shape kernel of its shape (fewer live registers).
template <class C1, class C2>
inline int launch_lowp_shapes(const JobList& JL, int* ws, hipStream_t stream) {
  bool t = false, d = false;
  for (const Job& j : JL.jobs) ((j.flags & kJobTileT) ? t : d) = true;
  if (t && d) return launch_lowp<C1, C2>(JL, ws, stream);
  if (t) return launch_lowp<C2, C2>(JL, ws, stream);
  return launch_lowp<C1, C1>(JL, ws, stream);
}

// `one`: every product of the launch has single +-1 operand terms (plain
// GEMMs, Naive products): the low-precision types then run their
// single-term kernels (deeper slices, no operand-sum code).
inline int launch_jobs_dtype(int dtype, const JobList& JL, int max_terms, int* ws, hipStream_t stream,
                             bool one = false) {
  bool t = false;
  for (const Job& j : JL.jobs) t = t || (j.flags & kJobTileT) != 0;
  const bool use_one = one && !t;
  switch (dtype) {
    case kF64: return launch_jobs(JL, max_terms, ws, stream);
    case kF32:
      return use_one ? launch_lowp<CfgF32One, CfgF32One>(JL, ws, stream)
                     : launch_lowp_shapes<CfgF32, CfgF32T>(JL, ws, stream);
    case kBF16:
      if (use_one && h16_direct()) return launch_jobs_with(h16d_kernel<CfgBF16D>, CfgBF16D::NT, JL, ws, stream);
      return use_one ? launch_lowp<CfgBF16One, CfgBF16One>(JL, ws, stream)
                     : launch_lowp_shapes<CfgBF16, CfgBF16T>(JL, ws, stream);
    case kF16:
      if (use_one && h16_direct()) return launch_jobs_with(h16d_kernel<CfgF16D>, CfgF16D::NT, JL, ws, stream);
      return use_one ? launch_lowp<CfgF16One, CfgF16One>(JL, ws, stream)
                     : launch_lowp_shapes<CfgF16, CfgF16T>(JL, ws, stream);
    case kBF16x3: return launch_lowp<CfgBF16x3, CfgBF16x3>(JL, ws, stream);
    default: return kEUnsupported;
  }
}

// --- C-stationary launches (FP64, one C tile of every C block on chip) -------
//
// The FP64 Naive pack type runs its products C-stationary (fmm_cs.hpp) when
// the algorithm has at most 9 C blocks and the leaf is shallow: measured at
// m = n = 14400 (k = 480, leaf depth 160-240) it beats the ordered product
// launch + M_Add by 12-35% (<2,2,2>, <2,3,2>, <3,3,3>), from k = 960 (leaf
// depth 320-480) the ordered launch is faster.  With operand sums (ABC) the
// C-stationary kernel measured slower than the ordered one at every depth,
// so ABC keeps the ordered kernel.  FMM_CS=0 / 1 in the environment forces
// the ordered / the C-stationary product launch for Naive (when the C
// blocks fit).
constexpr int kCsMaxBlocks = 9;
inline bool cs_fits(const Plan& P) { return P.Mx * P.Nx <= kCsMaxBlocks; }
#ifndef FMM_CS_MAX_DEPTH
#define FMM_CS_MAX_DEPTH 256
#endif
inline bool use_cs(const Plan& P, int dtype, int ks) {
  if (dtype != kF64 || !cs_fits(P) || ks <= 0) return false;
  const char* e = getenv("FMM_CS");
  if (e && *e) return atoi(e) != 0;
  return ks <= FMM_CS_MAX_DEPTH;
}

template <class Cfg>
inline int launch_cs_k(CsJob J, hipStream_t stream) {
  DevInfo di;
  int rc = dev_info(reinterpret_cast<const void*>(&cs_f64_kernel<Cfg>), Cfg::NT, &di);
  if (rc != kOK) return rc;
  J.tiles_m = tiles(J.ms, Cfg::TM);
  J.tiles_n = tiles(J.ns, Cfg::TN);
  const int units = J.tiles_m * J.tiles_n;
  if (units <= 0 || J.nprod <= 0) return kOK;
  int grid = std::max(1, std::min(units, di.cus * di.blocks_per_cu));
  int nq = 1;
  if (grid >= 8 * kMaxQueues) {
    grid -= grid % kMaxQueues;
    nq = kMaxQueues;
  }
  LaunchTiming& tm = launch_timing();
  if (tm.on && tm.ev[0]) (void)hipEventRecord(tm.ev[0], stream);
  hipLaunchKernelGGL(cs_f64_kernel<Cfg>, dim3(grid), dim3(Cfg::NT), 0, stream, J, nq);
  if (hipGetLastError() != hipSuccess) return kEHip;
  if (tm.on && tm.ev[1]) (void)hipEventRecord(tm.ev[1], stream);
  return kOK;
}

// Single-term products (Naive over S / T, or single-term A / B blocks) by
// C-block count: <= 4 blocks 128 x 64 positions, <= 9 blocks 64 x 64.
template <class Cfg>
inline int launch_csd_k(CsJob J, hipStream_t stream) {
  DevInfo di;
  int rc = dev_info(reinterpret_cast<const void*>(&csd_f64_kernel<Cfg>), Cfg::NT, &di);
  if (rc != kOK) return rc;
  J.tiles_m = tiles(J.ms, Cfg::TM);
  J.tiles_n = tiles(J.ns, Cfg::TN);
  const int units = J.tiles_m * J.tiles_n;
  if (units <= 0 || J.nprod <= 0) return kOK;
  int grid = std::max(1, std::min(units, di.cus * di.blocks_per_cu));
  int nq = 1;
  if (grid >= 8 * kMaxQueues) {
    grid -= grid % kMaxQueues;
    nq = kMaxQueues;
  }
  LaunchTiming& tm = launch_timing();
  if (tm.on && tm.ev[0]) (void)hipEventRecord(tm.ev[0], stream);
  hipLaunchKernelGGL(csd_f64_kernel<Cfg>, dim3(grid), dim3(Cfg::NT), 0, stream, J, nq);
  if (hipGetLastError() != hipSuccess) return kEHip;
  if (tm.on && tm.ev[1]) (void)hipEventRecord(tm.ev[1], stream);
  return kOK;
}

// FMM_CSD=0 selects the register-staged C-stationary kernel (read per call).
inline bool csd_enabled() {
  const char* e = getenv("FMM_CSD");
  return !(e && e[0] == '0');
}

#ifndef FMM_CSD4_CFG
#define FMM_CSD4_CFG 4, 32, 2
#endif
#ifndef FMM_CSD9_CFG
#define FMM_CSD9_CFG 9, 32, 2
#endif

inline int launch_cs(const CsJob& J, hipStream_t stream) {
  const int ncb = J.mx * J.nx;
  if (ncb > kCsMaxBlocks) return kEInval;
  if (csd_enabled() && J.nprod <= kCsdMaxProd && J.nterms <= kCsdMaxTerms) {
    if (ncb <= 4) return launch_csd_k<CsdCfg<FMM_CSD4_CFG>>(J, stream);
    return launch_csd_k<CsdCfg<FMM_CSD9_CFG>>(J, stream);
  }
  if (ncb <= 4) return launch_cs_k<CsCfg<4, 32, 1>>(J, stream);
  return launch_cs_k<CsCfg<9, 32, 1>>(J, stream);
}

// The C-stationary core job of an FMM call (its geometry from the partition).
inline CsJob cs_core_job(const Plan& P, const Prod* dp, const Term* dt, int m, int n, int k, const void* A,
                         long long lda, const void* B, long long ldb, void* C, long long ldc) {
  CsJob j{};
  const int md = m - m % (P.Mx * kMR), kd = k - k % P.Kx, nd = n - n % (P.Nx * kNR);
  j.A = static_cast<const double*>(A);
  j.B = static_cast<const double*>(B);
  j.C = static_cast<double*>(C);
  j.lda = lda;
  j.ldb = ldb;
  j.ldc = ldc;
  j.ms = md / P.Mx;
  j.ns = nd / P.Nx;
  j.ks = kd / P.Kx;
  j.kr = k - kd;
  j.nprod = P.R + (j.kr > 0 ? P.Mx * P.Nx : 0);
  j.nmain = P.R;
  j.mx = P.Mx;
  j.nx = P.Nx;
  j.vec = vec_flags(kGeoF64, A, lda, B, ldb, j.ks, C, ldc);
  j.prods = dp;
  j.terms = dt;
  return j;
}

// --- AB and Naive pack types ---------------------------------------------
//
// Pack types of the reference (meta/ab_gen_mix.py, meta/naive_gen_mix.py):
//   ABC    operand sums formed while packing, every product updates its C
//          blocks directly (ordered per tile);
//   AB     operand sums formed while packing, product r stored to M_r, then
//          C_j += w_j M_r (M_Add);
//   Naive  S_r and T_r summed into workspaces first, M_r = S_r T_r by the
//          plain GEMM path, then M_Add.
// AB and Naive need no ordering inside the product launch (every product owns
// its M_r); the price is the M (and S, T) traffic through HBM.  All three
// apply the same C updates in the same order with the same arithmetic, so
// they agree bit for bit.
enum Pack : int { kPackABC = 0, kPackAB = 1, kPackNaive = 2 };

struct PackLayout {
  int md = 0, kd = 0, nd = 0, ms = 0, ks = 0, ns = 0, kr = 0;
  int nlive = 0;                       // AB: M blocks written (R [+ Mx*Nx])
  long long ldm = 0, lds = 0, ldt = 0; // leading dimensions of M, S, T
  size_t m_off = 0, s_off = 0, t_off = 0, o_off = 0, bytes = 0;
  int o_ints = 0;                      // ints of the status header at offset 0: the
                                       // ordered launch's ws (fused Naive), else
                                       // the error word area alone
};

inline size_t round256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace layout of an AB / Naive call; false when the shape has no FMM
// core (then the call is the ABC path's plain GEMMs).
// Element sizes: sin of A / B (and of the S, T sums), sc of C (and of M).
// The default (FP64) layout is the largest, so its size bounds every dtype's.
// Naive with its C updates fused into the product launch (the low-precision
// types, where it measured 5-20% faster than M + M_Add), or M + M_Add (FP64,
// 2% faster: the FP64 product launch idles its CU for each C read-modify-write
// of a unit, M_Add streams).
inline bool naive_fused(int dtype) { return dtype != kF64; }

// cs: FP64 Naive with a C-stationary product launch (use_cs): S and T only.
inline bool pack_layout(const Plan& P, int pack, int m, int n, int k, PackLayout* L, int sin = 8, int sc = 8,
                        bool fused = false, bool cs = false) {
  PackLayout o;
  o.md = m - m % (P.Mx * kMR);
  o.kd = k - k % P.Kx;
  o.nd = n - n % (P.Nx * kNR);
  o.ms = o.md / P.Mx;
  o.ks = o.kd / P.Kx;
  o.ns = o.nd / P.Nx;
  o.kr = k - o.kd;
  if (o.ms <= 0 || o.ns <= 0 || o.ks <= 0) return false;
  o.nlive = P.R + (o.kr > 0 ? P.Mx * P.Nx : 0);
  o.ldm = o.ms;                        // ms % 8 == 0: 16-byte columns
  o.lds = o.ms;
  o.ldt = (o.ks + 7) & ~7;             // T columns stay 16-byte aligned for every type
  // The workspace starts with a status header whose error word sits where
  // an ABC workspace has it (fmm_ws_status reads it for every pack type):
  // for the fused Naive launch the whole ordered workspace (order counters,
  // one per C tile of the smallest tile shape, 128 x 128, of any type), else
  // the error word area alone, cleared by the call.
  o.o_off = 0;
  o.o_ints = (pack == kPackNaive && fused) ? kWsCounters + tiles(o.ms, 128) * tiles(o.ns, 128) : kWsCounters;
  size_t off = round256(sizeof(int) * (size_t)o.o_ints);
  if (pack == kPackNaive) {
    o.s_off = off;
    off += round256((size_t)sin * (size_t)o.lds * o.ks * P.R);
    o.t_off = off;
    off += round256((size_t)sin * (size_t)o.ldt * o.ns * P.R);
  }
  if (!(pack == kPackNaive && (fused || cs))) {
    o.m_off = off;
    off += round256((size_t)sc * (size_t)o.ldm * o.ns * o.nlive);
  }
  o.bytes = off;
  *L = o;
  return true;
}

inline int ew_grid(int rows, int cols) { return tiles(rows, kEwRows) * tiles(cols, kEwCols); }

// FMM_PRESUM_ALL=0 in the environment runs the FP64 pre-sums one product
// per grid row (presum_kernel) instead of one pass per operand
// (presum_all_kernel); the sums are the same bit for bit.
inline bool presum_all_enabled() {
  static const bool on = [] {
    const char* e = getenv("FMM_PRESUM_ALL");
    return !(e && e[0] == '0');
  }();
  return on;
}

// Naive pre-sums S_r = sum_t u_t A_t and T_r = sum_t v_t B_t into the
// workspace blocks (0, r) of S and T, with the ABC kernel's packing arithmetic.
inline int launch_presums(const Plan* P, int dtype, const PackLayout& Lo, const Prod* dp, const Term* dt,
                          const void* A, long long lda, const void* B, long long ldb, void* S, void* T,
                          hipStream_t stream, bool skip_single = false) {
  const dim3 ga(ew_grid(Lo.ms, Lo.ks), P->R), gb(ew_grid(Lo.ks, Lo.ns), P->R);
  if (dtype == kF64) {
    SumJob sa{};
    sa.X = static_cast<const double*>(A);
    sa.Y = static_cast<double*>(S);
    sa.ldx = lda;
    sa.ldy = Lo.lds;
    sa.rows = Lo.ms;
    sa.cols = Lo.ks;
    sa.ny = 1;
    sa.side = 0;
    sa.vec = ((uintptr_t)A & 15) == 0 && (lda & 1) == 0 && ((uintptr_t)S & 15) == 0;
    sa.skip1 = skip_single ? 1 : 0;
    sa.prods = dp;
    sa.terms = dt;
    SumJob sb = sa;
    sb.X = static_cast<const double*>(B);
    sb.Y = static_cast<double*>(T);
    sb.ldx = ldb;
    sb.ldy = Lo.ldt;
    sb.rows = Lo.ks;
    sb.cols = Lo.ns;
    sb.side = 1;
    sb.vec = ((uintptr_t)B & 15) == 0 && (ldb & 1) == 0 && (Lo.ks & 1) == 0 && ((uintptr_t)T & 15) == 0;
    // one pass over each operand block for all products when its blocks fit
    // the LDS of a workgroup; else one pass per product
    for (int side = 0; side < 2; ++side) {
      const SumJob& sj = side ? sb : sa;
      const int nbr = side ? P->Kx : P->Mx, nbc = side ? P->Nx : P->Kx;
      if (presum_all_enabled() && nbr * nbc <= kSaMaxBlocks) {
        SumAllJob q{};
        q.X = sj.X;
        q.Y = sj.Y;
        q.ldx = sj.ldx;
        q.ldy = sj.ldy;
        q.rows = sj.rows;
        q.cols = sj.cols;
        q.ny = sj.ny;
        q.side = sj.side;
        q.vec = sj.vec;
        q.skip1 = sj.skip1;
        q.nbr = nbr;
        q.nbc = nbc;
        q.nprod = P->R;
        q.prods = sj.prods;
        q.terms = sj.terms;
        const dim3 g(tiles(q.rows, kEwRows) * q.cols);
        const size_t lds = (size_t)nbr * nbc * 64 * sizeof(f64x2);
        static const hipError_t attr = hipFuncSetAttribute(
            reinterpret_cast<const void*>(presum_all_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
            kSaMaxBlocks * 64 * (int)sizeof(f64x2));
        if (attr != hipSuccess) return kEHip;
        hipLaunchKernelGGL(presum_all_kernel, g, dim3(64 * kSaWaves), lds, stream, q);
      } else {
        presum_kernel<<<side ? gb : ga, dim3(kEwThreads), 0, stream>>>(sj);
      }
      if (hipGetLastError() != hipSuccess) return kEHip;
    }
  } else {
    SumJobL sa{};
    sa.X = A;
    sa.Y = S;
    sa.ldx = lda;
    sa.ldy = Lo.lds;
    sa.rows = Lo.ms;
    sa.cols = Lo.ks;
    sa.ny = 1;
    sa.side = 0;
    sa.prods = dp;
    sa.terms = dt;
    SumJobL sb = sa;
    sb.X = B;
    sb.Y = T;
    sb.ldx = ldb;
    sb.ldy = Lo.ldt;
    sb.rows = Lo.ks;
    sb.cols = Lo.ns;
    sb.side = 1;
    // paired loads: two consecutive elements of a column in one access
    const int es = (dtype == kBF16 || dtype == kF16) ? 2 : 4;
    auto pairs = [&](const void* X, long long ldx, int rows) {
      return ((uintptr_t)X % (2 * es)) == 0 && (ldx & 1) == 0 && (rows & 1) == 0 ? 1 : 0;
    };
    sa.vec = pairs(A, lda, Lo.ms);
    sb.vec = pairs(B, ldb, Lo.ks);
    auto two = [&](auto k0, auto k1) {
      // k0: element-wise loads, k1: paired loads
      if (sa.vec) hipLaunchKernelGGL(k1, ga, dim3(kEwThreads), 0, stream, sa);
      else hipLaunchKernelGGL(k0, ga, dim3(kEwThreads), 0, stream, sa);
      if (sb.vec) hipLaunchKernelGGL(k1, gb, dim3(kEwThreads), 0, stream, sb);
      else hipLaunchKernelGGL(k0, gb, dim3(kEwThreads), 0, stream, sb);
    };
    if (dtype == kBF16) {
      two(presum_lowp_kernel<kInBF16, false, false>, presum_lowp_kernel<kInBF16, false, true>);
    } else if (dtype == kF16) {
      two(presum_lowp_kernel<kInF16, false, false>, presum_lowp_kernel<kInF16, false, true>);
    } else if (dtype == kBF16x3 && h16_direct()) {
      // split planes for the direct kernel: hi at S (T), lo one plane later
      sa.plane = (long long)Lo.lds * Lo.ks * P->R;
      sb.plane = (long long)Lo.ldt * Lo.ns * P->R;
      two(presum_lowp_kernel<kInF32, true, false>, presum_lowp_kernel<kInF32, true, true>);
    } else {
      two(presum_lowp_kernel<kInF32, false, false>, presum_lowp_kernel<kInF32, false, true>);
    }
  }
  if (hipGetLastError() != hipSuccess) return kEHip;
  return kOK;
}

// One fused Naive call C += A*B on `stream`: the pre-sums, then one product launch
// M_r = S_r T_r whose epilogue applies C_j = fma(w_jr, M_r, C_j) for every
// C block of product r straight from the accumulators, in product order per
// C tile (the ordered hand-over of the ABC kernel), with the right and bottom
// fringe GEMMs as further jobs; then, when k has a fringe, the Mx*Nx k-fringe
// products of `main` in a launch of their own.  The reference stores M_r and
// adds it in a pass of its own (meta/naive_gen_mix.py:60-137); the C updates
// and their order are the same, so the result equals ABC's bit for bit.
inline int launch_naive_fused(Plan* P, int dtype, const PackLayout& Lo, int m, int n, int k, const void* A,
                              long long lda, const void* B, long long ldb, void* C, long long ldc, void* ws,
                              hipStream_t stream) {
  // Split-BF16 with the direct kernel: the pre-sums are stored as hi / lo
  // BF16 planes, which only the product jobs read; the fringe jobs (the
  // caller's FP32 operands) run in the second launch with the
  // register-staged split kernel (geometry GF).
  const bool sd = dtype == kBF16x3 && h16_direct();
  KGeo GF;
  if (!dtype_geo(dtype, &GF, 0, nullptr, true)) return kEUnsupported;
  const KGeo G = sd ? dgeo<CfgSplitD>() : GF;
  Prod *dp, *np, *ip;
  Term *dt, *nt, *it;
  int rc = device_tables(P, P->main, &dp, &dt);
  if (rc == kOK) rc = device_tables(P, P->naive, &np, &nt);
  if (rc == kOK) rc = device_tables(identity_plan(), &ip, &it);
  if (rc != kOK) return rc;
  char* w = static_cast<char*>(ws);
  void* S = w + Lo.s_off;
  void* T = w + Lo.t_off;
  rc = launch_presums(P, dtype, Lo, dp, dt, A, lda, B, ldb, S, T, stream);
  if (rc != kOK) return rc;
  JobList JL;   // the product launch
  JobList JF;   // the launch after it: the k-fringe products (and, split, the fringe GEMMs)
  JobList& JG = sd ? JF : JL;
  add_gemm_job(JG, GF, ip, it, m, n - Lo.nd, k, A, lda, eoff(B, (long long)Lo.nd * ldb, GF.sin), ldb,
               eoff(C, (long long)Lo.nd * ldc, GF.sc), ldc);
  add_gemm_job(JG, GF, ip, it, m - Lo.md, Lo.nd, k, eoff(A, Lo.md, GF.sin), lda, B, ldb, eoff(C, Lo.md, GF.sc),
               ldc);
  Job j{};
  j.A = S;
  j.B = T;
  j.C = C;
  j.lda = Lo.lds;
  j.ldb = Lo.ldt;
  j.ldc = ldc;
  j.ms = Lo.ms;
  j.ns = Lo.ns;
  j.ks = Lo.ks;
  j.kr = 0;
  j.tiles_m = tiles(Lo.ms, G.bm);
  j.tiles_n = tiles(Lo.ns, G.bn);
  j.nprod = P->R;
  j.gsize = 1;
  j.ngroups = P->R;
  j.unit_base = JL.total_units;
  j.order_base = P->R > 1 ? JL.counters : -1;
  j.flags = vec_flags(G, S, Lo.lds, T, Lo.ldt, 0, C, ldc);   // T_r is block (0, r): no row offset
  j.prods = np;
  j.terms = nt;
  if (sd) {
    j.alo = (long long)Lo.lds * Lo.ks * P->R;
    j.blo = (long long)Lo.ldt * Lo.ns * P->R;
  }
  if (P->R > 1) {
    JL.counters += j.tiles_m * j.tiles_n;
    JL.ordered = true;
  }
  JL.total_units += j.nprod * j.tiles_m * j.tiles_n;
  JL.jobs.push_back(j);
  if (ws_ints(JL) > (size_t)Lo.o_ints) return kEInval;   // the layout reserves 128 x 128 tiles
  int* wo = reinterpret_cast<int*>(w + Lo.o_off);
  rc = sd ? launch_jobs_with(h16d_kernel<CfgSplitD>, CfgSplitD::NT, JL, wo, stream)
          : launch_jobs_dtype(dtype, JL, 1, wo, stream, true);
  if (rc != kOK) return rc;
  if (Lo.kr > 0) {
    // C(I,J) += A(I, Kx) B(Kx, J) over depth kr: one product per C block, so
    // the launch needs no ordering.
    Job f = j;
    f.A = A;
    f.B = B;
    f.lda = lda;
    f.ldb = ldb;
    f.kr = Lo.kr;
    f.tiles_m = tiles(Lo.ms, GF.bm);
    f.tiles_n = tiles(Lo.ns, GF.bn);
    f.nprod = P->Mx * P->Nx;
    f.ngroups = f.nprod;
    f.unit_base = JF.total_units;
    f.order_base = -1;
    f.flags = vec_flags(GF, A, lda, B, ldb, Lo.ks, C, ldc);
    f.prods = dp + P->R;
    f.terms = dt;
    f.alo = f.blo = 0;
    JF.total_units += f.nprod * f.tiles_m * f.tiles_n;
    JF.jobs.push_back(f);
  }
  return launch_jobs_dtype(dtype, JF, 1, nullptr, stream, true);
}

// One FP64 Naive call with a C-stationary product launch: the pre-sums of
// the operands with more than one term into S, T; the right and bottom
// fringe GEMMs; then every product (over S_r / T_r, or A / B blocks for
// single-term operands) and the k-fringe products per position, C updated in
// registers in product order and written once -- the C updates of ABC in the
// same order, so the result equals ABC's bit for bit.
inline int launch_naive_cs(Plan* P, const PackLayout& Lo, int m, int n, int k, const void* A, long long lda,
                           const void* B, long long ldb, void* C, long long ldc, void* ws, hipStream_t stream) {
  Prod *dp, *cp, *ip;
  Term *dt, *ct, *it;
  int rc = device_tables(P, P->main, &dp, &dt);
  if (rc == kOK) rc = device_tables(P, P->naive_cs, &cp, &ct);
  if (rc == kOK) rc = device_tables(identity_plan(), &ip, &it);
  if (rc != kOK) return rc;
  char* w = static_cast<char*>(ws);
  hipLaunchKernelGGL(ws_clear_kernel, dim3(1), dim3(256), 0, stream, reinterpret_cast<int*>(w + Lo.o_off), Lo.o_ints);
  if (hipGetLastError() != hipSuccess) return kEHip;
  void* S = w + Lo.s_off;
  void* T = w + Lo.t_off;
  rc = launch_presums(P, kF64, Lo, dp, dt, A, lda, B, ldb, S, T, stream, true);
  if (rc != kOK) return rc;
  JobList JL;
  add_gemm_job(JL, kGeoF64, ip, it, m, n - Lo.nd, k, A, lda, eoff(B, (long long)Lo.nd * ldb, 8), ldb,
               eoff(C, (long long)Lo.nd * ldc, 8), ldc);
  add_gemm_job(JL, kGeoF64, ip, it, m - Lo.md, Lo.nd, k, eoff(A, Lo.md, 8), lda, B, ldb, eoff(C, Lo.md, 8), ldc);
  if (JL.total_units > 0) {
    LaunchTiming& tm = launch_timing();
    const bool on = tm.on;
    tm.on = false;   // time the product launch below, not the fringe GEMMs
    rc = launch_jobs_dtype(kF64, JL, 1, nullptr, stream, true);
    tm.on = on;
    if (rc != kOK) return rc;
  }
  CsJob J = cs_core_job(*P, cp, ct, m, n, k, A, lda, B, ldb, C, ldc);
  J.S = static_cast<const double*>(S);
  J.T = static_cast<const double*>(T);
  J.lds = Lo.lds;
  J.ldt = Lo.ldt;
  J.nterms = (int)P->naive_cs.terms.h.size();
  return launch_cs(J, stream);
}

// One AB or (M-writing) Naive call C += A*B on `stream`: [pre-sums,] one
// product launch (the products into M -- AB with its operand sums formed
// while packing, Naive over S, T --, the k-fringe products into M, the right
// and bottom fringe GEMMs straight into C), then M_Add into the C core.  Lo
// must be the layout of this dtype (pack_layout with its element sizes and
// naive_fused).
inline int launch_pack(Plan* P, int pack, int dtype, const PackLayout& Lo, int m, int n, int k, const void* A,
                       long long lda, const void* B, long long ldb, void* C, long long ldc, void* ws,
                       hipStream_t stream) {
  const bool naive = pack == kPackNaive;
  if (naive && naive_fused(dtype))
    return launch_naive_fused(P, dtype, Lo, m, n, k, A, lda, B, ldb, C, ldc, ws, stream);
  if (naive && use_cs(*P, dtype, Lo.ks))
    return launch_naive_cs(P, Lo, m, n, k, A, lda, B, ldb, C, ldc, ws, stream);
  KGeo G;
  if (!dtype_geo(dtype, &G, 0, nullptr, naive)) return kEUnsupported;
  Prod *dp, *xp, *ip;
  Term *dt, *xt, *it;
  int rc = device_tables(P, P->main, &dp, &dt);
  if (rc == kOK) rc = device_tables(P, naive ? P->diag : P->ab, &xp, &xt);
  if (rc == kOK) rc = device_tables(identity_plan(), &ip, &it);
  int *cptr, *cprod;
  double* cw;
  if (rc == kOK) rc = device_clists(P, &cptr, &cprod, &cw);
  if (rc != kOK) return rc;
  char* w = static_cast<char*>(ws);
  // no ordered launch in this call: its status header only needs a clear error word
  hipLaunchKernelGGL(ws_clear_kernel, dim3(1), dim3(256), 0, stream, reinterpret_cast<int*>(w + Lo.o_off),
                     Lo.o_ints);
  if (hipGetLastError() != hipSuccess) return kEHip;
  void* M = w + Lo.m_off;
  const void* pa = A;
  const void* pb = B;
  long long la = lda, lb = ldb, kblk = Lo.ks;
  // FP64 Naive: products with a single A (B) term read that block of A (B)
  // directly (scaled by its coefficient in the kernel, as the pre-sum would
  // scale it), so their pre-sums are skipped; one job per combination.
  const bool split = naive && dtype == kF64;
  if (naive) {
    void* S = w + Lo.s_off;
    void* T = w + Lo.t_off;
    rc = launch_presums(P, dtype, Lo, dp, dt, A, lda, B, ldb, S, T, stream, split);
    if (rc != kOK) return rc;
    pa = S;
    la = Lo.lds;
    pb = T;
    lb = Lo.ldt;
    kblk = 0;   // T_r is block (0, r): no row offset
  }
  JobList JL;
  add_gemm_job(JL, G, ip, it, m, n - Lo.nd, k, A, lda, eoff(B, (long long)Lo.nd * ldb, G.sin), ldb,
               eoff(C, (long long)Lo.nd * ldc, G.sc), ldc);
  add_gemm_job(JL, G, ip, it, m - Lo.md, Lo.nd, k, eoff(A, Lo.md, G.sin), lda, B, ldb, eoff(C, Lo.md, G.sc), ldc);
  auto add_m_job = [&](const void* a, long long ja, const void* b, long long jb, long long kb, int kr, int nprod,
                       const Prod* prods, const Term* terms) {
    Job j{};
    j.A = a;
    j.B = b;
    j.C = M;
    j.lda = ja;
    j.ldb = jb;
    j.ldc = Lo.ldm;
    j.ms = Lo.ms;
    j.ns = Lo.ns;
    j.ks = Lo.ks;
    j.kr = kr;
    j.tiles_m = tiles(Lo.ms, G.bm);
    j.tiles_n = tiles(Lo.ns, G.bn);
    j.nprod = nprod;
    j.gsize = 1;
    j.ngroups = nprod;
    j.unit_base = JL.total_units;
    j.order_base = -1;
    j.flags = vec_flags(G, a, ja, b, jb, kb, M, Lo.ldm) | kJobStore;
    j.prods = prods;
    j.terms = terms;
    JL.total_units += nprod * j.tiles_m * j.tiles_n;
    JL.jobs.push_back(j);
  };
  if (split) {
    for (int c = 0; c < 4; ++c) {
      const int n = (int)P->diag_by[c].prods.h.size();
      if (n == 0) continue;
      Prod* cp;
      Term* ct;
      rc = device_tables(P, P->diag_by[c], &cp, &ct);
      if (rc != kOK) return rc;
      const bool da = (c & 1) != 0, db = (c & 2) != 0;
      add_m_job(da ? A : pa, da ? lda : la, db ? B : pb, db ? ldb : lb, db ? (long long)Lo.ks : 0, 0, n, cp, ct);
    }
  } else {
    add_m_job(pa, la, pb, lb, kblk, 0, P->R, xp, xt);
  }
  if (Lo.kr > 0) add_m_job(A, lda, B, ldb, Lo.ks, Lo.kr, P->Mx * P->Nx, xp + P->R, xt);
  rc = launch_jobs_dtype(dtype, JL, naive ? 1 : P->max_nnz, nullptr, stream, naive);
  if (rc != kOK) return rc;
  const dim3 gm(ew_grid(Lo.ms, Lo.ns), P->Mx * P->Nx);
  const int ncb = P->Mx * P->Nx;
  if (dtype == kF64 && madd_all_enabled() && ncb <= kMaMaxBlocks && Lo.nlive > 0 && Lo.ms >= 2) {
    int *rptr, *rblk;
    double* rw;
    rc = device_rlists(P, &rptr, &rblk, &rw);
    if (rc != kOK) return rc;
    AddAllJob ad{};
    ad.M = static_cast<const double*>(M);
    ad.C = static_cast<double*>(C);
    ad.ldm = Lo.ldm;
    ad.ldc = ldc;
    ad.rows = Lo.ms;
    ad.cols = Lo.ns;
    ad.mx = 1;
    ad.nx = P->Nx;
    ad.ncb = ncb;
    ad.nlive = Lo.nlive;
    ad.vec = ((uintptr_t)C & 15) == 0 && (ldc & 1) == 0 && ((uintptr_t)M & 15) == 0;
    ad.rptr = rptr;
    ad.rblk = rblk;
    ad.rw = rw;
    const size_t lds = (size_t)ncb * 1024;
    // Few C blocks: many workgroups per CU, 16 M runs in flight each; many
    // C blocks (LDS-bound occupancy): deeper prefetch, registers to spare.
    auto go = [&](auto kern) {
      static_assert(sizeof(kern) > 0, "");
      if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                              kMaMaxBlocks * 1024) != hipSuccess)
        return kEHip;
      const long long nwg = (long long)tiles(Lo.ms, 128) * Lo.ns;
      hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(64), lds, stream, ad);
      return kOK;
    };
    const bool hi = ncb > FMM_MADD_PF_SPLIT;
    if (hi && madd_all4_enabled()) {
      // many C blocks: the 4-wave kernel, each wave a quarter of the blocks
      int *wbeg, *wprod, *wptr, *wblk;
      double* ww;
      rc = device_wlists(P, &wbeg, &wprod, &wptr, &wblk, &ww);
      if (rc != kOK) return rc;
      AddAll4Job a4{};
      a4.M = ad.M;
      a4.C = ad.C;
      a4.ldm = ad.ldm;
      a4.ldc = ad.ldc;
      a4.rows = ad.rows;
      a4.cols = ad.cols;
      a4.mx = ad.mx;
      a4.nx = ad.nx;
      a4.ncb = ad.ncb;
      a4.nlive = ad.nlive;
      a4.wbeg = wbeg;
      a4.wprod = wprod;
      a4.wptr = wptr;
      a4.wblk = wblk;
      a4.ww = ww;
      auto go4 = [&](auto kern) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(kern), hipFuncAttributeMaxDynamicSharedMemorySize,
                                kMaMaxBlocks * 1024) != hipSuccess)
          return kEHip;
        const long long nwg = (long long)tiles(Lo.ms, 128) * Lo.ns;
        hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(64 * kMa4Waves), lds, stream, a4);
        return kOK;
      };
      rc = ad.vec ? go4(madd_all4_kernel<FMM_MADD_PF_LO, true>) : go4(madd_all4_kernel<FMM_MADD_PF_LO, false>);
    } else if (ad.vec) {
      rc = hi ? go(madd_all_kernel<FMM_MADD_PF_HI, true>) : go(madd_all_kernel<FMM_MADD_PF_LO, true>);
    } else {
      rc = hi ? go(madd_all_kernel<FMM_MADD_PF_HI, false>) : go(madd_all_kernel<FMM_MADD_PF_LO, false>);
    }
    if (rc != kOK) return rc;
  } else if (dtype == kF64) {
    AddJob ad{};
    ad.M = static_cast<const double*>(M);
    ad.C = static_cast<double*>(C);
    ad.ldm = Lo.ldm;
    ad.ldc = ldc;
    ad.rows = Lo.ms;
    ad.cols = Lo.ns;
    ad.mx = 1;
    ad.nx = P->Nx;
    ad.nlive = Lo.nlive;
    ad.vec = ((uintptr_t)C & 15) == 0 && (ldc & 1) == 0 && ((uintptr_t)M & 15) == 0;
    ad.cptr = cptr;
    ad.cprod = cprod;
    ad.cw = cw;
    madd_kernel<<<gm, dim3(kEwThreads), 0, stream>>>(ad);
  } else {
    AddJobL ad{};
    ad.M = static_cast<const float*>(M);
    ad.C = static_cast<float*>(C);
    ad.ldm = Lo.ldm;
    ad.ldc = ldc;
    ad.rows = Lo.ms;
    ad.cols = Lo.ns;
    ad.mx = 1;
    ad.nx = P->Nx;
    ad.nlive = Lo.nlive;
    ad.vec = ((uintptr_t)C & 7) == 0 && (ldc & 1) == 0 && ((uintptr_t)M & 7) == 0;
    ad.cptr = cptr;
    ad.cprod = cprod;
    ad.cw = cw;
    madd_lowp_kernel<<<gm, dim3(kEwThreads), 0, stream>>>(ad);
  }
  if (hipGetLastError() != hipSuccess) return kEHip;
  return kOK;
}

}  // namespace fmm
