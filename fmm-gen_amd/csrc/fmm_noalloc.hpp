// fmm_noalloc.hpp -- fmm_abc_noalloc: the ABC call that allocates nothing.
//
// For callers that cannot call hipMalloc (code under HIP graph capture,
// tight serving loops, harnesses that forbid allocation).  fmm_abc takes the
// workspace of its ordered launch (ticket, error flags, one order counter per
// C tile of the leaf block) from the caller or allocates it on the stream;
// this variant takes it from `scratch`, a file-scope device array of fixed
// size that every GPU the library runs on holds until the process ends.  The
// product tables of the plan (and of the plain-GEMM plan its fringes use)
// are read where an earlier call on the device left them, or else copied
// into the scratch array too.  The kernels are fmm_abc's, with the same
// arguments except the workspace address, so results equal fmm_abc's bit for
// bit.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (other streams, other host threads) race and give wrong results; the
// array keeps the last call's contents, and each call clears what has to
// start at zero (the ticket, flags and counters, by a fill kernel).
//
// Host code never uses the array's name as an address: on the host it is a
// stand-in symbol, and a kernel given it would fault.  The device address of
// each GPU's copy is looked up once per device (hipGetSymbolAddress, at the
// entry point's first call on that device) and kept per device.
#pragma once
#include <atomic>

#include "fmm_host.hpp"

namespace fmm {

// Scratch layout, in ints:
//   [0, kNaWsInts)                       launch workspace (kWs* layout), cleared per call
//   [kNaWsInts, + kNaStickyInts)         word 0: flags of the calls since the last status query
//   [kNaTableOff, + kNaTableBytes / 4)   product tables copied from the host when not resident
constexpr int kNaWsInts = 1 << 16;                  // 256 KiB: counters for 65,272 C tiles of a leaf block
constexpr int kNaStickyInts = 64;
constexpr int kNaTableOff = kNaWsInts + kNaStickyInts;
constexpr size_t kNaTableBytes = 1u << 20;          // 1 MiB of Prod / Term tables
constexpr size_t kNaScratchInts = (size_t)kNaTableOff + kNaTableBytes / sizeof(int);
static_assert(kNaTableBytes / sizeof(int) >= (size_t)kXcdProbeInts, "probe buffer");

static __device__ __align__(256) int scratch[kNaScratchInts];

struct NoAllocDev {
  int* ptr = nullptr;          // device address of this GPU's scratch
  unsigned long long tables = 0;   // uid of the plan whose tables the scratch holds (0: none)
};

struct NoAllocState {
  std::mutex mu;
  std::vector<NoAllocDev> dev;
};
inline NoAllocState& noalloc_state() {
  static NoAllocState s;
  return s;
}

// Device address of the current GPU's scratch (set-up on first use: symbol
// lookup, a clear of the flag words, and the XCD placement probe run in the
// table area, so the later launches of the entry point need no set-up).
inline int noalloc_device(int dev, NoAllocDev** out) {
  NoAllocState& st = noalloc_state();
  std::lock_guard<std::mutex> g(st.mu);
  if ((int)st.dev.size() <= dev) st.dev.resize(dev + 1);
  NoAllocDev& d = st.dev[dev];
  if (!d.ptr) {
    int* ptr = nullptr;
    if (hipGetSymbolAddress(reinterpret_cast<void**>(&ptr), HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return kEHip;
    hipLaunchKernelGGL(ws_clear_kernel, dim3(1), dim3(64), 0, 0, ptr + kNaWsInts, kNaStickyInts);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return kEHip;
    (void)ensure_xcd_probe(dev, ptr + kNaTableOff);
    d.ptr = ptr;
    d.tables = 0;
  }
  *out = &d;
  return kOK;
}

// Flags the allocation-free calls left on the current device since the last
// query; cleared by the query (synchronises the device).
inline int noalloc_take_flags(int* flags) {
  *flags = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  int* w = nullptr;
  {
    NoAllocState& st = noalloc_state();
    std::lock_guard<std::mutex> g(st.mu);
    if (dev < (int)st.dev.size()) w = st.dev[dev].ptr;
  }
  if (!w) return kOK;
  if (hipDeviceSynchronize() != hipSuccess) return kEHip;
  if (hipMemcpy(flags, w + kNaWsInts, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) return kEHip;
  if (*flags != 0) {
    const int zero = 0;
    if (hipMemcpy(w + kNaWsInts, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return kEHip;
  }
  return kOK;
}

template <class T>
inline T* resident(Plan* P, DevVec<T>& v, int dev) {
  std::lock_guard<std::mutex> g(P->mu);
  return dev < (int)v.d.size() ? v.d[dev] : nullptr;
}

inline size_t align16(size_t x) { return (x + 15) & ~(size_t)15; }

inline int run_abc_noalloc(Plan* P, int dtype, int m, int n, int k, const void* A, int lda, const void* B, int ldb,
                           void* C, int ldc, hipStream_t s) {
  if (bad_dims(m, n, k, lda, ldb, ldc)) return kEInval;
  KGeo G;
  if (!dtype_geo(dtype, &G, 0, nullptr, P->unit_single)) return kEUnsupported;
  if (m == 0 || n == 0 || k == 0) return kOK;
  // Limits first, on the host: the workspace of this shape and the size of
  // the tables a non-resident plan would copy into the scratch array.
  {
    JobList JL0;
    add_fmm_jobs(JL0, G, *P, nullptr, nullptr, nullptr, nullptr, m, n, k, A, lda, B, ldb, C, ldc);
    if (ws_ints(JL0) > (size_t)kNaWsInts) return kETooBig;
  }
  Plan* I = identity_plan();
  const std::vector<Prod>& hp = P->main.prods.h;
  const std::vector<Term>& ht = P->main.terms.h;
  const std::vector<Prod>& hip_ = I->main.prods.h;
  const std::vector<Term>& hit = I->main.terms.h;
  const size_t o1 = align16(sizeof(Prod) * hp.size());
  const size_t o2 = o1 + align16(sizeof(Term) * ht.size());
  const size_t o3 = o2 + align16(sizeof(Prod) * hip_.size());
  const size_t end = o3 + align16(sizeof(Term) * hit.size());
  if (end > kNaTableBytes) return kETooBig;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return kEHip;
  NoAllocDev* d = nullptr;
  int rc = noalloc_device(dev, &d);
  if (rc != kOK) return rc;
  Prod* dp = resident(P, P->main.prods, dev);
  Term* dt = resident(P, P->main.terms, dev);
  Prod* ip = resident(I, I->main.prods, dev);
  Term* it = resident(I, I->main.terms, dev);
  if (!dp || !dt || !ip || !it) {
    // not resident on this device: the four tables go to the scratch table area
    char* base = reinterpret_cast<char*>(d->ptr + kNaTableOff);
    {
      NoAllocState& st = noalloc_state();
      std::lock_guard<std::mutex> g(st.mu);
      if (d->tables != P->uid) {
        if (hipMemcpyAsync(base, hp.data(), sizeof(Prod) * hp.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(base + o1, ht.data(), sizeof(Term) * ht.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
            hipMemcpyAsync(base + o2, hip_.data(), sizeof(Prod) * hip_.size(), hipMemcpyHostToDevice, s) !=
                hipSuccess ||
            hipMemcpyAsync(base + o3, hit.data(), sizeof(Term) * hit.size(), hipMemcpyHostToDevice, s) != hipSuccess) {
          d->tables = 0;
          return kEHip;
        }
        d->tables = P->uid;
      }
    }
    dp = reinterpret_cast<Prod*>(base);
    dt = reinterpret_cast<Term*>(base + o1);
    ip = reinterpret_cast<Prod*>(base + o2);
    it = reinterpret_cast<Term*>(base + o3);
  }
  JobList JL;
  add_fmm_jobs(JL, G, *P, dp, dt, ip, it, m, n, k, A, lda, B, ldb, C, ldc);
  const size_t need = ws_ints(JL);
  rc = launch_jobs_dtype(dtype, JL, P->max_nnz, need > 0 ? d->ptr : nullptr, s, P->unit_single);
  if (rc == kOK && need > 0) rc = fold_error(d->ptr + kNaWsInts, d->ptr, s);
  return rc;
}

}  // namespace fmm
