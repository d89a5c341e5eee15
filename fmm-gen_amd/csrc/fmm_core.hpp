// fmm_core.hpp -- device side of the MI355X (gfx950) ABC fast-matrix-multiply.
//
// One persistent kernel runs a list of *jobs*.  A job is C += A*B over a leaf
// grid: an FMM core (R products of an <Mx,Kx,Nx> algorithm, plus the k-fringe
// as Mx*Nx extra one-term products), or a plain GEMM (the one-product identity
// plan).  The reference runs the same arithmetic as a chain of OpenMP loops per
// product (meta/abc_gen_mix.py:302-391, 448-472); here the work is cut into
// *units* = (product q, C-tile position p) and handed to resident workgroups
// through an atomic ticket.
//
// Per unit (all in registers / LDS, no workspace):
//   * fused packing: sum_t u_t * A_t and sum_t v_t * B_t are formed from
//     coalesced global loads in registers and stored to LDS, one BK-deep slice
//     at a time (the reference's packA/B_add_stra_abc, meta/abc_gen_mix.py:
//     171-251), with the next slice's loads in flight during the MFMAs;
//   * micro-kernel: v_mfma_f64_16x16x4f64 over the whole leaf depth into a
//     128x128 FP64 accumulator tile (8 waves x 64x32), computed transposed
//     (B^T A^T) so the accumulators map onto contiguous runs of C;
//   * W scatter: C_j = C_j + w_j * acc for every C block the product feeds
//     (meta/abc_micro_kernel_gen.py:147-196; bl_dgemm_asm_8x4_mulstrassen.c:
//     324-383), straight from the accumulators.
//
// Determinism: products that update the same C tile are applied in product
// order 0..R-1 (then the k-fringe), enforced by one counter per tile
// position: unit (q,p) does its MFMA work at once but waits for counter[p]==q
// before its read-modify-write of C, then sets counter[p]=q+1.  Tickets are
// taken in unit order and a unit only ever waits for a lower-numbered unit
// that an already-running workgroup holds, so the wait always ends; a bounded
// spin (wall clock) turns an impossible wait into an error flag instead of a
// hang.  Results are therefore bit-for-bit reproducible run to run.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmm {

// One coefficient of a product: block (bi, bj) of the operand, scaled by coef.
struct Term {
  double coef;
  int bi;
  int bj;
};

// A product: ranges into the term array; `fringe` products run over the
// k-fringe depth (kr) instead of the leaf depth (ks).
struct Prod {
  int a0, na, b0, nb, c0, nc, fringe, pad;
};

// Operand pointers are untyped: the kernel that runs the job knows the
// element types (FP64 here; FP32 / BF16 / FP16 inputs with FP32 C in the
// low-precision kernels).
struct Job {
  const void* A;
  const void* B;
  void* C;
  long long lda, ldb, ldc;
  int ms, ns, ks, kr;       // leaf block and k-fringe depth
  int tiles_m, tiles_n;     // C tiles over one leaf block
  int nprod;                // products run by this job (R [+ Mx*Nx fringe])
  int gsize;                // products per unit (consecutive, same C tile)
  int ngroups;              // ceil(nprod / gsize)
  int unit_base;            // first unit of this job in the launch
  int order_base;           // first order counter, <0 when ngroups == 1
  int flags;                // bit0/1/2: A/B/C allow 16-byte vector access;
                            // kJobStore: C = w*AB instead of C += w*AB
  const Prod* prods;
  const Term* terms;
  long long alo, blo;       // split-BF16 planes (direct kernel): element offset
                            // of the lo plane from the hi plane of A / B
};

constexpr int kInlineJobs = 8;
constexpr int kJobStore = 8;
// Low-precision kernels: run this job with the kernel's second C-tile shape
// (128 x 256 instead of 256 x 128).
constexpr int kJobTileT = 16;

struct Launch {
  int njobs;
  int total_units;
  int use_ticket;           // 1: units are claimed from per-queue tickets in ws
  int nq;                   // unit queues (1 or 8, one per XCD)
  int local_handover;       // 1: each queue runs on one XCD (see wait_turn)
  int pad2;
  const Job* jobs;          // device job table, or nullptr -> inl[]
  const int* hdr;           // with a job table: {njobs, total_units} in device
                            // memory (written before the launch), so one
                            // captured launch serves any table contents
  int* ws;                  // workspace, layout kWs* below
  long long spin_limit;     // bounded wait, in wall_clock64() ticks
  Job inl[kInlineJobs];
};

enum : int { kErrSpin = 1 };

// Workspace layout (ints): queue x's ticket at kWsTicket*x (one 128-byte line
// each), the error flags at kWsError, the per-tile order counters from
// kWsCounters on.
constexpr int kMaxQueues = 8;
constexpr int kWsTicket = 32;
constexpr int kWsError = kWsTicket * kMaxQueues;
constexpr int kWsExit = kWsError + 4;      // workgroups that finished their queue
constexpr int kWsCounters = kWsError + 8;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Pointers reached through the job table are generic; accesses through them
// would compile to flat_* instructions, which also count against the LDS
// wait counter and would make every LDS wait drain the in-flight global
// prefetch.  All job memory is device (global) memory, so cast explicitly.
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) const f64x2 gcf64x2;
typedef __attribute__((address_space(1))) const int gcint;
__device__ __forceinline__ gcdouble* gptr(const double* p) { return (gcdouble*)p; }
__device__ __forceinline__ gdouble* gptr(double* p) { return (gdouble*)p; }
__device__ __forceinline__ Term term_at(const Job& J, int i) {
  const Term* t = J.terms + i;
  Term r;
  r.coef = *(gcdouble*)&t->coef;
  r.bi = *(gcint*)&t->bi;
  r.bj = *(gcint*)&t->bj;
  return r;
}
__device__ __forceinline__ Prod prod_at(const Job& J, int q) {
  gcint* p = (gcint*)(J.prods + q);
  Prod r;
  r.a0 = p[0];
  r.na = p[1];
  r.b0 = p[2];
  r.nb = p[3];
  r.c0 = p[4];
  r.nc = p[5];
  r.fringe = p[6];
  r.pad = 0;
  return r;
}

// Tile geometry of the FP64 kernel: a 128x128 C tile per workgroup of 8
// waves (2 per SIMD), each wave owning a 64x32 sub-tile (4x2 MFMA blocks of
// 16x16); the contraction advances BK deep per LDS slice.
constexpr int NT = 512;
constexpr int WAVES_M = 2;
constexpr int WAVES_N = 4;
constexpr int BM = 128;
constexpr int BN = 128;
constexpr int WTM = BM / WAVES_M;      // 64 rows per wave
constexpr int WTN = BN / WAVES_N;      // 32 columns per wave
constexpr int FM = WTM / 16;           // 4 MFMA blocks down
constexpr int FN = WTN / 16;           // 2 MFMA blocks across
// A slice (BM x BK, column-major source): thread t owns rows 2*(t%64)+{0,1}
// of columns t/64 + A_CPP*i: each wave reads one contiguous 1 KiB column run.
constexpr int A_CPP = NT / 64;         // columns covered per pass
// B slice (BK x BN, column-major source): BK/2 consecutive lanes cover the
// BK-deep run of one column (16 bytes each), so every load instruction reads
// whole cache lines; a thread takes the same k pair of B_IT columns.
#ifndef FMM_BLINE
#define FMM_BLINE 1
#endif
constexpr int B_TPC = NT / BN;

// Slice depth BK is a template parameter: 32 suits products whose operand
// sums have at most 2 terms (fewer barriers per flop), 16 leaves registers
// for 4 terms per chunk.
template <int BK_>
struct TileK {
  static constexpr int BK = BK_;
  static constexpr int BKP = BK + 1;          // padded k stride of the B tile in LDS
  static constexpr int A_IT = BK / A_CPP;     // A passes per slice
  static constexpr int AE = 2 * A_IT;         // A values per thread per slice
  static constexpr int B_KPT = BK / B_TPC;    // B k-run per thread
  static constexpr int B_LPC = BK / 2;        // lanes per column (line layout)
  static constexpr int B_CPP = NT / B_LPC;    // columns per pass (line layout)
  static constexpr int B_IT = BN / B_CPP;     // passes (line layout)
  static constexpr bool LINE = FMM_BLINE && BK_ <= 16;
  static constexpr int BE = B_KPT;            // B values per thread per slice
  static_assert(BM == 128 && A_CPP * A_IT == BK && B_TPC * B_KPT == BK, "staging layout");
  static_assert(B_CPP * B_IT == BN && 2 * B_IT == BE, "B line layout");
};
static_assert(WAVES_M * WAVES_N * 64 == NT, "wave layout");

__device__ __forceinline__ const Job& job_at(const Launch& L, int j) {
  return L.jobs ? L.jobs[j] : L.inl[j];
}

// Job count and unit total of a launch: from the device header when the jobs
// come from a device table, else from the launch arguments.
__device__ __forceinline__ void launch_dims(const Launch& L, int* njobs, int* total) {
  if (L.hdr) {
    *njobs = __builtin_amdgcn_readfirstlane(*(gcint*)&L.hdr[0]);
    *total = __builtin_amdgcn_readfirstlane(*(gcint*)&L.hdr[1]);
  } else {
    *njobs = L.njobs;
    *total = L.total_units;
  }
}

// Largest j < njobs with unit_base <= u (jobs are sorted by unit_base).
__device__ __forceinline__ int find_job(const Launch& L, int njobs, int u) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (job_at(L, mid).unit_base <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <int BK>
struct Slice {
  double a[TileK<BK>::AE];
  double b[TileK<BK>::BE];
};

template <int BK, int CH>
struct Raw {
  double a[CH][TileK<BK>::AE];
  double b[CH][TileK<BK>::BE];
};

struct UnitGeo {
  int row0;        // first leaf row of this tile
  int col0;        // first leaf column of this tile
  int depth;       // k extent (ks or kr)
  int arow;        // this thread's A row (even) within the tile
  int acol;        // this thread's first A column within the slice
  int bcol;        // this thread's B column within the tile
  int bk;          // this thread's first B k within the slice
  bool interior;   // whole tile inside the leaf block and 16-byte accessible
  long long aoff;  // this thread's A element offset within a block at k = 0
  long long boff;  // this thread's B element offset within a block at k = 0
};

__device__ __forceinline__ gcdouble* a_block(const Job& J, const Term& t) {
  return gptr(static_cast<const double*>(J.A)) + (long long)t.bi * J.ms + (long long)t.bj * J.ks * J.lda;
}
__device__ __forceinline__ gcdouble* b_block(const Job& J, const Term& t) {
  return gptr(static_cast<const double*>(J.B)) + (long long)t.bi * J.ks + (long long)t.bj * J.ns * J.ldb;
}

// Uniform (wave-invariant) copies of values the compiler cannot prove
// uniform, so they live in scalar registers.
__device__ __forceinline__ long long uni64(long long x) {
  const int lo = __builtin_amdgcn_readfirstlane((int)(x & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(x >> 32));
  return ((long long)hi << 32) | (unsigned int)lo;
}
__device__ __forceinline__ double unid(double x) { return __longlong_as_double(uni64(__double_as_longlong(x))); }
template <class T>
__device__ __forceinline__ T* unip(T* p) { return (T*)uni64((long long)p); }

// The source blocks of one chunk of a product (up to CH terms of A and of B):
// block base pointers and coefficients, all uniform across the workgroup.
// A thread's element is base + its fixed offset (UnitGeo::aoff / boff).
template <int CH>
struct Srcs {
  gcdouble* a[CH];
  gcdouble* b[CH];
  double u[CH];
  double v[CH];
  int na, nb;
};

template <int CH>
__device__ __forceinline__ void make_srcs(const Job& J, int ta0, int na, int tb0, int nb, Srcs<CH>& S) {
  S.na = na < CH ? na : CH;
  S.nb = nb < CH ? nb : CH;
#pragma unroll
  for (int t = 0; t < CH; ++t) {
    if (t < S.na) {
      const Term tm = term_at(J, ta0 + t);
      S.a[t] = unip(a_block(J, tm));
      S.u[t] = unid(tm.coef);
    } else {
      S.a[t] = nullptr;
      S.u[t] = 0.0;
    }
    if (t < S.nb) {
      const Term tm = term_at(J, tb0 + t);
      S.b[t] = unip(b_block(J, tm));
      S.v[t] = unid(tm.coef);
    } else {
      S.b[t] = nullptr;
      S.v[t] = 0.0;
    }
  }
}

// Issue the global loads of slice kb for the chunk's sources.  A full slice
// of an interior tile takes unpredicated 16-byte loads; any other slice is
// loaded element-wise with zero fill outside the leaf block and depth.
template <int BK, int CH>
__device__ __forceinline__ void load_raw(const Job& J, const Srcs<CH>& S, int kb, const UnitGeo& g,
                                         Raw<BK, CH>& r) {
  constexpr int A_IT = TileK<BK>::A_IT, B_KPT = TileK<BK>::B_KPT;
  const bool full = g.interior && kb + BK <= g.depth;
  if (full) {
    const long long ka = (long long)kb * J.lda;
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      if (t < S.na) {
        gcdouble* base = S.a[t] + ka;
#pragma unroll
        for (int i = 0; i < A_IT; ++i) {
          const f64x2 v = *(gcf64x2*)(base + (long long)(A_CPP * i) * J.lda + g.aoff);
          r.a[t][2 * i] = v.x;
          r.a[t][2 * i + 1] = v.y;
        }
      }
      if (t < S.nb) {
        gcdouble* base = S.b[t] + kb;
        if constexpr (TileK<BK>::LINE) {
#pragma unroll
          for (int i = 0; i < TileK<BK>::B_IT; ++i) {
            const f64x2 v = *(gcf64x2*)(base + (long long)(TileK<BK>::B_CPP * i) * J.ldb + g.boff);
            r.b[t][2 * i] = v.x;
            r.b[t][2 * i + 1] = v.y;
          }
        } else {
#pragma unroll
          for (int e = 0; e < B_KPT; e += 2) {
            const f64x2 v = *(gcf64x2*)(base + e + g.boff);
            r.b[t][e] = v.x;
            r.b[t][e + 1] = v.y;
          }
        }
      }
    }
    return;
  }
  if (CH == 1 && g.interior) {
    // Last (partial) slice of an interior tile: every row and column is in
    // range, only k runs out, so the loads stay 16-byte wide and only the
    // k predicate remains (a leaf depth that is no multiple of BK, e.g. the
    // 240-deep leaves of a 222 product at k = 480, ends every unit here).
    // Single-term kernel only: in the operand-sum kernels the extra path
    // cost registers (a spill) and measured slower.
    const long long ka = (long long)kb * J.lda;
    const int krem = g.depth - (kb + g.bk);
#pragma unroll
    for (int t = 0; t < CH; ++t) {
      if (t < S.na) {
        gcdouble* base = S.a[t] + ka;
#pragma unroll
        for (int i = 0; i < A_IT; ++i) {
          f64x2 v = {0.0, 0.0};
          if (g.acol + kb + A_CPP * i < g.depth) v = *(gcf64x2*)(base + (long long)(A_CPP * i) * J.lda + g.aoff);
          r.a[t][2 * i] = v.x;
          r.a[t][2 * i + 1] = v.y;
        }
      }
      if (t < S.nb) {
        gcdouble* base = S.b[t] + kb;
        if constexpr (TileK<BK>::LINE) {
#pragma unroll
          for (int i = 0; i < TileK<BK>::B_IT; ++i) {
            gcdouble* p = base + (long long)(TileK<BK>::B_CPP * i) * J.ldb + g.boff;
            f64x2 v = {0.0, 0.0};
            if (1 < krem) {
              v = *(gcf64x2*)p;
            } else if (0 < krem) {
              v.x = p[0];
            }
            r.b[t][2 * i] = v.x;
            r.b[t][2 * i + 1] = v.y;
          }
        } else {
#pragma unroll
          for (int e = 0; e < B_KPT; e += 2) {
            f64x2 v = {0.0, 0.0};
            if (e + 1 < krem) {
              v = *(gcf64x2*)(base + e + g.boff);
            } else if (e < krem) {
              v.x = base[e + g.boff];
            }
            r.b[t][e] = v.x;
            r.b[t][e + 1] = v.y;
          }
        }
      }
    }
    return;
  }
  const int row = g.row0 + g.arow;
  const bool r0ok = row < J.ms, r1ok = row + 1 < J.ms;
  const bool bcok = g.col0 + g.bcol < J.ns;
  const int krem = g.depth - (kb + g.bk);   // valid k in this thread's B run
#pragma unroll
  for (int t = 0; t < CH; ++t) {
    if (t < S.na) {
#pragma unroll
      for (int i = 0; i < A_IT; ++i) {
        const int col = kb + A_CPP * i;
        const bool cok = g.acol + col < g.depth;
        gcdouble* p = S.a[t] + (long long)col * J.lda + g.aoff;
        r.a[t][2 * i] = (cok && r0ok) ? p[0] : 0.0;
        r.a[t][2 * i + 1] = (cok && r1ok) ? p[1] : 0.0;
      }
    }
    if (t < S.nb) {
      gcdouble* p = S.b[t] + kb + g.boff;
      if constexpr (TileK<BK>::LINE) {
#pragma unroll
        for (int i = 0; i < TileK<BK>::B_IT; ++i) {
          const bool cok = g.col0 + g.bcol + TileK<BK>::B_CPP * i < J.ns;
          gcdouble* q = p + (long long)(TileK<BK>::B_CPP * i) * J.ldb;
          r.b[t][2 * i] = (cok && 0 < krem) ? q[0] : 0.0;
          r.b[t][2 * i + 1] = (cok && 1 < krem) ? q[1] : 0.0;
        }
      } else {
#pragma unroll
        for (int e = 0; e < B_KPT; ++e) r.b[t][e] = (bcok && e < krem) ? p[e] : 0.0;
      }
    }
  }
}

// s = u0*x0, then s = s + u_t*x_t left to right (fma == add of the exact
// product for the dyadic table coefficients, i.e. the reference's packing
// arithmetic, meta/abc_gen_mix.py:229-251).
template <int BK, int CH>
__device__ __forceinline__ void accumulate(const Srcs<CH>& S, bool first, const Raw<BK, CH>& r, Slice<BK>& s) {
  constexpr int AE = TileK<BK>::AE, BE = TileK<BK>::BE;
#pragma unroll
  for (int t = 0; t < CH; ++t) {
    if (t < S.na) {
#pragma unroll
      for (int e = 0; e < AE; ++e)
        s.a[e] = (first && t == 0) ? S.u[t] * r.a[t][e] : __builtin_fma(S.u[t], r.a[t][e], s.a[e]);
    }
    if (t < S.nb) {
#pragma unroll
      for (int e = 0; e < BE; ++e)
        s.b[e] = (first && t == 0) ? S.v[t] * r.b[t][e] : __builtin_fma(S.v[t], r.b[t][e], s.b[e]);
    }
  }
}

// Full operand sums of slice kb: `r` already holds chunk 0 (issued earlier);
// products with more than CH terms load their remaining chunks here.
template <int BK, int CH>
__device__ __forceinline__ void finish_sums(const Job& J, const Prod& P, const Srcs<CH>& S0, int kb,
                                            const UnitGeo& g, Raw<BK, CH>& r, Slice<BK>& s) {
  accumulate<BK, CH>(S0, true, r, s);
  const int nmax = P.na > P.nb ? P.na : P.nb;
  for (int c = CH; c < nmax; c += CH) {
    Srcs<CH> S;
    make_srcs<CH>(J, P.a0 + c, P.na - c, P.b0 + c, P.nb - c, S);
    load_raw<BK, CH>(J, S, kb, g, r);
    accumulate<BK, CH>(S, false, r, s);
  }
}

template <int BK>
__device__ __forceinline__ void store_slice(double (*As)[BM], double (*Bs)[TileK<BK>::BKP], const UnitGeo& g,
                                            const double (&sa)[TileK<BK>::AE], const double (&sb)[TileK<BK>::BE]) {
  constexpr int A_IT = TileK<BK>::A_IT, B_KPT = TileK<BK>::B_KPT;
#pragma unroll
  for (int i = 0; i < A_IT; ++i) {
    double2 v;
    v.x = sa[2 * i];
    v.y = sa[2 * i + 1];
    *reinterpret_cast<double2*>(&As[g.acol + A_CPP * i][g.arow]) = v;
  }
  if constexpr (TileK<BK>::LINE) {
#pragma unroll
    for (int i = 0; i < TileK<BK>::B_IT; ++i) {
      Bs[g.bcol + TileK<BK>::B_CPP * i][g.bk] = sb[2 * i];
      Bs[g.bcol + TileK<BK>::B_CPP * i][g.bk + 1] = sb[2 * i + 1];
    }
  } else {
#pragma unroll
    for (int e = 0; e < B_KPT; ++e) Bs[g.bcol][g.bk + e] = sb[e];
  }
}

// Operand sums of slice kb (chunk 0 already loaded into r) stored to LDS.
// The single-term kernel scales the loaded values in place, and only when
// the coefficient is not 1 (plain GEMMs and Naive products have 1): the FP64
// multiplies and register copies this saves cost MFMA throughput.
template <int BK, int CH>
__device__ __forceinline__ void pack_slice(const Job& J, const Prod& P, const Srcs<CH>& S, int kb, const UnitGeo& g,
                                           Raw<BK, CH>& r, Slice<BK>& s, double (*As)[BM],
                                           double (*Bs)[TileK<BK>::BKP]) {
  if constexpr (CH == 1) {
    if (S.u[0] != 1.0) {
#pragma unroll
      for (int e = 0; e < TileK<BK>::AE; ++e) r.a[0][e] = S.u[0] * r.a[0][e];
    }
    if (S.v[0] != 1.0) {
#pragma unroll
      for (int e = 0; e < TileK<BK>::BE; ++e) r.b[0][e] = S.v[0] * r.b[0][e];
    }
    store_slice<BK>(As, Bs, g, r.a[0], r.b[0]);
  } else {
    finish_sums<BK, CH>(J, P, S, kb, g, r, s);
    store_slice<BK>(As, Bs, g, s.a, s.b);
  }
}

// MFMAs of k-steps [KK0, KK1) of the slice.  v_mfma_f64_16x16x4f64 computes
// D[i][j] = sum_k X[i][k] Y[k][j] with X[i][k] in lane i+16k, Y[k][j] in lane
// j+16k and element v of lane l holding D[l/16 + 4v][l%16] (layout measured
// on gfx950).  X = B^T and Y = A^T, so D is the transposed C block: element v
// of lane l is C row l%16, column l/16 + 4v.
//
// Scheduling per kernel variant (FP64 measurements, 9600^3): the single-term
// kernel (CH = 1; plain GEMMs, Naive products) reads the fragments of step
// kk+1 while the MFMAs of step kk issue and runs the two waves of a SIMD in
// opposite phase order (pingpong, see contract), at equal wave priority
// (raising it for the MFMA phases cost 1%); the operand-sum kernels have too
// few free registers for either and run faster without both (ABC 333: 49.6
// -> 51.4 TFLOP/s), but gain from the raised priority (s_setprio: 49.6 ->
// 51.5).
#ifndef FMM_ABC_SCHED
#define FMM_ABC_SCHED 0
#endif
template <int CH>
struct F64Sched {
#ifndef FMM_ABC_FP
#define FMM_ABC_FP 0
#endif
  static constexpr bool frag_pipe = CH == 1 || FMM_ABC_FP != 0;
  static constexpr bool pingpong = CH == 1 || FMM_ABC_SCHED != 0;
  static constexpr bool setprio = CH != 1 && FMM_ABC_SCHED != 2;
};

template <int BK, int KK0, int KK1, bool FRAG_PIPE, bool SETPRIO>
__device__ __forceinline__ void mfma_slice(const double (*As)[BM], const double (*Bs)[TileK<BK>::BKP], int lane, int wm,
                                           int wn, f64x4 (&acc)[FM][FN]) {
  const int l16 = lane & 15, lk = lane >> 4;
  if constexpr (SETPRIO) __builtin_amdgcn_s_setprio(1);
  if constexpr (FRAG_PIPE) {
    // Fragments of step kk+1 are read while the MFMAs of step kk issue.
    double a[2][FM], b[2][FN];
#pragma unroll
    for (int mi = 0; mi < FM; ++mi) a[0][mi] = As[KK0 * 4 + lk][wm * WTM + mi * 16 + l16];
#pragma unroll
    for (int ni = 0; ni < FN; ++ni) b[0][ni] = Bs[wn * WTN + ni * 16 + l16][KK0 * 4 + lk];
#pragma unroll
    for (int kk = KK0; kk < KK1; ++kk) {
      const int c = (kk - KK0) & 1;
      if (kk + 1 < KK1) {
#pragma unroll
        for (int mi = 0; mi < FM; ++mi) a[c ^ 1][mi] = As[(kk + 1) * 4 + lk][wm * WTM + mi * 16 + l16];
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) b[c ^ 1][ni] = Bs[wn * WTN + ni * 16 + l16][(kk + 1) * 4 + lk];
      }
#pragma unroll
      for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[c][ni], a[c][mi], acc[mi][ni], 0, 0, 0);
      if (kk + 1 < KK1) {
        // interleave: one LDS read after every two MFMAs
#pragma unroll
        for (int x = 0; x < (FM + FN) / 2; ++x) {
          __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);
          __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
        }
        __builtin_amdgcn_sched_group_barrier(0x008, FM * FN, 0);
      }
    }
  } else {
#pragma unroll
    for (int kk = KK0; kk < KK1; ++kk) {
      double a[FM], b[FN];
#pragma unroll
      for (int mi = 0; mi < FM; ++mi) a[mi] = As[kk * 4 + lk][wm * WTM + mi * 16 + l16];
#pragma unroll
      for (int ni = 0; ni < FN; ++ni) b[ni] = Bs[wn * WTN + ni * 16 + l16][kk * 4 + lk];
#pragma unroll
      for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ni], a[mi], acc[mi][ni], 0, 0, 0);
    }
  }
  if constexpr (SETPRIO) __builtin_amdgcn_s_setprio(0);
}

// Hand-over of a C tile between consecutive product units.
//
// Reader: an agent-scope acquire (buffer_inv sc1: L1 and non-coherent L2 lines
// invalidated) -- cheap.  An L1-only invalidate (buffer_inv sc0) was measured
// not to suffice outside threadgroup-split dispatch: a CU re-reading a tile it
// had cached returned stale C.
//
// Writer: with agent scope the release writes back the whole XCD L2
// (buffer_wbl2 sc1), which flushes every dirty line of the XCD and measured
// 8-13% of an FMM launch.  With local_handover all units of a tile run in the
// queue of one XCD (a workgroup serves the queue of the XCD it runs on, read
// from HW_REG_XCC_ID; see run_units), so writer and reader share that L2 and
// the release only has to wait until the writer's C stores are acknowledged
// by it (s_waitcnt vmcnt(0)).  The counter is always accessed with
// agent-scope atomics.
__device__ __forceinline__ void acquire_c() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }

__device__ __forceinline__ void release_c(bool local) {
  if (local) {
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
  } else {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  }
}

// Wait (thread 0, bounded) until counter reaches `want`, polling with relaxed
// loads (an acquiring poll would invalidate caches on every iteration); then
// every thread acquires once so its C reads see the previous product's writes.
__device__ __forceinline__ void wait_turn(int* ctr, int want, const Launch& L) {
  if (threadIdx.x == 0) {
    if (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
      const long long t0 = wall_clock64();
      while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < want) {
        __builtin_amdgcn_s_sleep(4);
        // A failed wait anywhere in the launch ends every later wait at once,
        // so a broken ordering costs one time bound, not one per unit.
        if (__hip_atomic_load(L.ws + kWsError, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (wall_clock64() - t0 > L.spin_limit) {
          atomicOr(L.ws + kWsError, (int)kErrSpin);
          break;
        }
      }
    }
  }
  __syncthreads();
  acquire_c();
}

__device__ __forceinline__ void pass_turn(int* ctr, int next, const Launch& L) {
  release_c(L.local_handover != 0);
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(ctr, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Hardware XCC (XCD) id of the executing workgroup: HW_REG_XCC_ID, bits 3:0.
__device__ __forceinline__ int xcc_id() {
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) & 15;
}

// C_j = C_j + w_j * acc for every C term of the product (C_j = w_j * acc for
// a kJobStore job, whose C is a product workspace), straight from the
// accumulators: for each element v a wave touches 4 columns x 16 consecutive
// rows, i.e. four full 128-byte runs of the column-major C.  All reads of a
// term's elements are issued before its first write: C is written through
// the pointer it is read through and, with a run-time ldc, the compiler
// cannot prove two elements distinct, so a read placed after a write waited
// for it (32 serial memory round trips per term, 8% of a Naive product
// launch when measured).
__device__ __forceinline__ void scatter_c(const Job& J, const Prod& P, int row0, int col0, int lane, int wm,
                                          int wn, const f64x4 (&acc)[FM][FN]) {
  const int l16 = lane & 15, lk = lane >> 4;
  const bool store = (J.flags & kJobStore) != 0;
  const int rb = row0 + wm * WTM + l16;
  const int cb0 = col0 + wn * WTN + lk;
  for (int t = 0; t < P.nc; ++t) {
    const Term tm = term_at(J, P.c0 + t);
    const double w = tm.coef;
    gdouble* cb = gptr(static_cast<double*>(J.C)) + (long long)tm.bi * J.ms + (long long)tm.bj * J.ns * J.ldc;
    double old[FM][FN][4];
    if (!store) {
#pragma unroll
      for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int row = rb + mi * 16, col = cb0 + ni * 16 + 4 * e;
            old[mi][ni][e] = (row < J.ms && col < J.ns) ? cb[row + (long long)col * J.ldc] : 0.0;
          }
    }
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < FN; ++ni)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int row = rb + mi * 16, col = cb0 + ni * 16 + 4 * e;
          if (row < J.ms && col < J.ns)
            cb[row + (long long)col * J.ldc] =
                store ? w * acc[mi][ni][e] : __builtin_fma(w, acc[mi][ni][e], old[mi][ni][e]);
        }
  }
}

// Contraction of one unit over its whole depth into acc, double-buffered in
// LDS: while the MFMAs of slice st run on buffer `cur`, slice st+1 (loaded one
// slice earlier) is summed and stored into the other buffer and the loads of
// slice st+2 are issued; one barrier per slice.  A wave does MFMA half 1,
// sums, MFMA half 2.  With pingpong (F64Sched) the second wave of each SIMD
// (waves w and w+4 share one, measured) does the sums first and then all
// MFMAs, so one wave's VALU work and load waits overlap the other wave's
// MFMAs.
template <int BK, int CH>
__device__ __forceinline__ void sums_phase(const Job& J, const Prod& P, const Srcs<CH>& S, int st, int nsteps,
                                           const UnitGeo& g, Raw<BK, CH>& r, Slice<BK>& s, double (*As)[BK][BM],
                                           double (*Bs)[BN][TileK<BK>::BKP], int nxt) {
  if (st + 1 < nsteps) {
    pack_slice<BK, CH>(J, P, S, (st + 1) * BK, g, r, s, As[nxt], Bs[nxt]);
    if (st + 2 < nsteps) load_raw<BK, CH>(J, S, (st + 2) * BK, g, r);
  }
}

template <int BK, int CH>
__device__ __forceinline__ void contract(const Job& J, const Prod& P, const UnitGeo& g, double (*As)[BK][BM],
                                         double (*Bs)[BN][TileK<BK>::BKP], int lane, int wm, int wn, int grp,
                                         f64x4 (&acc)[FM][FN]) {
  const int nsteps = (g.depth + BK - 1) / BK;
  if (!(P.na > 0 && P.nb > 0 && nsteps > 0)) return;
  Srcs<CH> S;
  make_srcs<CH>(J, P.a0, P.na, P.b0, P.nb, S);
  Raw<BK, CH> r;
  Slice<BK> s;
  load_raw<BK, CH>(J, S, 0, g, r);
  pack_slice<BK, CH>(J, P, S, 0, g, r, s, As[0], Bs[0]);
  if (nsteps > 1) load_raw<BK, CH>(J, S, BK, g, r);
  __syncthreads();
  constexpr bool FP = F64Sched<CH>::frag_pipe, PR = F64Sched<CH>::setprio;
  for (int st = 0; st < nsteps; ++st) {
    const int cur = st & 1;
    if (F64Sched<CH>::pingpong && grp != 0) {
      sums_phase<BK, CH>(J, P, S, st, nsteps, g, r, s, As, Bs, cur ^ 1);
      mfma_slice<BK, 0, BK / 4, FP, PR>(As[cur], Bs[cur], lane, wm, wn, acc);
    } else {
      mfma_slice<BK, 0, BK / 8, FP, PR>(As[cur], Bs[cur], lane, wm, wn, acc);
      sums_phase<BK, CH>(J, P, S, st, nsteps, g, r, s, As, Bs, cur ^ 1);
      mfma_slice<BK, BK / 8, BK / 4, FP, PR>(As[cur], Bs[cur], lane, wm, wn, acc);
    }
    __syncthreads();
  }
}

// --- operand sums in staggered rounds (CH == kRounds) ------------------------
//
// The 32-deep single-term pipeline (fragments of step kk+1 read during the
// MFMAs of step kk) with operand sums of up to kRoundTerms terms per operand:
// the sums of slice st+1 are built during the MFMAs of slice st one term at a
// time.  Round t (term t of A and of B) is loaded into the one raw register
// set, then folded into the running sums after two further k-steps:
//
//   end of slice st-1: load round 0 of slice st+1
//   after k-step 1:    s = u0*x0            ; load round 1
//   after k-step 3:    s = fma(u1, x1, s)   ; load round 2
//   after k-step 5:    s = fma(u2, x2, s)   ; load round 3
//   after k-step 7:    s = fma(u3, x3, s)
//   end of slice st:   store s to LDS       ; load round 0 of slice st+2
//
// so only one term's raw values are live at a time (16 instead of 64 FP64
// values per thread for 4 terms) and the 32-deep slice (half the barriers
// of a 16-deep one) fits the register file.  The sums are formed in the same
// order as in the chunked kernels (s = u0*x0, then s = fma(u_t, x_t, s)), so
// the results are bit-identical.  Terms beyond kRoundTerms (tables with more
// than 4 terms per operand) are loaded and folded at the end of the slice
// without prefetch.
constexpr int kRounds = 0;
constexpr int kRoundTerms = 8;

// k-step after which round t of NR is folded: every second k-step from k-step
// 1 while the rounds fit (measured best for up to 4 rounds), else one round
// per k-step.
template <int KK, int NR>
constexpr int round_pos(int t) {
  return 2 * NR <= KK ? 2 * t + 1 : t;
}

template <int N>
__device__ __forceinline__ Srcs<1> round_src(const Srcs<N>& S, int t) {
  Srcs<1> R;
  R.na = t < S.na ? 1 : 0;
  R.nb = t < S.nb ? 1 : 0;
  R.a[0] = S.a[t];
  R.b[0] = S.b[t];
  R.u[0] = S.u[t];
  R.v[0] = S.v[t];
  return R;
}

// MFMAs of one whole slice with fragment pipelining, calling hook(kk) after
// the MFMAs of k-step kk are issued (kk is a constant after unrolling).
template <int BK, class Hook>
__device__ __forceinline__ void mfma_slice_hooked(const double (*As)[BM], const double (*Bs)[TileK<BK>::BKP], int lane,
                                                  int wm, int wn, f64x4 (&acc)[FM][FN], Hook&& hook) {
  constexpr int KK = BK / 4;
  const int l16 = lane & 15, lk = lane >> 4;
  double a[2][FM], b[2][FN];
#pragma unroll
  for (int mi = 0; mi < FM; ++mi) a[0][mi] = As[lk][wm * WTM + mi * 16 + l16];
#pragma unroll
  for (int ni = 0; ni < FN; ++ni) b[0][ni] = Bs[wn * WTN + ni * 16 + l16][lk];
#pragma unroll
  for (int kk = 0; kk < KK; ++kk) {
    const int c = kk & 1;
    if (kk + 1 < KK) {
#pragma unroll
      for (int mi = 0; mi < FM; ++mi) a[c ^ 1][mi] = As[(kk + 1) * 4 + lk][wm * WTM + mi * 16 + l16];
#pragma unroll
      for (int ni = 0; ni < FN; ++ni) b[c ^ 1][ni] = Bs[wn * WTN + ni * 16 + l16][(kk + 1) * 4 + lk];
    }
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < FN; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[c][ni], a[c][mi], acc[mi][ni], 0, 0, 0);
    hook(kk);
  }
}

// Rounds beyond kRoundTerms of slice kb, folded into s without prefetch.
template <int BK>
__device__ __forceinline__ void extra_rounds(const Job& J, const Prod& P, int kb, const UnitGeo& g, Raw<BK, 1>& r,
                                             Slice<BK>& s) {
  const int nr = P.na > P.nb ? P.na : P.nb;
  for (int t = kRoundTerms; t < nr; ++t) {
    Srcs<1> S;
    make_srcs<1>(J, P.a0 + t, P.na - t, P.b0 + t, P.nb - t, S);
    load_raw<BK, 1>(J, S, kb, g, r);
    accumulate<BK, 1>(S, false, r, s);
  }
}

// Unpredicated loads of one operand term's slice kb (interior tile, whole
// slice inside the leaf depth), addressed as a uniform 64-bit base (the term's
// block, the tile's first row or column and kb: scalar arithmetic) plus a
// per-thread 32-bit byte offset that is the same for every term, so the
// loads take the SGPR-base + VGPR-offset form and no per-term address lives
// in vector registers.
struct ThreadOff {
  long long lda;   // uniform copies of the job's leading dimensions
  long long ldb;
  int row0, col0;  // the tile's first row / column (uniform)
  unsigned a;      // (arow + acol * lda) * 8
  unsigned astep;  // A_CPP * lda * 8 (uniform)
  unsigned b;      // (bcol * ldb + bk) * 8
  unsigned a0;     // arow * 8 (the row part of a)
  unsigned b0;     // bcol * ldb * 8 (the column part of b)
  int acol, bk;    // this thread's first A column / B k within a slice
};

__device__ __forceinline__ ThreadOff thread_off(const Job& J, const UnitGeo& g) {
  ThreadOff o;
  o.lda = uni64(J.lda);
  o.ldb = uni64(J.ldb);
  o.row0 = __builtin_amdgcn_readfirstlane(g.row0);
  o.col0 = __builtin_amdgcn_readfirstlane(g.col0);
  // Rows and columns past the leaf block (edge tiles) are clamped to the
  // last row pair / column: they load real, finite values, and the MFMA
  // results of those rows and columns are never stored (scatter_c bounds).
  // Needs an even leaf row count, so the clamped row pair stays 16-byte
  // aligned; the depth must be a multiple of the slice (no k predicate).
  const int arow = min(g.row0 + g.arow, J.ms - 2) - g.row0;
  const int bcol = min(g.col0 + g.bcol, J.ns - 1) - g.col0;
  o.a = (unsigned)((arow + (long long)g.acol * o.lda) * 8);
  o.astep = (unsigned)__builtin_amdgcn_readfirstlane((int)((long long)A_CPP * o.lda * 8));
  o.b = (unsigned)(((long long)bcol * o.ldb + g.bk) * 8);
  o.a0 = (unsigned)(arow * 8);
  o.b0 = (unsigned)((long long)bcol * o.ldb * 8);
  o.acol = g.acol;
  o.bk = g.bk;
  return o;
}

typedef __attribute__((address_space(1))) const char gcchar;

template <int BK>
__device__ __forceinline__ void load_a_full(gcdouble* block, int kb, const ThreadOff& o,
                                            double (&ra)[TileK<BK>::AE]) {
  gcchar* base = (gcchar*)(block + o.row0 + (long long)kb * o.lda);
#pragma unroll
  for (int i = 0; i < TileK<BK>::A_IT; ++i) {
    const f64x2 v = *(gcf64x2*)(base + (o.a + (unsigned)i * o.astep));
    ra[2 * i] = v.x;
    ra[2 * i + 1] = v.y;
  }
}
template <int BK>
__device__ __forceinline__ void load_b_full(gcdouble* block, int kb, const ThreadOff& o,
                                            double (&rb)[TileK<BK>::BE]) {
  static_assert(!TileK<BK>::LINE, "k-run layout");
  gcchar* base = (gcchar*)(block + (long long)o.col0 * o.ldb + kb);
#pragma unroll
  for (int e = 0; e < TileK<BK>::B_KPT; e += 2) {
    const f64x2 v = *(gcf64x2*)(base + (o.b + (unsigned)(e * 8)));
    rb[e] = v.x;
    rb[e + 1] = v.y;
  }
}

// The partial last slice kb (depth - kb < BK): columns / k at or past the
// depth read the last valid column / k pair instead and are zeroed.  The
// depth is even (16-byte B access needs an even leaf depth), so a B k pair is
// valid or invalid as a whole.
template <int BK>
__device__ __forceinline__ void load_a_tail(gcdouble* block, int kb, int depth, const ThreadOff& o,
                                            double (&ra)[TileK<BK>::AE]) {
  gcchar* base = (gcchar*)(block + o.row0 + (long long)kb * o.lda);
  const unsigned last = o.a0 + (unsigned)((long long)(depth - 1 - kb) * o.lda * 8);
#pragma unroll
  for (int i = 0; i < TileK<BK>::A_IT; ++i) {
    const bool ok = kb + o.acol + A_CPP * i < depth;
    const f64x2 v = *(gcf64x2*)(base + (ok ? o.a + (unsigned)i * o.astep : last));
    ra[2 * i] = ok ? v.x : 0.0;
    ra[2 * i + 1] = ok ? v.y : 0.0;
  }
}
template <int BK>
__device__ __forceinline__ void load_b_tail(gcdouble* block, int kb, int depth, const ThreadOff& o,
                                            double (&rb)[TileK<BK>::BE]) {
  gcchar* base = (gcchar*)(block + (long long)o.col0 * o.ldb + kb);
  const unsigned last = o.b0 + (unsigned)((depth - 2 - kb) * 8);
#pragma unroll
  for (int e = 0; e < TileK<BK>::B_KPT; e += 2) {
    const bool ok = kb + o.bk + e < depth;
    const f64x2 v = *(gcf64x2*)(base + (ok ? o.b + (unsigned)(e * 8) : last));
    rb[e] = ok ? v.x : 0.0;
    rb[e + 1] = ok ? v.y : 0.0;
  }
}

template <int N>
__device__ __forceinline__ void fold(double u, bool first, const double (&x)[N], double (&acc)[N]) {
#pragma unroll
  for (int e = 0; e < N; ++e) acc[e] = first ? u * x[e] : __builtin_fma(u, x[e], acc[e]);
}

// contract_rounds for a unit at least BK deep, with 16-byte operand access
// (so an even leaf depth) and an even leaf row count, whose product has NA A
// terms and NB B terms (both <= kRoundTerms), over its `nsteps` full slices
// and, with `tail`, a partial last slice: every
// load, fold and wait is fixed at compile time, with no branch inside a
// slice but the one for the last slice, so each fold waits only for its own
// round's loads (with run-time term counts the waits degrade to vmcnt(0)
// after every round and the loads of the next round are waited for too).
template <int BK, int NA, int NB>
__device__ __forceinline__ void contract_rounds_fixed(const Job& J, const Prod& P, int nsteps, bool tail,
                                                      const UnitGeo& g, double (*As)[BK][BM],
                                                      double (*Bs)[BN][TileK<BK>::BKP], int lane, int wm, int wn,
                                                      f64x4 (&acc)[FM][FN]) {
  constexpr int NR = NA > NB ? NA : NB;
  constexpr int KK = BK / 4;
  static_assert(NA >= 1 && NB >= 1 && NR <= kRoundTerms && NR <= KK, "round layout");
  Srcs<NR> S4;   // the product's sources, in scalar registers
  make_srcs<NR>(J, P.a0, P.na, P.b0, P.nb, S4);
  double ra[TileK<BK>::AE], rb[TileK<BK>::BE];
  Slice<BK> s;
  const ThreadOff o = thread_off(J, g);
  const int klast = (nsteps - 1) * BK;   // the next slice's loads are clamped here: the one
                                         // past-the-end round 0 re-reads the last slice, unused
  auto issue = [&](int t, int kb) {
    if (t < NA) load_a_full<BK>(S4.a[t], kb, o, ra);
    if (t < NB) load_b_full<BK>(S4.b[t], kb, o, rb);
  };
  auto consume = [&](int t) {
    if (t < NA) fold(S4.u[t], t == 0, ra, s.a);
    if (t < NB) fold(S4.v[t], t == 0, rb, s.b);
  };
#pragma unroll
  for (int t = 0; t < NR; ++t) {
    issue(t, 0);
    consume(t);
  }
  store_slice<BK>(As[0], Bs[0], g, s.a, s.b);
  issue(0, min(BK, klast));
  __syncthreads();
  int st = 0;
  for (; st + 1 < nsteps; ++st) {
    const int cur = st & 1;
    const int kb1 = (st + 1) * BK;
    mfma_slice_hooked<BK>(As[cur], Bs[cur], lane, wm, wn, acc, [&](int kk) {
#pragma unroll
      for (int t = 0; t < NR; ++t) {
        if (kk == round_pos<KK, NR>(t)) {
          // keep the machine scheduler from hoisting later rounds' loads
          // (each hoisted round would hold another raw register set)
          __builtin_amdgcn_sched_barrier(0);
          consume(t);
          __builtin_amdgcn_sched_barrier(0);
          if (t + 1 < NR) issue(t + 1, kb1);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    });
    store_slice<BK>(As[cur ^ 1], Bs[cur ^ 1], g, s.a, s.b);
    issue(0, min(kb1 + BK, klast));
    __syncthreads();
  }
  mfma_slice_hooked<BK>(As[st & 1], Bs[st & 1], lane, wm, wn, acc, [](int) {});
  __syncthreads();
  if (tail) {
    // partial last slice, after the full ones (the order of the other
    // kernels), loaded without overlap
    const int kb = nsteps * BK;
#pragma unroll
    for (int t = 0; t < NR; ++t) {
      if (t < NA) load_a_tail<BK>(S4.a[t], kb, g.depth, o, ra);
      if (t < NB) load_b_tail<BK>(S4.b[t], kb, g.depth, o, rb);
      consume(t);
    }
    store_slice<BK>(As[nsteps & 1], Bs[nsteps & 1], g, s.a, s.b);
    __syncthreads();
    mfma_slice_hooked<BK>(As[nsteps & 1], Bs[nsteps & 1], lane, wm, wn, acc, [](int) {});
    __syncthreads();
  }
}

template <int BK>
__device__ __forceinline__ void contract_rounds(const Job& J, const Prod& P, const UnitGeo& g, double (*As)[BK][BM],
                                                double (*Bs)[BN][TileK<BK>::BKP], int lane, int wm, int wn,
                                                f64x4 (&acc)[FM][FN]) {
  static_assert(BK / 4 >= kRoundTerms, "one round per k-step");
  const int nsteps = (g.depth + BK - 1) / BK;
  if (!(P.na > 0 && P.nb > 0 && nsteps > 0)) return;
  const int nr = __builtin_amdgcn_readfirstlane(P.na > P.nb ? P.na : P.nb);
  if ((J.flags & 3) == 3 && (J.ms & 1) == 0 && g.depth >= BK && nr <= kRoundTerms) {
    const int nfull = g.depth / BK;
    const bool tail = nfull * BK != g.depth;
    switch (P.na * 16 + P.nb) {
#define FMM_ROUNDS_CASE(a, b) \
  case (a) * 16 + (b): contract_rounds_fixed<BK, a, b>(J, P, nfull, tail, g, As, Bs, lane, wm, wn, acc); return;
      // every term-count pair of <2,2,2>, <2,3,2>, <3,2,3>, <3,3,3> at one
      // level, <2,2,2> at two and <2,2,2> x <3,2,3> (others: generic path)
      FMM_ROUNDS_CASE(1, 1) FMM_ROUNDS_CASE(1, 2) FMM_ROUNDS_CASE(1, 3) FMM_ROUNDS_CASE(1, 4)
      FMM_ROUNDS_CASE(2, 1) FMM_ROUNDS_CASE(2, 2) FMM_ROUNDS_CASE(2, 3) FMM_ROUNDS_CASE(2, 4)
      FMM_ROUNDS_CASE(3, 1) FMM_ROUNDS_CASE(3, 2) FMM_ROUNDS_CASE(3, 3) FMM_ROUNDS_CASE(3, 4)
      FMM_ROUNDS_CASE(4, 1) FMM_ROUNDS_CASE(4, 2) FMM_ROUNDS_CASE(4, 3) FMM_ROUNDS_CASE(4, 4)
      FMM_ROUNDS_CASE(2, 8) FMM_ROUNDS_CASE(3, 6) FMM_ROUNDS_CASE(4, 8) FMM_ROUNDS_CASE(6, 3)
      FMM_ROUNDS_CASE(6, 6) FMM_ROUNDS_CASE(8, 2) FMM_ROUNDS_CASE(8, 4)
#undef FMM_ROUNDS_CASE
      default: break;
    }
  }
  Srcs<kRoundTerms> S4;
  make_srcs<kRoundTerms>(J, P.a0, P.na, P.b0, P.nb, S4);
  Raw<BK, 1> r;
  Slice<BK> s;
  // slice 0, without overlap
#pragma unroll
  for (int t = 0; t < kRoundTerms; ++t) {
    if (t < nr) {
      const Srcs<1> R = round_src(S4, t);
      load_raw<BK, 1>(J, R, 0, g, r);
      accumulate<BK, 1>(R, t == 0, r, s);
    }
  }
  if (nr > kRoundTerms) extra_rounds<BK>(J, P, 0, g, r, s);
  store_slice<BK>(As[0], Bs[0], g, s.a, s.b);
  if (nsteps > 1) load_raw<BK, 1>(J, round_src(S4, 0), BK, g, r);
  __syncthreads();
  for (int st = 0; st < nsteps; ++st) {
    const int cur = st & 1;
    const bool next = st + 1 < nsteps;
    const int kb1 = (st + 1) * BK;
    mfma_slice_hooked<BK>(As[cur], Bs[cur], lane, wm, wn, acc, [&](int kk) {
      const int t = kk;
      if (next && t < kRoundTerms && t < nr) {
        __builtin_amdgcn_sched_barrier(0);
        accumulate<BK, 1>(round_src(S4, t), t == 0, r, s);
        __builtin_amdgcn_sched_barrier(0);
        if (t + 1 < kRoundTerms && t + 1 < nr) load_raw<BK, 1>(J, round_src(S4, t + 1), kb1, g, r);
        __builtin_amdgcn_sched_barrier(0);
      }
    });
    if (next) {
      if (nr > kRoundTerms) extra_rounds<BK>(J, P, kb1, g, r, s);
      store_slice<BK>(As[cur ^ 1], Bs[cur ^ 1], g, s.a, s.b);
      if (st + 2 < nsteps) load_raw<BK, 1>(J, round_src(S4, 0), kb1 + BK, g, r);
    }
    __syncthreads();
  }
}

// Map queue-local unit index i of queue x to (job, product group, tile).
// Each job's tiles are split into nq contiguous chunks; queue x runs, job by
// job, all product groups over its chunk (groups in order, tiles innermost),
// so every unit a unit waits for (same tile, earlier group) sits earlier in
// the same queue.  Jobs in a device table (many small problems) use one queue.
__device__ __forceinline__ bool decode_unit(const Launch& L, int njobs, int total, int x, int i, int* jo, int* qo,
                                            int* po) {
  if (L.jobs) {
    if (i >= total) return false;
    const int j = find_job(L, njobs, i);
    const Job& J = L.jobs[j];
    const int ntile = J.tiles_m * J.tiles_n;
    const int ul = i - J.unit_base;
    *jo = j;
    *qo = ul / ntile;
    *po = ul - *qo * ntile;
    return true;
  }
  for (int j = 0; j < L.njobs; ++j) {
    const Job& J = L.inl[j];
    const int ntile = J.tiles_m * J.tiles_n;
    const int c0 = (int)(((long long)x * ntile) / L.nq);
    const int c1 = (int)(((long long)(x + 1) * ntile) / L.nq);
    const int cs = c1 - c0;
    const int n = cs * J.ngroups;
    if (i < n) {
      *jo = j;
      *qo = i / cs;
      *po = c0 + (i - *qo * cs);
      return true;
    }
    i -= n;
  }
  return false;
}

// C tile of tile index p (0 <= p < tiles_m * tiles_n) in grouped order: runs
// of G tile rows are walked column by column, so the workgroups of a queue
// (one XCD) working on consecutive tile indices share a few A row panels and
// a few B column panels in that XCD's L2, instead of one A panel and as many
// B panels as workgroups.  A bijection for any grid; G = 1 is row order.
// Measured: +9% for the BF16 ABC kernel at 8192^3 with G = 4 (low-precision
// kernels), no gain for the FP64 kernel (row order kept).
template <int G>
__device__ __forceinline__ void tile_coords(int tiles_m, int tiles_n, int p, int* tmi, int* tni) {
  if constexpr (G <= 1) {
    *tmi = p / tiles_n;
    *tni = p - *tmi * tiles_n;
  } else {
    const int per = G * tiles_n;             // tiles per full group of rows
    const int grp = p / per;
    const int r0 = grp * G;
    const int rows = min(G, tiles_m - r0);   // rows in this group (the last may be short)
    const int q = p - grp * per;
    *tni = q / rows;
    *tmi = r0 + (q - *tni * rows);
  }
}

// The persistent loop of a workgroup: claim units and run them with
// unit(j, q, p) until no unit is left.  With XCD-local C hand-over every unit
// of a queue must run on one XCD, so a workgroup serves the queue of the XCD
// it actually runs on (HW_REG_XCC_ID), not one derived from its index: the
// hardware usually places workgroup i on XCD i % 8, but not always (it was
// seen to differ after other launches), and a misplaced writer would leave C
// in another XCD's L2.  A queue whose XCD received no workgroup is drained by
// the last workgroup to finish its own queue; none of that queue's tiles was
// touched in this launch, so that workgroup reads them coherently from memory.
// The unit body has a single call site (it is inlined once).
template <class F>
__device__ __forceinline__ void run_units(const Launch& L, int* s_unit, F&& unit) {
  int njobs, total;
  launch_dims(L, &njobs, &total);
  const int nq = L.nq;
  int x = L.local_handover ? __builtin_amdgcn_readfirstlane(xcc_id() % nq) : (int)(blockIdx.x % (unsigned)nq);
  const int wq = (int)(blockIdx.x / (unsigned)nq);     // index within the queue (no-ticket launches)
  const int nw = (int)(gridDim.x / (unsigned)nq);      // workgroups per queue (no-ticket launches)
  bool drain = false;                                  // draining the queues of XCDs without workgroups
  int iter = 0;
  while (true) {
    int i;
    if (L.use_ticket) {
      if (threadIdx.x == 0) *s_unit = atomicAdd(L.ws + kWsTicket * x, 1);
      __syncthreads();
      i = __builtin_amdgcn_readfirstlane(*s_unit);
      __syncthreads();
    } else {
      i = wq + iter * nw;
    }
    ++iter;
    int j, q, p;
    if (!decode_unit(L, njobs, total, x, i, &j, &q, &p)) {
      if (!L.local_handover) break;
      if (!drain) {
        if (threadIdx.x == 0) *s_unit = atomicAdd(L.ws + kWsExit, 1) == (int)gridDim.x - 1 ? 1 : 0;
        __syncthreads();
        const int last = __builtin_amdgcn_readfirstlane(*s_unit);
        __syncthreads();
        if (!last) break;
        drain = true;
        x = 0;
      } else if (++x >= nq) {
        break;
      }
      continue;
    }
    unit(__builtin_amdgcn_readfirstlane(j), __builtin_amdgcn_readfirstlane(q), __builtin_amdgcn_readfirstlane(p));
  }
}

template <int BK, int CH>
__global__ void __launch_bounds__(NT) abc_f64_kernel(Launch L) {
  constexpr int B_KPT = TileK<BK>::B_KPT;
  __shared__ double As[2][BK][BM];
  __shared__ double Bs[2][BN][TileK<BK>::BKP];
  __shared__ int s_unit;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;

  run_units(L, &s_unit, [&](int j, int q, int p) __attribute__((always_inline)) {
    const Job& J = job_at(L, j);
    int tmi, tni;
    // grouped tile order for the round kernel (measured: 222 L2 16384^3
    // ABC 134.5 -> 127.3 ms, 333 L1 9600^3 27.35 -> 26.93 ms); row order for
    // the others (the single-term kernel measured no gain)
    tile_coords<CH == kRounds ? 4 : 1>(J.tiles_m, J.tiles_n, p, &tmi, &tni);
    UnitGeo g;
    g.row0 = tmi * BM;
    g.col0 = tni * BN;
    g.arow = 2 * (tid & 63);
    g.acol = tid >> 6;
    if constexpr (TileK<BK>::LINE) {
      g.bcol = tid / TileK<BK>::B_LPC;
      g.bk = 2 * (tid % TileK<BK>::B_LPC);
    } else {
      g.bcol = tid / B_TPC;
      g.bk = B_KPT * (tid % B_TPC);
    }
    g.interior = (J.flags & 3) == 3 && g.row0 + BM <= J.ms && g.col0 + BN <= J.ns;
    g.aoff = g.row0 + g.arow + (long long)g.acol * J.lda;
    g.boff = (long long)(g.col0 + g.bcol) * J.ldb + g.bk;

    // The group's products run back to back on this tile; the workgroup
    // applies their C updates itself, in order, so only the first update of
    // the group waits for the previous group and only the last one signals.
    const bool ordered = J.order_base >= 0;
    int* ctr = L.ws + kWsCounters + (ordered ? J.order_base + p : 0);
    const int q0 = q * J.gsize;
    const int q1 = min(J.nprod, q0 + J.gsize);
    for (int qq = q0; qq < q1; ++qq) {
      Prod P = prod_at(J, qq);
      P.a0 = __builtin_amdgcn_readfirstlane(P.a0);
      P.na = __builtin_amdgcn_readfirstlane(P.na);
      P.b0 = __builtin_amdgcn_readfirstlane(P.b0);
      P.nb = __builtin_amdgcn_readfirstlane(P.nb);
      P.c0 = __builtin_amdgcn_readfirstlane(P.c0);
      P.nc = __builtin_amdgcn_readfirstlane(P.nc);
      P.fringe = __builtin_amdgcn_readfirstlane(P.fringe);
      g.depth = __builtin_amdgcn_readfirstlane(P.fringe ? J.kr : J.ks);

      f64x4 acc[FM][FN];
#pragma unroll
      for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

      if constexpr (CH == kRounds) {
        contract_rounds<BK>(J, P, g, As, Bs, lane, wm, wn, acc);
      } else {
        contract<BK, CH>(J, P, g, As, Bs, lane, wm, wn, wave >> 2, acc);
      }

      if (ordered && q > 0 && qq == q0) wait_turn(ctr, q, L);
      scatter_c(J, P, g.row0, g.col0, lane, wm, wn, acc);
    }
    if (ordered) pass_turn(ctr, q + 1, L);
  });
}

// --- element-wise kernels of the AB and Naive pack types --------------------
//
// Both walk one leaf block (rows x cols, column-major) per grid.y index in
// tiles of kEwRows rows x kEwCols columns: a wave covers 128 consecutive rows
// of one column (two per lane), the 4 waves of a workgroup take every 4th
// column of the tile.
constexpr int kEwThreads = 256;
constexpr int kEwRows = 128;
constexpr int kEwCols = 16;

// Operand sums of the Naive pack type (meta/naive_gen_mix.py:60-137): for
// product r, Y_r = sum_t u_t X(bi_t, bj_t) with the arithmetic of the ABC
// kernel's packing (s = u0*x0, then s = fma(u_t, x_t, s)), so the sums equal
// the ones the ABC kernel forms in LDS bit for bit.  Block (bi, bj) of X sits
// at X + bi*rows + bj*cols*ldx; Y_r is block (r % ny, r / ny) of Y.
struct SumJob {
  const double* X;
  double* Y;
  long long ldx, ldy;
  int rows, cols;      // leaf block extent
  int ny;              // blocks per block column of Y
  int side;            // 0: the A terms of each product, 1: the B terms
  int vec;             // 1: 16-byte accesses allowed (even rows/ld, aligned)
  int skip1;           // 1: leave out products with a single term on this side
                       //    (their product jobs read the operand directly)
  const Prod* prods;
  const Term* terms;
};

__device__ __forceinline__ void ew_tile(int rows, int cols, int* r0, int* c0) {
  const int tr = (rows + kEwRows - 1) / kEwRows;
  const int t = (int)blockIdx.x;
  *r0 = (t % tr) * kEwRows;
  *c0 = (t / tr) * kEwCols;
  (void)cols;
}

__global__ void __launch_bounds__(kEwThreads) presum_kernel(SumJob S) {
  const int r = (int)blockIdx.y;
  int r0, c0;
  ew_tile(S.rows, S.cols, &r0, &c0);
  if (c0 >= S.cols) return;
  const Prod* pp = S.prods + r;
  const int t0 = S.side ? pp->b0 : pp->a0;
  const int nt = S.side ? pp->nb : pp->na;
  if (S.skip1 && nt == 1) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = r0 + 2 * lane;
  gdouble* y = gptr(S.Y) + (long long)(r % S.ny) * S.rows + (long long)(r / S.ny) * S.cols * S.ldy;
  for (int c = c0 + wave; c < min(c0 + kEwCols, S.cols); c += kEwThreads / 64) {
    double s0 = 0.0, s1 = 0.0;
    for (int t = 0; t < nt; ++t) {
      const Term tm = S.terms[t0 + t];
      gcdouble* x = (gcdouble*)S.X + (long long)tm.bi * S.rows + (long long)tm.bj * S.cols * S.ldx +
                    (long long)c * S.ldx;
      double x0 = 0.0, x1 = 0.0;
      if (S.vec && i + 1 < S.rows) {
        const f64x2 v = *(gcf64x2*)(x + i);
        x0 = v.x;
        x1 = v.y;
      } else {
        if (i < S.rows) x0 = x[i];
        if (i + 1 < S.rows) x1 = x[i + 1];
      }
      s0 = t == 0 ? tm.coef * x0 : __builtin_fma(tm.coef, x0, s0);
      s1 = t == 0 ? tm.coef * x1 : __builtin_fma(tm.coef, x1, s1);
    }
    gdouble* yc = y + (long long)c * S.ldy;
    if (S.vec && i + 1 < S.rows) {
      f64x2 v;
      v.x = s0;
      v.y = s1;
      *(__attribute__((address_space(1))) f64x2*)(yc + i) = v;
    } else {
      if (i < S.rows) yc[i] = s0;
      if (i + 1 < S.rows) yc[i + 1] = s1;
    }
  }
}

// The Naive pre-sums of one side for every product in one pass: a
// workgroup stages one column run (kEwRows rows) of every source block of
// the operand (nbr x nbc blocks) into LDS once -- its waves split the blocks
// -- and then writes that run of every product's sum, the waves splitting
// the products.  Each source element is read from HBM once instead of once
// per product term that names it (333: 9 block reads per element instead of
// 49; <2,2,2> x <3,2,3>: 24 instead of 384).  Same arithmetic and term order
// as presum_kernel, so the sums are equal bit for bit.  Dynamic LDS: 1 KiB
// per source block.
constexpr int kSaWaves = 4;
constexpr int kSaBatch = 8;
constexpr int kSaMaxBlocks = 128;
struct SumAllJob {
  const double* X;
  double* Y;
  long long ldx, ldy;
  int rows, cols;      // leaf block extent
  int ny;              // blocks per block column of Y
  int side;            // 0: the A terms of each product, 1: the B terms
  int vec;             // 1: 16-byte accesses allowed (even rows/ld, aligned)
  int skip1;           // 1: leave out products with a single term on this side
  int nbr, nbc;        // source blocks: (bi, bj) with bi < nbr, bj < nbc
  int nprod;           // products
  int pad;
  const Prod* prods;
  const Term* terms;
};

__global__ void __launch_bounds__(64 * kSaWaves) presum_all_kernel(SumAllJob S) {
  extern __shared__ f64x2 sa_lds[];
  const int tr = (S.rows + kEwRows - 1) / kEwRows;
  const int t = (int)blockIdx.x;
  const int r0 = (t % tr) * kEwRows;
  const int c = t / tr;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = r0 + 2 * lane;
  const bool v2 = S.vec && i + 1 < S.rows;
  const int nb = S.nbr * S.nbc;
  // the wave's source blocks, kSaBatch loads in flight at a time (clamped
  // indices past the last block load a valid block, not stored)
  for (int b0 = wave; b0 < nb; b0 += kSaWaves * kSaBatch) {
    f64x2 v[kSaBatch];
#pragma unroll
    for (int q = 0; q < kSaBatch; ++q) {
      const int b = min(b0 + q * kSaWaves, nb - 1);
      const int bi = b / S.nbc, bj = b - bi * S.nbc;
      gcdouble* x = (gcdouble*)S.X + (long long)bi * S.rows + (long long)bj * S.cols * S.ldx + (long long)c * S.ldx;
      v[q] = f64x2{0.0, 0.0};
      if (v2) {
        v[q] = *(gcf64x2*)(x + i);
      } else {
        if (i < S.rows) v[q].x = x[i];
        if (i + 1 < S.rows) v[q].y = x[i + 1];
      }
    }
#pragma unroll
    for (int q = 0; q < kSaBatch; ++q) {
      const int b = b0 + q * kSaWaves;
      if (b < nb) sa_lds[b * 64 + lane] = v[q];
    }
  }
  __syncthreads();
  // product and term records as scalar loads (the stores below would
  // otherwise make the compiler read them as vector loads, each waited for)
  const auto* prods = (const __attribute__((address_space(4))) Prod*)S.prods;
  const auto* terms = (const __attribute__((address_space(4))) Term*)S.terms;
  for (int r = wave; r < S.nprod; r += kSaWaves) {
    const int t0 = S.side ? prods[r].b0 : prods[r].a0;
    const int nt = S.side ? prods[r].nb : prods[r].na;
    if (S.skip1 && nt == 1) continue;
    double s0 = 0.0, s1 = 0.0;
    for (int u = 0; u < nt; ++u) {
      Term tm;
      tm.coef = terms[t0 + u].coef;
      tm.bi = terms[t0 + u].bi;
      tm.bj = terms[t0 + u].bj;
      const f64x2 x = sa_lds[(tm.bi * S.nbc + tm.bj) * 64 + lane];
      s0 = u == 0 ? tm.coef * x.x : __builtin_fma(tm.coef, x.x, s0);
      s1 = u == 0 ? tm.coef * x.y : __builtin_fma(tm.coef, x.y, s1);
    }
    gdouble* yc = gptr(S.Y) + (long long)(r % S.ny) * S.rows + (long long)(r / S.ny) * S.cols * S.ldy +
                  (long long)c * S.ldy;
    if (v2) {
      f64x2 v;
      v.x = s0;
      v.y = s1;
      *(__attribute__((address_space(1))) f64x2*)(yc + i) = v;
    } else {
      if (i < S.rows) yc[i] = s0;
      if (i + 1 < S.rows) yc[i + 1] = s1;
    }
  }
}

// M_Add of the AB and Naive pack types (meta/naive_gen_mix.py:244-361,
// meta/ab_gen_mix.py:341-372): for C block b = I*Nx + J (grid.y),
// C_b = fma(w, M_r, C_b) over the list cptr[b]..cptr[b+1] in product order --
// the C update sequence of the ABC kernel, so the results agree bit for bit.
// M_r is block (r % mx, r / mx) of M; list entries with r >= nlive are skipped.
struct AddJob {
  const double* M;
  double* C;
  long long ldm, ldc;
  int rows, cols;      // leaf block extent (ms x ns)
  int mx, nx;          // C blocks per block column / row
  int nlive;           // products present in M
  int vec;
  const int* cptr;
  const int* cprod;
  const double* cw;
};

__global__ void __launch_bounds__(kEwThreads) madd_kernel(AddJob A) {
  const int b = (int)blockIdx.y;
  int r0, c0;
  ew_tile(A.rows, A.cols, &r0, &c0);
  if (c0 >= A.cols) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = r0 + 2 * lane;
  const int l0 = A.cptr[b], l1 = A.cptr[b + 1];
  gdouble* cb = gptr(A.C) + (long long)(b / A.nx) * A.rows + (long long)(b % A.nx) * A.cols * A.ldc;
  for (int c = c0 + wave; c < min(c0 + kEwCols, A.cols); c += kEwThreads / 64) {
    gdouble* cc = cb + (long long)c * A.ldc;
    const bool v2 = A.vec && i + 1 < A.rows;
    double c0v = 0.0, c1v = 0.0;
    if (v2) {
      const f64x2 v = *(gcf64x2*)(cc + i);
      c0v = v.x;
      c1v = v.y;
    } else {
      if (i < A.rows) c0v = cc[i];
      if (i + 1 < A.rows) c1v = cc[i + 1];
    }
    for (int l = l0; l < l1; ++l) {
      const int r = A.cprod[l];
      if (r >= A.nlive) continue;
      const double w = A.cw[l];
      gcdouble* m = (gcdouble*)A.M + (long long)(r % A.mx) * A.rows + (long long)(r / A.mx) * A.cols * A.ldm +
                    (long long)c * A.ldm;
      double m0 = 0.0, m1 = 0.0;
      if (v2) {
        const f64x2 v = *(gcf64x2*)(m + i);
        m0 = v.x;
        m1 = v.y;
      } else {
        if (i < A.rows) m0 = m[i];
        if (i + 1 < A.rows) m1 = m[i + 1];
      }
      c0v = __builtin_fma(w, m0, c0v);
      c1v = __builtin_fma(w, m1, c1v);
    }
    if (v2) {
      f64x2 v;
      v.x = c0v;
      v.y = c1v;
      *(__attribute__((address_space(1))) f64x2*)(cc + i) = v;
    } else {
      if (i < A.rows) cc[i] = c0v;
      if (i + 1 < A.rows) cc[i + 1] = c1v;
    }
  }
}

// M_Add in one pass over M: a one-wave workgroup keeps one column run
// (128 rows) of every C block in LDS, walks the products in order reading
// each M_r run once (kMaPf products ahead), applies C_b = fma(w, M_r, C_b)
// for each C term (b, w) of product r, and writes the C runs back once.
// Per C block the updates come in product order -- the sequence of
// madd_kernel -- so the results agree bit for bit, while each M_r element
// is read once instead of once per W term (333: 23 M reads per element
// instead of 45; <2,2,2> x <3,2,3>: 105 instead of 360).  Product r's C
// terms are rblk / rw [rptr[r], rptr[r+1]); the k-fringe product R + b
// (present when nlive > R) feeds block b alone, with weight 1.
// Dynamic LDS: 1 KiB per C block.
constexpr int kMaMaxBlocks = 64;
struct AddAllJob {
  const double* M;
  double* C;
  long long ldm, ldc;
  int rows, cols;      // leaf block extent (ms x ns)
  int mx, nx;          // M blocks per block column; C blocks per block row
  int ncb;             // C blocks (Mx * Nx)
  int nlive;           // products present in M (>= 1)
  int vec;
  const int* rptr;
  const int* rblk;
  const double* rw;
};

// VEC: 16-byte accesses (C, M aligned, even leading dimensions; rows, a
// multiple of 8, is even).  Every load is issued unconditionally -- lanes
// past the block and prefetches past the last product read a clamped valid
// address, never used -- so each wave has a fixed number of loads in flight
// and the waits before each use count only the older ones.
template <int kMaPf, bool VEC>
__global__ void __launch_bounds__(64) madd_all_kernel(AddAllJob A) {
  extern __shared__ f64x2 ma_lds[];
  const int tr = (A.rows + 127) / 128;
  const int t = (int)blockIdx.x;
  const int r0 = (t % tr) * 128;
  const int c = t / tr;
  const int lane = threadIdx.x;
  const int i = r0 + 2 * lane;
  const int ia = VEC ? min(i, A.rows - 2) : min(i, A.rows - 1);
  const int ib = min(i + 1, A.rows - 1);
  auto ld2 = [&](gcdouble* p) {
    f64x2 v;
    if (VEC) {
      v = *(gcf64x2*)(p + ia);
    } else {
      v.x = p[ia];
      v.y = p[ib];
    }
    return v;
  };
  auto cblk = [&](int b) {
    return gptr(A.C) + (long long)(b / A.nx) * A.rows + (long long)(b % A.nx) * A.cols * A.ldc +
           (long long)c * A.ldc;
  };
  auto mblk = [&](int r) {
    return (gcdouble*)A.M + (long long)(r % A.mx) * A.rows + (long long)(r / A.mx) * A.cols * A.ldm +
           (long long)c * A.ldm;
  };
  for (int b = 0; b < A.ncb; ++b) ma_lds[b * 64 + lane] = ld2(cblk(b));
  f64x2 mr[kMaPf];
#pragma unroll
  for (int q = 0; q < kMaPf; ++q) mr[q] = ld2(mblk(min(q, A.nlive - 1)));
  for (int rb = 0; rb < A.nlive; rb += kMaPf) {
#pragma unroll
    for (int q = 0; q < kMaPf; ++q) {
      const int r = rb + q;
      const f64x2 m = mr[q];
      mr[q] = ld2(mblk(min(r + kMaPf, A.nlive - 1)));
      if (r < A.nlive) {
        const int l1 = A.rptr[r + 1];
        for (int l = A.rptr[r]; l < l1; ++l) {
          const double w = A.rw[l];
          f64x2* cv = ma_lds + A.rblk[l] * 64 + lane;
          f64x2 x = *cv;
          x.x = __builtin_fma(w, m.x, x.x);
          x.y = __builtin_fma(w, m.y, x.y);
          *cv = x;
        }
      }
    }
  }
  for (int b = 0; b < A.ncb; ++b) {
    const f64x2 v = ma_lds[b * 64 + lane];
    gdouble* p = cblk(b);
    if (VEC && i + 1 < A.rows) {
      *(__attribute__((address_space(1))) f64x2*)(p + i) = v;
    } else {
      if (i < A.rows) p[i] = v.x;
      if (i + 1 < A.rows) p[i + 1] = v.y;
    }
  }
}

// M_Add in one pass over M for many C blocks: a 4-wave workgroup keeps one
// column run of every C block in LDS (shared), and wave w owns the blocks b
// with b % 4 == w: it walks, in product order, only the products with a
// term on its blocks (its own list: wprod, wptr, wblk, ww from wbeg[w]),
// reading those M_r runs (kMaPf ahead) and applying their terms, then
// writes its blocks back.  Each block sees the updates of madd_kernel in
// the same order (bit for bit); against the one-wave kernel the workgroup
// holds four times the waves per CU for the same LDS and each wave a
// quarter of the read-modify-write chain.
constexpr int kMa4Waves = 4;
struct AddAll4Job {
  const double* M;
  double* C;
  long long ldm, ldc;
  int rows, cols;
  int mx, nx;
  int ncb;
  int nlive;
  const int* wbeg;
  const int* wprod;
  const int* wptr;
  const int* wblk;
  const double* ww;
};

template <int kMaPf, bool VEC>
__global__ void __launch_bounds__(64 * kMa4Waves) madd_all4_kernel(AddAll4Job A) {
  extern __shared__ f64x2 ma_lds[];
  const int tr = (A.rows + 127) / 128;
  const int t = (int)blockIdx.x;
  const int r0 = (t % tr) * 128;
  const int c = t / tr;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = r0 + 2 * lane;
  const int ia = VEC ? min(i, A.rows - 2) : min(i, A.rows - 1);
  const int ib = min(i + 1, A.rows - 1);
  auto ld2 = [&](gcdouble* p) {
    f64x2 v;
    if (VEC) {
      v = *(gcf64x2*)(p + ia);
    } else {
      v.x = p[ia];
      v.y = p[ib];
    }
    return v;
  };
  auto cblk = [&](int b) {
    return gptr(A.C) + (long long)(b / A.nx) * A.rows + (long long)(b % A.nx) * A.cols * A.ldc +
           (long long)c * A.ldc;
  };
  auto mblk = [&](int r) {
    return (gcdouble*)A.M + (long long)(r % A.mx) * A.rows + (long long)(r / A.mx) * A.cols * A.ldm +
           (long long)c * A.ldm;
  };
  const auto* wbeg = (const __attribute__((address_space(4))) int*)A.wbeg;
  const auto* wprod = (const __attribute__((address_space(4))) int*)A.wprod;
  const auto* wptr = (const __attribute__((address_space(4))) int*)A.wptr;
  const auto* wblk = (const __attribute__((address_space(4))) int*)A.wblk;
  const auto* ww = (const __attribute__((address_space(4))) double*)A.ww;
  for (int b = wave; b < A.ncb; b += kMa4Waves) ma_lds[b * 64 + lane] = ld2(cblk(b));
  const int e0 = wbeg[wave], n = wbeg[wave + 1] - e0;
  if (n > 0) {
    auto prod_m = [&](int e) { return min(wprod[e0 + min(e, n - 1)], A.nlive - 1); };
    f64x2 mr[kMaPf];
#pragma unroll
    for (int q = 0; q < kMaPf; ++q) mr[q] = ld2(mblk(prod_m(q)));
    for (int eb = 0; eb < n; eb += kMaPf) {
#pragma unroll
      for (int q = 0; q < kMaPf; ++q) {
        const int e = eb + q;
        const f64x2 m = mr[q];
        mr[q] = ld2(mblk(prod_m(e + kMaPf)));
        if (e < n && wprod[e0 + e] < A.nlive) {
          const int l1 = wptr[e0 + e + 1];
          for (int l = wptr[e0 + e]; l < l1; ++l) {
            const double w = ww[l];
            const int ix = wblk[l] * 64 + lane;
            f64x2 x = ma_lds[ix];
            x.x = __builtin_fma(w, m.x, x.x);
            x.y = __builtin_fma(w, m.y, x.y);
            ma_lds[ix] = x;
          }
        }
      }
    }
  }
  for (int b = wave; b < A.ncb; b += kMa4Waves) {
    const f64x2 v = ma_lds[b * 64 + lane];
    gdouble* p = cblk(b);
    if (VEC && i + 1 < A.rows) {
      *(__attribute__((address_space(1))) f64x2*)(p + i) = v;
    } else {
      if (i < A.rows) p[i] = v.x;
      if (i + 1 < A.rows) p[i + 1] = v.y;
    }
  }
}

// Records the XCC id of each workgroup (one launch per device at set-up):
// the host enables XCD-local hand-over only if the ids look like an 8-XCD part
// (workgroup i on XCC i % 8 for a fresh launch).
__global__ void xcc_probe_kernel(int* out) {
  if (threadIdx.x == 0) out[blockIdx.x] = xcc_id();
}

// Zero the ticket, error flag and order counters of a launch (a plain fill
// kernel: it can be captured in a graph, unlike a memset node).
__global__ void ws_clear_kernel(int* ws, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ws[i] = 0;
}

// dst |= error word of workspace ws (one thread): carries the flags of a
// launch whose workspace is about to be released, or of a sub-call, to the
// word the caller reads.
__global__ void err_fold_kernel(int* dst, const int* ws) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    const int v = __hip_atomic_load(ws + kWsError, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v != 0) atomicOr(dst, v);
  }
}

// *dst = error word of ws (one thread); dst may be host-mapped pinned memory,
// read by the host once the launch is known to be complete.
__global__ void err_copy_kernel(int* dst, const int* ws) {
  if (blockIdx.x == 0 && threadIdx.x == 0)
    __hip_atomic_store(dst, __hip_atomic_load(ws + kWsError, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

}  // namespace fmm
