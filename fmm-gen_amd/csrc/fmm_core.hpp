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
//     128x128 FP64 accumulator tile (4 waves x 64x64), computed transposed
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

struct Job {
  const double* A;
  const double* B;
  double* C;
  long long lda, ldb, ldc;
  int ms, ns, ks, kr;       // leaf block and k-fringe depth
  int tiles_m, tiles_n;     // C tiles over one leaf block
  int nprod;                // products run by this job (R [+ Mx*Nx fringe])
  int unit_base;            // first unit of this job in the launch
  int order_base;           // first order counter, <0 when nprod == 1
  int flags;                // bit0/1/2: A/B/C allow 16-byte vector access
  const Prod* prods;
  const Term* terms;
};

constexpr int kInlineJobs = 4;
constexpr int kThreads = 256;

struct Launch {
  int njobs;
  int total_units;
  int use_ticket;           // 1: ws[0] is the unit ticket (ordered launches)
  int pad;
  const Job* jobs;          // device job table, or nullptr -> inl[]
  int* ws;                  // [0] ticket, [1] error flags, [2..] order counters
  long long spin_limit;     // bounded wait, in wall_clock64() ticks
  Job inl[kInlineJobs];
};

enum : int { kErrSpin = 1 };

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Tile geometry of the FP64 kernel.
constexpr int BM = 128;
constexpr int BN = 128;
constexpr int BK = 16;
constexpr int BKP = BK + 1;   // padded k stride of the B tile in LDS

__device__ __forceinline__ const Job& job_at(const Launch& L, int j) {
  return L.jobs ? L.jobs[j] : L.inl[j];
}

// Largest j with unit_base <= u (jobs are sorted by unit_base).
__device__ __forceinline__ int find_job(const Launch& L, int u) {
  int lo = 0, hi = L.njobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (job_at(L, mid).unit_base <= u) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// ---------------------------------------------------------------------------
// Global -> register staging of one BK-deep slice for up to CH source blocks.
//
// A slice (BM x BK, column-major source): thread t owns rows 2*(t%64)+{0,1}
// of columns t/64 + 4*i, i < BK/4  -> each wave reads one 1 KiB column run.
// B slice (BK x BN, column-major source): thread t owns column t/2 and the
// 8 consecutive k of half t%2 -> each wave reads 32 full 128-byte runs.
// ---------------------------------------------------------------------------
struct Slice {
  double a[BK / 2];   // 8 values of the A sum
  double b[BK / 2];   // 8 values of the B sum
};

template <int CH>
struct Raw {
  double a[CH][BK / 2];
  double b[CH][BK / 2];
};

struct UnitGeo {
  int row0;        // first leaf row of this tile
  int col0;        // first leaf column of this tile
  int depth;       // k extent (ks or kr)
  int arow;        // this thread's A row (even) within the tile
  int acol;        // this thread's first A column within the slice
  int bcol;        // this thread's B column within the tile
  int bk;          // this thread's first B k within the slice
};

__device__ __forceinline__ const double* a_block(const Job& J, const Term& t) {
  return J.A + (long long)t.bi * J.ms + (long long)t.bj * J.ks * J.lda;
}
__device__ __forceinline__ const double* b_block(const Job& J, const Term& t) {
  return J.B + (long long)t.bi * J.ks + (long long)t.bj * J.ns * J.ldb;
}

template <int CH>
__device__ __forceinline__ void load_raw(const Job& J, const Prod& P, int ta0, int na,
                                         int tb0, int nb, int kb, const UnitGeo& g,
                                         Raw<CH>& r) {
  const bool vecA = (J.flags & 1) != 0;
  const bool vecB = (J.flags & 2) != 0;
  const int row = g.row0 + g.arow;
  const bool r0ok = row < J.ms, r1ok = row + 1 < J.ms;
#pragma unroll
  for (int t = 0; t < CH; ++t) {
    if (t < na) {
      const Term tm = J.terms[ta0 + t];
      const double* src = a_block(J, tm) + row;
#pragma unroll
      for (int i = 0; i < BK / 4; ++i) {
        const int col = kb + g.acol + 4 * i;
        const bool cok = col < g.depth;
        const double* p = src + (long long)col * J.lda;
        if (vecA && cok && r1ok) {
          const double2 v = *reinterpret_cast<const double2*>(p);
          r.a[t][2 * i] = v.x;
          r.a[t][2 * i + 1] = v.y;
        } else {
          r.a[t][2 * i] = (cok && r0ok) ? p[0] : 0.0;
          r.a[t][2 * i + 1] = (cok && r1ok) ? p[1] : 0.0;
        }
      }
    }
    if (t < nb) {
      const Term tm = J.terms[tb0 + t];
      const int col = g.col0 + g.bcol;
      const bool cok = col < J.ns;
      const double* src = b_block(J, tm) + (long long)col * J.ldb + kb + g.bk;
      const int krem = g.depth - (kb + g.bk);   // valid k in this run
      if (vecB && cok && krem >= BK / 2) {
#pragma unroll
        for (int e = 0; e < BK / 2; e += 2) {
          const double2 v = *reinterpret_cast<const double2*>(src + e);
          r.b[t][e] = v.x;
          r.b[t][e + 1] = v.y;
        }
      } else {
#pragma unroll
        for (int e = 0; e < BK / 2; ++e) r.b[t][e] = (cok && e < krem) ? src[e] : 0.0;
      }
    }
  }
}

// s = u0*x0, then s = s + u_t*x_t left to right (fma == add of the exact
// product for the dyadic table coefficients).
template <int CH>
__device__ __forceinline__ void accumulate(const Job& J, int ta0, int na, int tb0, int nb,
                                           bool first, const Raw<CH>& r, Slice& s) {
#pragma unroll
  for (int t = 0; t < CH; ++t) {
    if (t < na) {
      const double u = J.terms[ta0 + t].coef;
#pragma unroll
      for (int e = 0; e < BK / 2; ++e)
        s.a[e] = (first && t == 0) ? u * r.a[t][e] : __builtin_fma(u, r.a[t][e], s.a[e]);
    }
    if (t < nb) {
      const double v = J.terms[tb0 + t].coef;
#pragma unroll
      for (int e = 0; e < BK / 2; ++e)
        s.b[e] = (first && t == 0) ? v * r.b[t][e] : __builtin_fma(v, r.b[t][e], s.b[e]);
    }
  }
}

// Build the full sums for slice kb: `r` already holds chunk 0 (issued
// earlier); further chunks are loaded here.
template <int CH>
__device__ __forceinline__ void finish_sums(const Job& J, const Prod& P, int kb, const UnitGeo& g,
                                            Raw<CH>& r, Slice& s) {
  accumulate<CH>(J, P.a0, P.na, P.b0, P.nb, true, r, s);
  const int nmax = P.na > P.nb ? P.na : P.nb;
  for (int c = CH; c < nmax; c += CH) {
    const int na = P.na - c, nb = P.nb - c;
    load_raw<CH>(J, P, P.a0 + c, na, P.b0 + c, nb, kb, g, r);
    accumulate<CH>(J, P.a0 + c, na, P.b0 + c, nb, false, r, s);
  }
}

__device__ __forceinline__ void store_slice(double (*As)[BM], double (*Bs)[BKP], const UnitGeo& g,
                                            const Slice& s) {
#pragma unroll
  for (int i = 0; i < BK / 4; ++i) {
    double2 v;
    v.x = s.a[2 * i];
    v.y = s.a[2 * i + 1];
    *reinterpret_cast<double2*>(&As[g.acol + 4 * i][g.arow]) = v;
  }
#pragma unroll
  for (int e = 0; e < BK / 2; ++e) Bs[g.bcol][g.bk + e] = s.b[e];
}

template <int KK0, int KK1>
__device__ __forceinline__ void mfma_slice(const double (*As)[BM], const double (*Bs)[BKP],
                                           int lane, int wm, int wn, f64x4 (&acc)[4][4]) {
  const int l16 = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = KK0; kk < KK1; ++kk) {
    double a[4], b[4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) a[mi] = As[kk * 4 + lk][wm * 64 + mi * 16 + l16];
#pragma unroll
    for (int ni = 0; ni < 4; ++ni) b[ni] = Bs[wn * 64 + ni * 16 + l16][kk * 4 + lk];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ni], a[mi], acc[mi][ni], 0, 0, 0);
  }
}

// Wait (thread 0, bounded) until counter reaches `want`; then every thread
// acquires at agent scope so C reads see the previous product's writes.
__device__ __forceinline__ void wait_turn(int* ctr, int want, const Launch& L) {
  if (threadIdx.x == 0) {
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(ctr, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want) {
      __builtin_amdgcn_s_sleep(4);
      // A failed wait anywhere in the launch ends every later wait at once,
      // so a broken ordering costs one time bound, not one per unit.
      if (__hip_atomic_load(L.ws + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
      if (wall_clock64() - t0 > L.spin_limit) {
        atomicOr(L.ws + 1, (int)kErrSpin);
        break;
      }
    }
  }
  __syncthreads();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

__device__ __forceinline__ void pass_turn(int* ctr, int next) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_store(ctr, next, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// C_j = C_j + w_j * acc for every C term of the product.
//
// v_mfma_f64_16x16x4f64 computes D[i][j] = sum_k X[i][k] Y[k][j] with X[i][k]
// in lane i+16k, Y[k][j] in lane j+16k, and element v of lane l holding
// D[l/16 + 4v][l%16] (layout measured on gfx950).
// The contraction feeds X = B^T and Y = A^T, so D = (A B)^T: element v of
// lane l is C row l%16, column l/16 + 4v of the 16x16 block.  For each v a
// wave then touches 4 columns x 16 consecutive rows: four full 128-byte runs
// of the column-major C.
__device__ __forceinline__ void scatter_c(const Job& J, const Prod& P, int row0, int col0, int lane,
                                          int wm, int wn, const f64x4 (&acc)[4][4]) {
  const int l16 = lane & 15, lk = lane >> 4;
  for (int t = 0; t < P.nc; ++t) {
    const Term tm = J.terms[P.c0 + t];
    const double w = tm.coef;
    double* cb = J.C + (long long)tm.bi * J.ms + (long long)tm.bj * J.ns * J.ldc;
#pragma unroll
    for (int mi = 0; mi < 4; ++mi) {
      const int row = row0 + wm * 64 + mi * 16 + l16;
      if (row >= J.ms) continue;
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) {
        const f64x4 v = acc[mi][ni];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int col = col0 + wn * 64 + ni * 16 + lk + 4 * e;
          if (col < J.ns) {
            double* c = cb + row + (long long)col * J.ldc;
            *c = __builtin_fma(w, v[e], *c);
          }
        }
      }
    }
  }
}

// Contraction of one unit over its whole depth into acc.
//   DB = false: one LDS buffer; the next slice's loads fly during the MFMAs,
//               then sum + store + two barriers per slice.
//   DB = true:  two LDS buffers; the next slice is summed and stored between
//               the two halves of the current slice's MFMAs, one barrier per
//               slice, loads issued one slice ahead.
template <int CH, bool DB>
__device__ __forceinline__ void contract(const Job& J, const Prod& P, const UnitGeo& g, double (*As)[BK][BM],
                                         double (*Bs)[BN][BKP], int lane, int wm, int wn, f64x4 (&acc)[4][4]) {
  const int nsteps = (g.depth + BK - 1) / BK;
  if (!(P.na > 0 && P.nb > 0 && nsteps > 0)) return;
  Raw<CH> r;
  Slice s;
  load_raw<CH>(J, P, P.a0, P.na, P.b0, P.nb, 0, g, r);
  finish_sums<CH>(J, P, 0, g, r, s);
  store_slice(As[0], Bs[0], g, s);
  if (!DB) {
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
      const bool more = st + 1 < nsteps;
      const int kb_next = (st + 1) * BK;
      if (more) load_raw<CH>(J, P, P.a0, P.na, P.b0, P.nb, kb_next, g, r);
      mfma_slice<0, BK / 4>(As[0], Bs[0], lane, wm, wn, acc);
      __syncthreads();
      if (more) {
        finish_sums<CH>(J, P, kb_next, g, r, s);
        store_slice(As[0], Bs[0], g, s);
        __syncthreads();
      }
    }
  } else {
    if (nsteps > 1) load_raw<CH>(J, P, P.a0, P.na, P.b0, P.nb, BK, g, r);
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) {
      const int cur = st & 1;
      mfma_slice<0, BK / 8>(As[cur], Bs[cur], lane, wm, wn, acc);
      if (st + 1 < nsteps) {
        finish_sums<CH>(J, P, (st + 1) * BK, g, r, s);
        store_slice(As[cur ^ 1], Bs[cur ^ 1], g, s);
        if (st + 2 < nsteps) load_raw<CH>(J, P, P.a0, P.na, P.b0, P.nb, (st + 2) * BK, g, r);
      }
      mfma_slice<BK / 8, BK / 4>(As[cur], Bs[cur], lane, wm, wn, acc);
      __syncthreads();
    }
  }
}

template <int CH, bool DB>
__global__ void __launch_bounds__(kThreads) abc_f64_kernel(Launch L) {
  __shared__ double As[DB ? 2 : 1][BK][BM];
  __shared__ double Bs[DB ? 2 : 1][BN][BKP];
  __shared__ int s_unit;

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;

  int iter = 0;
  while (true) {
    int u;
    if (L.use_ticket) {
      if (tid == 0) s_unit = atomicAdd(L.ws, 1);
      __syncthreads();
      u = __builtin_amdgcn_readfirstlane(s_unit);
      __syncthreads();
    } else {
      u = blockIdx.x + iter * gridDim.x;
    }
    ++iter;
    if (u >= L.total_units) break;

    const Job& J = job_at(L, find_job(L, u));
    const int ul = u - J.unit_base;
    const int ntile = J.tiles_m * J.tiles_n;
    const int q = ul / ntile;
    const int p = ul - q * ntile;
    const int tmi = p / J.tiles_n;
    const int tni = p - tmi * J.tiles_n;
    const Prod P = J.prods[q];

    UnitGeo g;
    g.row0 = tmi * BM;
    g.col0 = tni * BN;
    g.depth = P.fringe ? J.kr : J.ks;
    g.arow = 2 * (tid & 63);
    g.acol = tid >> 6;
    g.bcol = tid >> 1;
    g.bk = (BK / 2) * (tid & 1);

    f64x4 acc[4][4];
#pragma unroll
    for (int mi = 0; mi < 4; ++mi)
#pragma unroll
      for (int ni = 0; ni < 4; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

    contract<CH, DB>(J, P, g, As, Bs, lane, wm, wn, acc);

    const bool ordered = J.order_base >= 0;
    int* ctr = L.ws + 2 + (ordered ? J.order_base + p : 0);
    if (ordered && q > 0) wait_turn(ctr, q, L);
    scatter_c(J, P, g.row0, g.col0, lane, wm, wn, acc);
    if (ordered) pass_turn(ctr, q + 1);
  }
}

// Zero the ticket, error flag and order counters of a launch (a plain fill
// kernel: it can be captured in a graph, unlike a memset node).
__global__ void ws_clear_kernel(int* ws, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) ws[i] = 0;
}

}  // namespace fmm
