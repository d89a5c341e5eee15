// fmm_cs.hpp -- C-stationary FP64 kernel for one-level algorithms.
//
// The ordered kernel (abc_f64_kernel) runs one product per unit and
// read-modify-writes every C tile the product feeds, in product order across
// workgroups: 2 * nnz(W) passes over C blocks through HBM.  When the leaf is
// shallow (rank-k updates such as 14400^2 x 480) those passes cost as much as
// the products themselves, and with one workgroup per CU they do not overlap
// its MFMAs.  Here one workgroup owns one tile *position* of the leaf grid for
// the whole call: it loads that tile of all Mx*Nx C blocks into registers,
// runs every product (then the k-fringe products) over it, applies
// C_j = fma(w_jr, M_r, C_j) from the product's accumulators into the
// register-resident C tiles, and writes each C tile once.  No ordering, no
// workspace.  The updates are the ordered kernel's, in the same order, with
// the same MFMA chains and operand sums, so the results are identical bit
// for bit.  The reference applies the same updates per product
// (meta/abc_gen_mix.py:302-391, C_j += w_j*AB in the micro-kernel,
// meta/abc_micro_kernel_gen.py:147-196).
//
// The price is operand traffic: the position tile is smaller (all Mx*Nx C
// tiles must fit in registers: 4 blocks of 128 x 64, or 9 blocks of 64 x 64)
// and every product reads its own source panels.  So the kernel serves
// shallow leaves of the Naive pack type, whose products read one pre-summed
// S_r / T_r panel each (single-term products), not nnz(U_r) + nnz(V_r)
// panels as the ABC operand sums would.
//
// Product terms: a term with bi >= 0 is block (bi, bj) of A (B); bi < 0 names
// workspace block bj, i.e. S_bj = S + bj*ks*lds (T_bj = T + bj*ns*ldt) of the
// Naive pre-sums.  Terms carry their coefficient (the single-term kernel
// scales in place when it is not 1, as the pre-sum would).
#pragma once
#include "fmm_core.hpp"

namespace fmm {

struct CsJob {
  const double* A;
  const double* B;
  double* C;
  const double* S;          // Naive pre-sums (nullptr when no term names them)
  const double* T;
  long long lda, ldb, ldc, lds, ldt;
  int ms, ns, ks, kr;       // leaf block and k-fringe depth
  int tiles_m, tiles_n;     // position tiles over one leaf block
  int nprod;                // products, the k-fringe ones last
  int mx, nx;               // C blocks: mx x nx
  int vec;                  // bit0/1/2: A/B/C allow 16-byte accesses
  int nmain;                // products of the leaf depth ks (the k-fringe ones follow)
  int nterms;               // entries of the term table
  int pad;
  const Prod* prods;
  const Term* terms;
};

// Tile geometry by C-block count: 512 threads (8 waves, 2 per SIMD); wave
// tiles of 32 x 32 (<= 4 C blocks: 4 x 32 = 128 accumulator VGPRs) or
// 32 x 16 (<= 9 blocks: 9 x 16 = 144), plus the product's own accumulators.
//
// Measured variants that did not pay (9600^3 and 14400^2 x 480, Naive):
// reading the next k-step's fragments during the MFMAs (-1..-8%), the two
// waves of a SIMD in opposite phase order (-1..-7%), and 4 waves with one per
// SIMD, 64 x 32 wave tiles and the C tiles in AGPRs (-20..-30%).  The
// 32 x 32 / 32 x 16 wave tiles the register budget allows run the slice loop
// at ~40 TFLOP/s against ~62 for the ordered kernel's 64 x 32, which caps
// this kernel to shallow leaves, where the C traffic it saves dominates.
template <int NCB_, int BK_, int CH_>
struct CsCfg {
  static constexpr int NCB = NCB_, BK = BK_, CH = CH_;
  static constexpr int NT = 512;
  static constexpr int NWAVES = NT / 64;
  static constexpr int TM = NCB <= 4 ? 128 : 64;
  static constexpr int TN = 64;
  static constexpr int WAVES_M = TM / 32;
  static constexpr int WAVES_N = NWAVES / WAVES_M;
  static constexpr int WTM = TM / WAVES_M;
  static constexpr int WTN = TN / WAVES_N;
  static constexpr int FM = WTM / 16, FN = WTN / 16;
  // A slice (TM x BK, column-major): TPC_A threads per column, 2 rows each.
  static constexpr int TPC_A = TM / 2;
  static constexpr int CPP_A = NT / TPC_A;
  static constexpr int A_IT = BK / CPP_A;
  static constexpr int AE = 2 * A_IT;
  // B slice (BK x TN): B_TPC threads per column, B_KPT consecutive k each.
  static constexpr int B_TPC = NT / TN;
  static constexpr int B_KPT = BK / B_TPC;
  static constexpr int BE = B_KPT;
  static constexpr int BKP = BK + 1;
  static_assert(WAVES_M * WAVES_N == NWAVES && FM >= 1 && FN >= 1, "wave layout");
  static_assert(CH == 1, "single-term products only");
  static_assert(A_IT >= 1 && CPP_A * A_IT == BK && B_KPT >= 2 && B_KPT % 2 == 0 && B_TPC * B_KPT == BK,
                "staging layout");
};

// Uniform description of one product's sources for one chunk of terms.
template <int CH>
struct CsSrcs {
  gcdouble* a[CH];
  gcdouble* b[CH];
  long long la[CH], lb[CH];   // leading dimension of each source
  double u[CH];
  double v[CH];
  int na, nb;
};

__device__ __forceinline__ void cs_a_src(const CsJob& J, const Term& t, gcdouble** p, long long* ld) {
  if (t.bi < 0) {
    *p = unip(gptr(J.S) + (long long)t.bj * J.ks * J.lds);
    *ld = uni64(J.lds);
  } else {
    *p = unip(gptr(J.A) + (long long)t.bi * J.ms + (long long)t.bj * J.ks * J.lda);
    *ld = uni64(J.lda);
  }
}
__device__ __forceinline__ void cs_b_src(const CsJob& J, const Term& t, gcdouble** p, long long* ld) {
  if (t.bi < 0) {
    *p = unip(gptr(J.T) + (long long)t.bj * J.ns * J.ldt);
    *ld = uni64(J.ldt);
  } else {
    *p = unip(gptr(J.B) + (long long)t.bi * J.ks + (long long)t.bj * J.ns * J.ldb);
    *ld = uni64(J.ldb);
  }
}

__device__ __forceinline__ Term cs_term(const CsJob& J, int i) {
  const Term* t = J.terms + i;
  Term r;
  r.coef = unid(*(gcdouble*)&t->coef);
  r.bi = __builtin_amdgcn_readfirstlane(*(gcint*)&t->bi);
  r.bj = __builtin_amdgcn_readfirstlane(*(gcint*)&t->bj);
  return r;
}

__device__ __forceinline__ Prod cs_prod(const CsJob& J, int q) {
  gcint* p = (gcint*)(J.prods + q);
  Prod r;
  r.a0 = __builtin_amdgcn_readfirstlane(p[0]);
  r.na = __builtin_amdgcn_readfirstlane(p[1]);
  r.b0 = __builtin_amdgcn_readfirstlane(p[2]);
  r.nb = __builtin_amdgcn_readfirstlane(p[3]);
  r.c0 = __builtin_amdgcn_readfirstlane(p[4]);
  r.nc = __builtin_amdgcn_readfirstlane(p[5]);
  r.fringe = __builtin_amdgcn_readfirstlane(p[6]);
  r.pad = 0;
  return r;
}

template <int CH>
__device__ __forceinline__ void cs_make_srcs(const CsJob& J, int ta0, int na, int tb0, int nb, CsSrcs<CH>& S) {
  S.na = na < CH ? na : CH;
  S.nb = nb < CH ? nb : CH;
#pragma unroll
  for (int t = 0; t < CH; ++t) {
    if (t < S.na) {
      const Term tm = cs_term(J, ta0 + t);
      cs_a_src(J, tm, &S.a[t], &S.la[t]);
      S.u[t] = tm.coef;
    } else {
      S.a[t] = nullptr;
      S.la[t] = 0;
      S.u[t] = 0.0;
    }
    if (t < S.nb) {
      const Term tm = cs_term(J, tb0 + t);
      cs_b_src(J, tm, &S.b[t], &S.lb[t]);
      S.v[t] = tm.coef;
    } else {
      S.b[t] = nullptr;
      S.lb[t] = 0;
      S.v[t] = 0.0;
    }
  }
}

// Per-thread position within a tile (fixed for the kernel).
struct CsThread {
  int arow;     // A: first of two rows
  int acol;     // A: first column of the slice
  int bcol;     // B: column
  int bk;       // B: first k of the run
};

template <class Cfg>
struct CsRaw {
  double a[Cfg::CH][Cfg::AE];
  double b[Cfg::CH][Cfg::BE];
};

// Loads of slice kb (depth `depth`) of the chunk's sources for the position
// at (row0, col0).  Interior full slices use unpredicated 16-byte loads;
// everything else is element-wise with zero fill outside the leaf block.
template <class Cfg>
__device__ __forceinline__ void cs_load(const CsJob& J, const CsSrcs<Cfg::CH>& S, int kb, int depth, int row0,
                                        int col0, bool interior, const CsThread& th, CsRaw<Cfg>& r) {
  constexpr int A_IT = Cfg::A_IT, B_KPT = Cfg::B_KPT, CPP_A = Cfg::CPP_A;
  if (interior && kb + Cfg::BK <= depth) {
#pragma unroll
    for (int t = 0; t < Cfg::CH; ++t) {
      if (t < S.na) {
        gcdouble* base = S.a[t] + (row0 + th.arow) + (long long)(kb + th.acol) * S.la[t];
#pragma unroll
        for (int i = 0; i < A_IT; ++i) {
          const f64x2 v = *(gcf64x2*)(base + (long long)(CPP_A * i) * S.la[t]);
          r.a[t][2 * i] = v.x;
          r.a[t][2 * i + 1] = v.y;
        }
      }
      if (t < S.nb) {
        gcdouble* base = S.b[t] + (kb + th.bk) + (long long)(col0 + th.bcol) * S.lb[t];
#pragma unroll
        for (int e = 0; e < B_KPT; e += 2) {
          const f64x2 v = *(gcf64x2*)(base + e);
          r.b[t][e] = v.x;
          r.b[t][e + 1] = v.y;
        }
      }
    }
    return;
  }
  const int row = row0 + th.arow;
  const bool r0ok = row < J.ms, r1ok = row + 1 < J.ms;
  const bool bcok = col0 + th.bcol < J.ns;
  const int krem = depth - (kb + th.bk);
#pragma unroll
  for (int t = 0; t < Cfg::CH; ++t) {
    if (t < S.na) {
#pragma unroll
      for (int i = 0; i < A_IT; ++i) {
        const int col = kb + th.acol + CPP_A * i;
        const bool cok = col < depth;
        gcdouble* p = S.a[t] + row + (long long)col * S.la[t];
        r.a[t][2 * i] = (cok && r0ok) ? p[0] : 0.0;
        r.a[t][2 * i + 1] = (cok && r1ok) ? p[1] : 0.0;
      }
    }
    if (t < S.nb) {
      gcdouble* p = S.b[t] + (kb + th.bk) + (long long)(col0 + th.bcol) * S.lb[t];
#pragma unroll
      for (int e = 0; e < B_KPT; ++e) r.b[t][e] = (bcok && e < krem) ? p[e] : 0.0;
    }
  }
}

template <class Cfg>
__device__ __forceinline__ void cs_store(double (*As)[Cfg::TM], double (*Bs)[Cfg::BKP], const CsThread& th,
                                         const double (&sa)[Cfg::AE], const double (&sb)[Cfg::BE]) {
#pragma unroll
  for (int i = 0; i < Cfg::A_IT; ++i) {
    double2 v;
    v.x = sa[2 * i];
    v.y = sa[2 * i + 1];
    *reinterpret_cast<double2*>(&As[th.acol + Cfg::CPP_A * i][th.arow]) = v;
  }
#pragma unroll
  for (int e = 0; e < Cfg::B_KPT; ++e) Bs[th.bcol][th.bk + e] = sb[e];
}

// Slice kb of product P (loaded into r) into LDS, scaled by the product's
// coefficients where they are not 1 (the pre-sum arithmetic for one term).
template <class Cfg>
__device__ __forceinline__ void cs_pack(const CsSrcs<Cfg::CH>& S0, const CsThread& th, CsRaw<Cfg>& r,
                                        double (*As)[Cfg::TM], double (*Bs)[Cfg::BKP]) {
  if (S0.u[0] != 1.0) {
#pragma unroll
    for (int e = 0; e < Cfg::AE; ++e) r.a[0][e] = S0.u[0] * r.a[0][e];
  }
  if (S0.v[0] != 1.0) {
#pragma unroll
    for (int e = 0; e < Cfg::BE; ++e) r.b[0][e] = S0.v[0] * r.b[0][e];
  }
  cs_store<Cfg>(As, Bs, th, r.a[0], r.b[0]);
}

// MFMAs of k-steps [KK0, KK1) of the slice into the product accumulators.
// Operands as in the ordered kernel (X = B^T, Y = A^T), so element v of lane
// l of block (mi, ni) is tile row wm*WTM + mi*16 + l%16, column
// wn*WTN + ni*16 + l/16 + 4v, and every element's MFMA chain is the same.
template <class Cfg, int KK0, int KK1>
__device__ __forceinline__ void cs_mfma(const double (*As)[Cfg::TM], const double (*Bs)[Cfg::BKP], int lane, int wm,
                                        int wn, f64x4 (&acc)[Cfg::FM][Cfg::FN]) {
  const int l16 = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = KK0; kk < KK1; ++kk) {
    double a[Cfg::FM], b[Cfg::FN];
#pragma unroll
    for (int mi = 0; mi < Cfg::FM; ++mi) a[mi] = As[kk * 4 + lk][wm * Cfg::WTM + mi * 16 + l16];
#pragma unroll
    for (int ni = 0; ni < Cfg::FN; ++ni) b[ni] = Bs[wn * Cfg::WTN + ni * 16 + l16][kk * 4 + lk];
#pragma unroll
    for (int mi = 0; mi < Cfg::FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < Cfg::FN; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[ni], a[mi], acc[mi][ni], 0, 0, 0);
  }
}

// A (product, slice) step of a position's slice stream.
struct CsStep {
  int r;        // product index (r >= nprod: past the end)
  int st;       // slice within the product
};

__device__ __forceinline__ int cs_depth(const CsJob& J, const Prod& P) { return P.fringe ? J.kr : J.ks; }

template <class Cfg>
__global__ void __launch_bounds__(Cfg::NT) cs_f64_kernel(CsJob J, int nq) {
  constexpr int BK = Cfg::BK, FM = Cfg::FM, FN = Cfg::FN, NCB = Cfg::NCB;
  __shared__ double As[2][BK][Cfg::TM];
  __shared__ double Bs[2][Cfg::TN][Cfg::BKP];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % Cfg::WAVES_M, wn = wave / Cfg::WAVES_M;
  const int l16 = lane & 15, lk = lane >> 4;
  CsThread th;
  th.arow = 2 * (tid % Cfg::TPC_A);
  th.acol = tid / Cfg::TPC_A;
  th.bcol = tid / Cfg::B_TPC;
  th.bk = Cfg::B_KPT * (tid % Cfg::B_TPC);

  // Positions: split into nq contiguous chunks (one per XCD when nq = 8; a
  // chunk's consecutive positions share operand panels in that XCD's L2),
  // each served by the workgroups x, x + nq, ... in round robin.
  const int units = J.tiles_m * J.tiles_n;
  const int x = (int)(blockIdx.x % (unsigned)nq);
  const int w = (int)(blockIdx.x / (unsigned)nq);
  const int nw = (int)(gridDim.x / (unsigned)nq);
  const int lo = (int)(((long long)x * units) / nq), hi = (int)(((long long)(x + 1) * units) / nq);
  const int ncb = J.mx * J.nx;

  for (int u = lo + w; u < hi; u += nw) {
    int tmi, tni;
    tile_coords<4>(J.tiles_m, J.tiles_n, u, &tmi, &tni);
    const int row0 = tmi * Cfg::TM, col0 = tni * Cfg::TN;
    const bool interior = (J.vec & 3) == 3 && row0 + Cfg::TM <= J.ms && col0 + Cfg::TN <= J.ns;
    const int crow = row0 + wm * Cfg::WTM + l16;
    const int ccol = col0 + wn * Cfg::WTN + lk;

    // This position's tile of every C block, into registers.
    f64x4 cacc[NCB][FM][FN];
#pragma unroll
    for (int j = 0; j < NCB; ++j) {
      if (j < ncb) {
        gcdouble* cb = gptr(static_cast<const double*>(J.C)) + (long long)(j / J.nx) * J.ms +
                       (long long)(j % J.nx) * J.ns * J.ldc;
#pragma unroll
        for (int mi = 0; mi < FM; ++mi)
#pragma unroll
          for (int ni = 0; ni < FN; ++ni)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int row = crow + mi * 16, col = ccol + ni * 16 + 4 * e;
              cacc[j][mi][ni][e] = (row < J.ms && col < J.ns) ? cb[row + (long long)col * J.ldc] : 0.0;
            }
      }
    }

    // The slice stream over all products: `cur` is contracted from LDS while
    // `nxt` is summed into the other buffer and the loads of `ld` are issued.
    auto nslices = [&](const Prod& P) { return (cs_depth(J, P) + BK - 1) / BK; };
    auto advance = [&](CsStep s, Prod& P) {
      // successor of step s of product P (P updated when the product changes)
      if (s.st + 1 < nslices(P)) return CsStep{s.r, s.st + 1};
      int r = s.r + 1;
      while (r < J.nprod) {
        P = cs_prod(J, r);
        if (P.na > 0 && P.nb > 0 && nslices(P) > 0) break;
        ++r;
      }
      return CsStep{r, 0};
    };
    CsStep cur{0, 0};
    Prod Pc = cs_prod(J, 0);
    if (!(Pc.na > 0 && Pc.nb > 0 && nslices(Pc) > 0)) cur = advance(CsStep{0, 1 << 30}, Pc);
    if (cur.r >= J.nprod) {
      // nothing to contract: C unchanged
      continue;
    }
    Prod Pn = Pc;
    CsStep nxt = advance(cur, Pn);
    Prod Pl = Pn;
    CsStep ldn = nxt.r < J.nprod ? advance(nxt, Pl) : nxt;

    CsSrcs<Cfg::CH> Sc, Sn, Sl;
    cs_make_srcs<Cfg::CH>(J, Pc.a0, Pc.na, Pc.b0, Pc.nb, Sc);
    CsRaw<Cfg> raw;
    // prologue: slice cur into buffer 0, loads of nxt in flight
    cs_load<Cfg>(J, Sc, cur.st * BK, cs_depth(J, Pc), row0, col0, interior, th, raw);
    cs_pack<Cfg>(Sc, th, raw, As[0], Bs[0]);
    if (nxt.r < J.nprod) {
      cs_make_srcs<Cfg::CH>(J, Pn.a0, Pn.na, Pn.b0, Pn.nb, Sn);
      cs_load<Cfg>(J, Sn, nxt.st * BK, cs_depth(J, Pn), row0, col0, interior, th, raw);
    }
    if (ldn.r < J.nprod) cs_make_srcs<Cfg::CH>(J, Pl.a0, Pl.na, Pl.b0, Pl.nb, Sl);
    __syncthreads();

    f64x4 acc[FM][FN];
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

    int buf = 0;
    while (true) {
      cs_mfma<Cfg, 0, BK / 8>(As[buf], Bs[buf], lane, wm, wn, acc);
      if (nxt.r < J.nprod) {
        cs_pack<Cfg>(Sn, th, raw, As[buf ^ 1], Bs[buf ^ 1]);
        if (ldn.r < J.nprod) cs_load<Cfg>(J, Sl, ldn.st * BK, cs_depth(J, Pl), row0, col0, interior, th, raw);
      }
      cs_mfma<Cfg, BK / 8, BK / 4>(As[buf], Bs[buf], lane, wm, wn, acc);
      if (cur.st + 1 == nslices(Pc)) {
        // product Pc complete: C_j = fma(w_j, M, C_j) for its W terms, in
        // table order (one term per C block), then a fresh accumulator
        for (int t = 0; t < Pc.nc; ++t) {
          const Term tm = cs_term(J, Pc.c0 + t);
          const int jt = tm.bi * J.nx + tm.bj;
          const double wv = tm.coef;
#pragma unroll
          for (int j = 0; j < NCB; ++j) {
            if (j == jt) {
#pragma unroll
              for (int mi = 0; mi < FM; ++mi)
#pragma unroll
                for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                  for (int e = 0; e < 4; ++e) cacc[j][mi][ni][e] = __builtin_fma(wv, acc[mi][ni][e], cacc[j][mi][ni][e]);
            }
          }
        }
#pragma unroll
        for (int mi = 0; mi < FM; ++mi)
#pragma unroll
          for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};
      }
      __syncthreads();
      if (nxt.r >= J.nprod) break;
      // rotate the stream
      cur = nxt;
      Pc = Pn;
      Sc = Sn;
      nxt = ldn;
      Pn = Pl;
      Sn = Sl;
      if (ldn.r < J.nprod) {
        // the source descriptors change only with the product
        const int r0 = ldn.r;
        ldn = advance(ldn, Pl);
        if (ldn.r < J.nprod && ldn.r != r0) cs_make_srcs<Cfg::CH>(J, Pl.a0, Pl.na, Pl.b0, Pl.nb, Sl);
      }
      buf ^= 1;
    }

    // write this position's C tiles back (each once)
#pragma unroll
    for (int j = 0; j < NCB; ++j) {
      if (j < ncb) {
        gdouble* cb = gptr(J.C) + (long long)(j / J.nx) * J.ms + (long long)(j % J.nx) * J.ns * J.ldc;
#pragma unroll
        for (int mi = 0; mi < FM; ++mi)
#pragma unroll
          for (int ni = 0; ni < FN; ++ni)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int row = crow + mi * 16, col = ccol + ni * 16 + 4 * e;
              if (row < J.ms && col < J.ns) cb[row + (long long)col * J.ldc] = cacc[j][mi][ni][e];
            }
      }
    }
  }
}

}  // namespace fmm
