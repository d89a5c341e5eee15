// fmm_f64d.hpp -- single-term FP64 kernel with direct-to-LDS operand loads
// (gfx950).
//
// Serves the FP64 launches whose every product has one A and one B term
// (plain GEMMs, Naive products over the pre-sums or single blocks, k-fringe
// products): the units, C tiles (128 x 128, 8 waves of 64 x 32), MFMA chains,
// C updates and product order of abc_f64_kernel's single-term variant, with
// the operands moved from HBM to LDS by global_load_lds_dwordx4 (no register
// staging, no LDS stores by the waves) into a ring of NST stages BK deep.
//
//   * A stays column-major in LDS (As[k][i], one 1 KiB row per k, read as
//     the MFMA's 16 consecutive rows of one k); B stays k-contiguous
//     (Bs[j][k], 8*BK bytes per column), its 16-byte chunks XOR-swizzled
//     within the column so that the 16 columns of one fragment read hit
//     distinct banks.
//   * A product's operands with coefficient +-1 take the direct path and the
//     sign u*v is applied to the accumulators (exact); other coefficients,
//     rows / columns without 16-byte access and a partial last stage are
//     staged through registers (u*x, zero fill) into the same layout -- the
//     values the register-staged kernel stores, so results are bit for bit
//     those of abc_f64_kernel<32, 1>.
//   * The barrier that publishes stage s+1 sits inside stage s, after its
//     last fragment reads, so the last k-step's MFMAs cover it.
//
// Measured against the register-staged single-term kernel (9600^3): GEMM
// 66.3 -> 67.6-67.9 TFLOP/s, Naive 333 L1 67.1 -> 68.5-69.1, 16384^3 GEMM
// 66.1 -> 67.8, rank-k 14400^2 x 480 GEMM 53.5 -> 55.6; MFMA busy 0.85 ->
// 0.87.  16-deep stages in rings of 3 or 4 ran 3-4% slower than 32-deep
// stages in a ring of 2.  Letting the waves of an edge tile whose sub-tile
// lies outside the leaf block skip their MFMAs (a branch around the stage's
// fragment reads and MFMAs) saved 0.5% on <2,2,2> x <3,2,3> (leaf 1600, half
// tiles at every edge) and cost 3-4% on whole tiles, with or without
// pairing opposite row halves on one SIMD.
#pragma once
#include "fmm_h16d.hpp"

// tile rows per group of the unit order (tile_coords); 2, 8 and 16
// measured within 0.3% of 4 (333 L1 Naive 9600^3, GEMM 9600^3, 222 L2
// Naive 16384^3)
#ifndef FMM_F64D_GROUP
#define FMM_F64D_GROUP 4
#endif

namespace fmm {

template <int BK_, int NST_>
struct F64DCfg {
  static constexpr int BK = BK_, NST = NST_;
  static constexpr int NW = NT / 64;
  static constexpr int KS = BK / 4;             // MFMA k-steps per stage
  static constexpr int A_BYTES = BK * BM * 8;   // BK rows of 1 KiB
  static constexpr int B_ROW = BK * 8;          // bytes per B column
  static constexpr int B_CH = BK / 2;           // 16-byte chunks per B column
  static constexpr int B_CPI = 1024 / B_ROW;    // B columns per wave load
  static constexpr int B_BYTES = BN * B_ROW;
  static constexpr int STAGE = A_BYTES + B_BYTES;
  static constexpr int LDS_BYTES = NST * STAGE;
  static constexpr int A_LD = BK / NW;          // A rows (wave loads) per wave
  static constexpr int B_LD = B_BYTES / 1024 / NW;
  static constexpr int LPT = A_LD + B_LD;
  static_assert(BM == 128 && BN == 128 && NT == 512, "tile");
  static_assert(BK == 16 || BK == 32, "B swizzle written for 8 or 16 chunks per column");
  static_assert(A_LD * NW == BK && B_LD * NW * 1024 == B_BYTES, "whole wave loads");
  static_assert(NST >= 2 && (NST - 1) * LPT <= 63, "vmcnt range");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");
};

// Physical chunk of logical k pair kg of B column j: the 16 columns of one
// fragment read land in 16 distinct chunk slots of 256 bytes.  A rows need
// no swizzle: XOR patterns on the row chunks (rows k and k+1, or k and k+2,
// 128 bytes apart) measured the same bank-conflict count (a quarter of the
// LDS-active cycles, from the B reads of lanes k and k+2, which share a
// chunk half) and no faster.
template <int NCH>
__device__ __forceinline__ int f_swz_b(int j) {
  return j & (NCH - 1);
}

// This lane's direct loads: byte offsets from the stage's A base (A block +
// row0 + kb*lda) and B base (B block + col0*ldb + kb); rows / columns past
// the leaf block load the last valid chunk / column.
template <class C>
struct FGeo {
  unsigned a[C::A_LD];
  unsigned b[C::B_LD];
};

template <class C>
__device__ __forceinline__ void f_geo(const DDims& J, int row0, int col0, int wave, int lane, FGeo<C>& g) {
  const int rlast = (J.ms - row0) / 2 - 1;   // last chunk with valid rows (direct path: ms even)
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) {
    const int k = wave + C::NW * t;
    g.a[t] = (unsigned)(((long long)k * J.lda + 2 * min(lane, rlast)) * 8);
  }
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) {
    const int jj = (wave + C::NW * t) * C::B_CPI + lane / C::B_CH;
    const int kg = (lane % C::B_CH) ^ f_swz_b<C::B_CH>(jj);
    g.b[t] = (unsigned)(((long long)(min(col0 + jj, J.ns - 1) - col0) * J.ldb + 2 * kg) * 8);
  }
}

template <class C>
__device__ __forceinline__ void f_issue(const double* a, const double* b, long long lda, int kb, const FGeo<C>& g,
                                        char* st, int wave) {
  const void* ab = unip(a + (long long)kb * lda);
  const void* bb = unip(b + kb);
  const unsigned l0 = __builtin_amdgcn_readfirstlane(lds_u32(st) + wave * 1024);
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) d_dma16(l0 + t * C::NW * 1024, g.a[t], ab);
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) d_dma16(l0 + C::A_BYTES + t * C::NW * 1024, g.b[t], bb);
}

// Register-staged fill of stage kb (rem valid k; rows / columns past the
// leaf block and k past rem read as zero), operands scaled by u / v, into
// the same layout.  a, b: the A block at row0, the B block at column col0.
template <class C>
__device__ __forceinline__ void f_stage_slow(const DDims& J, const double* a, const double* b, double u, double v,
                                             int row0, int col0, int kb, int rem, char* st, int wave, int lane) {
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) {
    const int k = wave + C::NW * t;
    const int r = row0 + 2 * lane;
    const gcdouble* p = gptr(a + (long long)(kb + k) * J.lda + 2 * lane);
    const bool kok = k < rem;
    const double x0 = kok && r < J.ms ? p[0] : 0.0;
    const double x1 = kok && r + 1 < J.ms ? p[1] : 0.0;
    *reinterpret_cast<f64x2*>(st + (wave + C::NW * t) * 1024 + lane * 16) = f64x2{u * x0, u * x1};
  }
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) {
    const int jj = (wave + C::NW * t) * C::B_CPI + lane / C::B_CH;
    const int kg = (lane % C::B_CH) ^ f_swz_b<C::B_CH>(jj);
    const gcdouble* p = gptr(b + (long long)jj * J.ldb + kb + 2 * kg);
    const bool cok = col0 + jj < J.ns;
    const double x0 = cok && 2 * kg < rem ? p[0] : 0.0;
    const double x1 = cok && 2 * kg + 1 < rem ? p[1] : 0.0;
    *reinterpret_cast<f64x2*>(st + C::A_BYTES + (wave + C::NW * t) * 1024 + lane * 16) = f64x2{v * x0, v * x1};
  }
}

// Fragments of k-step kk: a[mi] = A rows wm*64 + mi*16 + l16 at k = 4kk + lk,
// b[ni] = B columns wn*32 + ni*16 + l16 at the same k.
template <class C>
__device__ __forceinline__ void f_frags(const char* st, int kk, const int (&aofs)[FM], int bcol, int lk,
                                        double (&a)[FM], double (&b)[FN]) {
#pragma unroll
  for (int mi = 0; mi < FM; ++mi) a[mi] = *reinterpret_cast<const double*>(st + kk * 4 * 1024 + aofs[mi]);
  const int k = 4 * kk + lk;
#pragma unroll
  for (int ni = 0; ni < FN; ++ni) {
    const int j = bcol + ni * 16;
    const int off = j * C::B_ROW + (((k >> 1) ^ f_swz_b<C::B_CH>(j)) << 4) + ((k & 1) << 3);
    b[ni] = *reinterpret_cast<const double*>(st + C::A_BYTES + off);
  }
}

// MFMAs of one stage; the fragments of k-step kk+1 are read while those of
// kk issue.  `hook` (the barrier publishing the next stage and the loads it
// frees this buffer for) runs once every fragment of the stage has been
// read, before the last k-step's MFMAs.
template <class C, class Hook>
__device__ __forceinline__ void f_stage(const char* st, const int (&aofs)[FM], int bcol, int lk,
                                        f64x4 (&acc)[FM][FN], Hook&& hook) {
  double a[2][FM], b[2][FN];
  f_frags<C>(st, 0, aofs, bcol, lk, a[0], b[0]);
#pragma unroll
  for (int kk = 0; kk < C::KS; ++kk) {
    const int c = kk & 1;
    if (kk + 1 < C::KS) f_frags<C>(st, kk + 1, aofs, bcol, lk, a[c ^ 1], b[c ^ 1]);
    if (kk == C::KS - 1) hook();
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < FN; ++ni)
        acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[c][ni], a[c][mi], acc[mi][ni], 0, 0, 0);
  }
}

template <class C>
__device__ __forceinline__ void f_contract(const Job& J, const Prod& P, int row0, int col0, int depth, char* lds,
                                           int wave, int lane, int wm, int wn, f64x4 (&acc)[FM][FN]) {
  if (!(P.na > 0 && P.nb > 0 && depth > 0)) return;
  const Term ta = term_at(J, P.a0), tb = term_at(J, P.b0);
  const double* a = unip(static_cast<const double*>(J.A) + (long long)ta.bi * J.ms + (long long)ta.bj * J.ks * J.lda);
  const double* b = unip(static_cast<const double*>(J.B) + (long long)tb.bi * J.ks + (long long)tb.bj * J.ns * J.ldb);
  const double u = unid(ta.coef), v = unid(tb.coef);
  DDims D;
  D.lda = uni64(J.lda);
  D.ldb = uni64(J.ldb);
  D.ms = __builtin_amdgcn_readfirstlane(J.ms);
  D.ns = __builtin_amdgcn_readfirstlane(J.ns);
  // direct path: unit coefficients, 16-byte access (vec_flags), whole row
  // pairs of A, 32-bit lane offsets
  const bool dma = __builtin_fabs(u) == 1.0 && __builtin_fabs(v) == 1.0 &&
                   (__builtin_amdgcn_readfirstlane(J.flags) & 3) == 3 && (D.ms & 1) == 0 && D.lda < (1ll << 21) &&
                   D.ldb < (1ll << 22);
  const double* at = unip(a + row0);
  const double* bt = unip(b + (long long)col0 * D.ldb);
  FGeo<C> g;
  if (dma) f_geo<C>(D, row0, col0, wave, lane, g);
  const int l16 = lane & 15, lk = lane >> 4;
  int aofs[FM];
#pragma unroll
  for (int mi = 0; mi < FM; ++mi) {
    const int i = wm * WTM + mi * 16 + l16;
    aofs[mi] = lk * 1024 + i * 8;
  }
  const int bcol = wn * WTN + l16;
  const int nfull = depth / C::BK;
  const int rem = depth - nfull * C::BK;
  auto stage = [&](int s) { return lds + (s % C::NST) * C::STAGE; };
  auto fill = [&](int s) {
    if (dma) f_issue<C>(at, bt, D.lda, s * C::BK, g, stage(s), wave);
    else f_stage_slow<C>(D, at, bt, u, v, row0, col0, s * C::BK, C::BK, stage(s), wave, lane);
  };
#pragma unroll
  for (int s = 0; s < C::NST; ++s)
    if (s < nfull) fill(s);
  if (nfull > 0) {
    d_wait_stages<C>(min(C::NST - 1, nfull - 1));
    __syncthreads();
  }
  for (int s = 0; s < nfull; ++s) {
    f_stage<C>(stage(s), aofs, bcol, lk, acc, [&]() {
      if (s + 1 < nfull) {
        d_wait_stages<C>(min(C::NST - 2, nfull - 2 - s));
        __syncthreads();
        if (s + C::NST < nfull) fill(s + C::NST);
      }
    });
  }
  if (rem > 0) {
    __syncthreads();
    f_stage_slow<C>(D, at, bt, dma ? 1.0 : u, dma ? 1.0 : v, row0, col0, nfull * C::BK, rem, stage(nfull), wave, lane);
    __syncthreads();
    f_stage<C>(stage(nfull), aofs, bcol, lk, acc, []() {});
  }
  __syncthreads();   // the next unit's first loads overwrite stages read here
  if (dma && u * v < 0.0) {
#pragma unroll
    for (int mi = 0; mi < FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = -acc[mi][ni];
  }
}

template <class C>
__global__ void __launch_bounds__(NT) f64d_kernel(Launch L) {
  __shared__ __align__(16) char lds[C::LDS_BYTES];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % WAVES_M, wn = wave / WAVES_M;
  run_units(L, reinterpret_cast<int*>(lds + (C::NST - 1) * C::STAGE), [&](int j, int q, int p) {
    const Job& J = job_at(L, j);
    int tmi, tni;
    tile_coords<FMM_F64D_GROUP>(J.tiles_m, J.tiles_n, p, &tmi, &tni);   // grouped order (Naive 333 L1 9600^3 +0.5%)
    const int row0 = tmi * BM, col0 = tni * BN;
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
      const int depth = __builtin_amdgcn_readfirstlane(P.fringe ? J.kr : J.ks);
      f64x4 acc[FM][FN];
#pragma unroll
      for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};
      f_contract<C>(J, P, row0, col0, depth, lds, wave, lane, wm, wn, acc);
      if (ordered && q > 0 && qq == q0) wait_turn(ctr, q, L);
      scatter_c(J, P, row0, col0, lane, wm, wn, acc);
    }
    if (ordered) pass_turn(ctr, q + 1, L);
  });
}

#ifndef FMM_F64D_CFG
#define FMM_F64D_CFG 32, 2
#endif
typedef F64DCfg<FMM_F64D_CFG> CfgF64D;

}  // namespace fmm
