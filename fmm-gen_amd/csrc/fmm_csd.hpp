// fmm_csd.hpp -- C-stationary FP64 kernel with direct-to-LDS operand loads
// (gfx950).
//
// The C-stationary scheme of cs_f64_kernel (one workgroup owns one tile
// position of the leaf grid for the whole call, keeps that tile of every
// C block in registers, runs every product over it and applies
// C_j = fma(w, M_r, C_j) from the product's accumulators, in product order,
// then writes each C tile once), with the operand movement of f64d_kernel:
// global_load_lds_dwordx4 into a ring of NST stages BK deep, no register
// staging, the barrier that publishes the next stage inside the current one.
//
//   * The stages of all products of a position form one stream (the leaf-
//     depth products, then the k-fringe products as a second stream), so the
//     ring stays full across product boundaries: a product of depth 160
//     (<3,3,3> at k = 480) is five 32-deep stages, and restarting the ring
//     per product would idle the MFMAs for a load latency each time.
//   * A product's operands with coefficient +-1 take the direct path; the
//     sign u*v folds into its W coefficients (exact).  A last stage of
//     depth rem < BK with rem % 4 == 0 is loaded direct with its k clamped
//     to valid rows and contracts rem / 4 k-steps; other coefficients,
//     operands without 16-byte access and other partial stages are staged
//     through registers (u*x, zero fill) into the same layout.  Every C
//     element therefore sees the MFMA chain and the update sequence of
//     cs_f64_kernel (the k-steps skipped are those whose A and B fragments
//     are all zero there).
//   * Geometry by C-block count as in cs_f64_kernel: <= 4 blocks 128 x 64
//     positions (8 waves of 32 x 32), <= 9 blocks 64 x 64 (8 waves of
//     32 x 16); the position tile of every C block lives in registers.
//
// Measured at 14400^2 x 480 (Naive, the whole call) against the register-
// staged kernel: <2,2,2> 4.94 -> 4.51 ms, <2,3,2> 4.82 -> 4.33, <3,3,3>
// 5.97 -> 4.67, <3,2,3> 6.13 -> 4.85 (rings of 2 stages, 3 and 4 measured
// 2-3% slower; whole stages unmasked); the plain GEMM takes 3.56 ms.  The
// MFMAs run about half the time: with the operand loads removed the kernel
// still takes 90% as long, with the MFMAs removed 45-60%, so the stage's
// fixed costs (barrier, next-stage issue, the position's C tiles) add to
// the MFMA time rather than hide under it.  Variants that did not pay:
// 16-deep stages (+25%), the last k-step's MFMAs issued before the refill,
// and 4 waves with one per SIMD (64 x 32 / 32 x 32 wave tiles, the C tiles
// in AGPRs: +6-11%).
#pragma once
#include "fmm_cs.hpp"
#include "fmm_f64d.hpp"

namespace fmm {

template <int NCB_, int BK_, int NST_, int NW_ = 8>
struct CsdCfg {
  static constexpr int NCB = NCB_, BK = BK_, NST = NST_;
  static constexpr int NW = NW_, NT = NW * 64;
  static constexpr int TM = NCB <= 4 ? 128 : 64, TN = 64;
  // 8 waves (two per SIMD): wave tiles 32 x 32 / 32 x 16; 4 waves (one per
  // SIMD, the register file of the SIMD to each): 64 x 32 / 32 x 32
  static constexpr int WAVES_M = NW == 8 ? TM / 32 : 2, WAVES_N = NW / WAVES_M;
  static constexpr int WTM = TM / WAVES_M, WTN = TN / WAVES_N;
  static constexpr int FM = WTM / 16, FN = WTN / 16;
  static constexpr int KS = BK / 4;             // MFMA k-steps per stage
  static constexpr int A_ROW = TM * 8;          // bytes per k row of A
  static constexpr int A_BYTES = BK * A_ROW;
  static constexpr int A_RPL = 1024 / A_ROW;    // k rows per wave load
  static constexpr int A_CPR = A_ROW / 16;      // 16-byte chunks per k row
  static constexpr int B_ROW = BK * 8;          // bytes per B column
  static constexpr int B_CH = BK / 2;           // 16-byte chunks per B column
  static constexpr int B_CPI = 1024 / B_ROW;    // B columns per wave load
  static constexpr int B_BYTES = TN * B_ROW;
  static constexpr int STAGE = A_BYTES + B_BYTES;
  static constexpr int LDS_BYTES = NST * STAGE;
  static constexpr int A_LD = A_BYTES / 1024 / NW;
  static constexpr int B_LD = B_BYTES / 1024 / NW;
  static constexpr int LPT = A_LD + B_LD;
  static_assert(NCB <= 9 && FM >= 1 && FN >= 1 && WAVES_M * WAVES_N == NW && (NW == 8 || NW == 4), "wave layout");
  static_assert(BK == 16 || BK == 32, "B swizzle written for 8 or 16 chunks per column");
  static_assert(A_LD * NW * 1024 == A_BYTES && B_LD * NW * 1024 == B_BYTES, "whole wave loads");
  static_assert(NST >= 2 && (NST - 1) * LPT <= 63, "vmcnt range");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");
};

// One product's operand blocks for the stage stream, and its W terms.  The
// records of all products are built once per workgroup in LDS (csd_tables):
// the stream reads them with LDS loads, which cost neither a scalar-memory
// round trip per stage nor a wait on the vector-memory counter (which also
// counts the direct loads in flight).
struct CsdSrc {
  const double* a;   // A-side block (row 0, k 0)
  const double* b;   // B-side block (k 0, column 0)
  long long la, lb;
  double u, v;       // operand scales of the register-staged loads (1 when direct)
  double sg;         // sign folded into the W coefficients (u*v when direct, else 1)
  int direct;        // unit coefficients and 16-byte access on both sides
  int c0, nc;        // W terms: entries [c0, c0 + nc) of the term table
  int pad;
};
struct CsdCt {
  double w;
  int jt;            // C block bi*nx + bj
  int pad;
};
constexpr int kCsdMaxProd = 64;
constexpr int kCsdMaxTerms = 384;

// Builds the LDS records (every thread a share; the caller synchronizes).
__device__ __forceinline__ void csd_tables(const CsJob& J, CsdSrc* sd, CsdCt* ct, int tid, int nt) {
  for (int r = tid; r < J.nprod; r += nt) {
    const Prod P = J.prods[r];
    const Term ta = J.terms[P.a0], tb = J.terms[P.b0];
    CsdSrc s;
    if (ta.bi < 0) {
      s.a = J.S + (long long)ta.bj * J.ks * J.lds;
      s.la = J.lds;
    } else {
      s.a = J.A + (long long)ta.bi * J.ms + (long long)ta.bj * J.ks * J.lda;
      s.la = J.lda;
    }
    if (tb.bi < 0) {
      s.b = J.T + (long long)tb.bj * J.ns * J.ldt;
      s.lb = J.ldt;
    } else {
      s.b = J.B + (long long)tb.bi * J.ks + (long long)tb.bj * J.ns * J.ldb;
      s.lb = J.ldb;
    }
    // workspace pre-sums (bi < 0) are 16-byte aligned with even leading
    // dimensions whenever ms is even
    const bool aok = ta.bi < 0 || (J.vec & 1) != 0;
    const bool bok = tb.bi < 0 || (J.vec & 2) != 0;
    const bool direct = __builtin_fabs(ta.coef) == 1.0 && __builtin_fabs(tb.coef) == 1.0 && aok && bok &&
                        (J.ms & 1) == 0 && s.la < (1ll << 21) && s.lb < (1ll << 22);
    s.direct = direct;
    s.u = direct ? 1.0 : ta.coef;
    s.v = direct ? 1.0 : tb.coef;
    s.sg = direct ? ta.coef * tb.coef : 1.0;
    s.c0 = P.c0;
    s.nc = P.nc;
    s.pad = 0;
    sd[r] = s;
  }
  for (int t = tid; t < J.nterms; t += nt) {
    const Term tm = J.terms[t];
    ct[t] = CsdCt{tm.coef, tm.bi * J.nx + tm.bj, 0};
  }
}

// Product r's record, uniform (into scalar registers).
__device__ __forceinline__ CsdSrc csd_src(const CsdSrc* sd, int r) {
  const CsdSrc& d = sd[r];
  CsdSrc s;
  s.a = unip(d.a);
  s.b = unip(d.b);
  s.la = uni64(d.la);
  s.lb = uni64(d.lb);
  s.u = unid(d.u);
  s.v = unid(d.v);
  s.sg = unid(d.sg);
  s.direct = __builtin_amdgcn_readfirstlane(d.direct);
  s.c0 = __builtin_amdgcn_readfirstlane(d.c0);
  s.nc = __builtin_amdgcn_readfirstlane(d.nc);
  s.pad = 0;
  return s;
}

// Direct loads of stage kb (rem valid k, rem % 4 == 0) of the product's
// position tile into stage buffer st.  Rows past the leaf block load the
// last valid row pair, columns the last valid column, k past rem the last
// valid k (never contracted).
template <class C>
__device__ __forceinline__ void csd_issue(const CsJob& J, const CsdSrc& s, int row0, int col0, int kb, int rem,
                                          char* st, int wave, int lane) {
  const int ms = __builtin_amdgcn_readfirstlane(J.ms), ns = __builtin_amdgcn_readfirstlane(J.ns);
  const void* ab = unip(s.a + row0 + (long long)kb * s.la);
  const void* bb = unip(s.b + (long long)col0 * s.lb + kb);
  const unsigned l0 = __builtin_amdgcn_readfirstlane(lds_u32(st) + wave * 1024);
  const int rlast = (ms - row0) / 2 - 1;
  const int ac = min(lane % C::A_CPR, rlast);
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) {
    const int k = min((wave + C::NW * t) * C::A_RPL + lane / C::A_CPR, rem - 1);
    d_dma16(l0 + t * C::NW * 1024, (unsigned)(((long long)k * s.la + 2 * ac) * 8), ab);
  }
  const int kgl = rem / 2 - 1;
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) {
    const int jj = (wave + C::NW * t) * C::B_CPI + lane / C::B_CH;
    const int kg = min((lane % C::B_CH) ^ f_swz_b<C::B_CH>(jj), kgl);
    d_dma16(l0 + C::A_BYTES + t * C::NW * 1024,
            (unsigned)(((long long)(min(col0 + jj, ns - 1) - col0) * s.lb + 2 * kg) * 8), bb);
  }
}

// Register-staged fill of stage kb (rem valid k; rows / columns past the
// leaf block and k past rem read as zero), operands scaled by u / v.
template <class C>
__device__ __forceinline__ void csd_stage_slow(const CsJob& J, const CsdSrc& s, double u, double v, int row0,
                                               int col0, int kb, int rem, char* st, int wave, int lane) {
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) {
    const int k = (wave + C::NW * t) * C::A_RPL + lane / C::A_CPR;
    const int r = row0 + 2 * (lane % C::A_CPR);
    const gcdouble* p = gptr(s.a + (long long)(kb + k) * s.la + r);
    const bool kok = k < rem;
    const double x0 = kok && r < J.ms ? p[0] : 0.0;
    const double x1 = kok && r + 1 < J.ms ? p[1] : 0.0;
    *reinterpret_cast<f64x2*>(st + (wave + C::NW * t) * 1024 + lane * 16) = f64x2{u * x0, u * x1};
  }
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) {
    const int jj = (wave + C::NW * t) * C::B_CPI + lane / C::B_CH;
    const int kg = (lane % C::B_CH) ^ f_swz_b<C::B_CH>(jj);
    const gcdouble* p = gptr(s.b + (long long)(col0 + jj) * s.lb + kb + 2 * kg);
    const bool cok = col0 + jj < J.ns;
    const double x0 = cok && 2 * kg < rem ? p[0] : 0.0;
    const double x1 = cok && 2 * kg + 1 < rem ? p[1] : 0.0;
    *reinterpret_cast<f64x2*>(st + C::A_BYTES + (wave + C::NW * t) * 1024 + lane * 16) = f64x2{v * x0, v * x1};
  }
}

// Fragments of k-step kk: a[mi] = A rows wm*32 + mi*16 + l16, b[ni] = B
// columns wn*WTN + ni*16 + l16, at k = 4kk + lk.
template <class C>
__device__ __forceinline__ void csd_frags(const char* st, int kk, const int (&aofs)[C::FM], int bcol, int lk,
                                          double (&a)[C::FM], double (&b)[C::FN]) {
#pragma unroll
  for (int mi = 0; mi < C::FM; ++mi) a[mi] = *reinterpret_cast<const double*>(st + kk * 4 * C::A_ROW + aofs[mi]);
  const int k = 4 * kk + lk;
#pragma unroll
  for (int ni = 0; ni < C::FN; ++ni) {
    const int j = bcol + ni * 16;
    const int off = j * C::B_ROW + (((k >> 1) ^ f_swz_b<C::B_CH>(j)) << 4) + ((k & 1) << 3);
    b[ni] = *reinterpret_cast<const double*>(st + C::A_BYTES + off);
  }
}

// The MFMAs of one k-step (operands as in the ordered kernel, X = B^T,
// Y = A^T); `ok` false zeroes this lane's fragments first (k past the valid
// depth of a product's last stage, whose rows were loaded clamped).
template <class C, bool MASK = true>
__device__ __forceinline__ void csd_mfma(const double (&a)[C::FM], const double (&b)[C::FN], bool ok,
                                         f64x4 (&acc)[C::FM][C::FN]) {
  double x[C::FM], y[C::FN];
#pragma unroll
  for (int mi = 0; mi < C::FM; ++mi) x[mi] = !MASK || ok ? a[mi] : 0.0;
#pragma unroll
  for (int ni = 0; ni < C::FN; ++ni) y[ni] = !MASK || ok ? b[ni] : 0.0;
#pragma unroll
  for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
    for (int ni = 0; ni < C::FN; ++ni)
      acc[mi][ni] = __builtin_amdgcn_mfma_f64_16x16x4f64(y[ni], x[mi], acc[mi][ni], 0, 0, 0);
}

// MFMAs of k-steps 0 .. KS-2 of one stage, the fragments of k-step kk+1
// read before those of kk issue (a scheduling barrier keeps the reads
// ahead); the last k-step's fragments are returned in (al, bl), so that the
// caller can publish the next stage once every fragment of this one has
// been read, and then issue the last k-step's MFMAs.  kv: valid k of the
// stage (BK but in a product's partial last stage, whose further k-steps
// contract zeros, as the register-staged kernels' zero fill does).
template <class C, bool MASK>
__device__ __forceinline__ void csd_stage_head(const char* st, int kv, const int (&aofs)[C::FM], int bcol, int lk,
                                               f64x4 (&acc)[C::FM][C::FN], double (&al)[C::FM],
                                               double (&bl)[C::FN]) {
  double a[2][C::FM], b[2][C::FN];
  csd_frags<C>(st, 0, aofs, bcol, lk, a[0], b[0]);
#pragma unroll
  for (int kk = 0; kk < C::KS - 1; ++kk) {
    const int c = kk & 1;
    if (kk + 1 < C::KS - 1) csd_frags<C>(st, kk + 1, aofs, bcol, lk, a[c ^ 1], b[c ^ 1]);
    else csd_frags<C>(st, kk + 1, aofs, bcol, lk, al, bl);
    __builtin_amdgcn_sched_barrier(0);
    csd_mfma<C, MASK>(a[c], b[c], 4 * kk + lk < kv, acc);
  }
}

template <class C>
__global__ void __launch_bounds__(C::NT) __attribute__((amdgpu_waves_per_eu(C::NW == 4 ? 1 : 2, 2)))
csd_f64_kernel(CsJob J, int nq) {
  constexpr int BK = C::BK, FM = C::FM, FN = C::FN, NCB = C::NCB, NST = C::NST;
  __shared__ __align__(16) char lds[C::LDS_BYTES];
  __shared__ CsdSrc s_src[kCsdMaxProd];
  __shared__ CsdCt s_ct[kCsdMaxTerms];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % C::WAVES_M, wn = wave / C::WAVES_M;
  const int l16 = lane & 15, lk = lane >> 4;
  int aofs[FM];
#pragma unroll
  for (int mi = 0; mi < FM; ++mi) aofs[mi] = lk * C::A_ROW + (wm * C::WTM + mi * 16 + l16) * 8;
  const int bcol = wn * C::WTN + l16;
  csd_tables(J, s_src, s_ct, tid, C::NT);
  __syncthreads();

  // Positions: nq contiguous chunks (one per XCD when nq = 8), each served
  // by the workgroups x, x + nq, ... in round robin.
  const int units = J.tiles_m * J.tiles_n;
  const int x = (int)(blockIdx.x % (unsigned)nq);
  const int w = (int)(blockIdx.x / (unsigned)nq);
  const int nw = (int)(gridDim.x / (unsigned)nq);
  const int lo = (int)(((long long)x * units) / nq), hi = (int)(((long long)(x + 1) * units) / nq);
  const int ncb = J.mx * J.nx;
  auto stage = [&](int s) { return lds + (s % NST) * C::STAGE; };

  for (int u = lo + w; u < hi; u += nw) {
    int tmi, tni;
    tile_coords<4>(J.tiles_m, J.tiles_n, u, &tmi, &tni);
    const int row0 = tmi * C::TM, col0 = tni * C::TN;
    const int crow = row0 + wm * C::WTM + l16;
    const int ccol = col0 + wn * C::WTN + lk;

    // The two streams: the leaf-depth products [0, nmain), then the k-fringe
    // products [nmain, nprod) of depth kr.
    int p0 = 0, p1 = J.nmain, depth = J.ks;
    auto nps_of = [&]() { return (depth + BK - 1) / BK; };
    int last_slow = -1;   // stages up to the last register-staged one are complete
    int fr = 0, fkb = 0;  // product and depth of the next stage to fill (fills go in order)
    auto fill = [&](int s) {
      const int r = p0 + fr;
      const int kb = fkb;
      const int rem = min(BK, depth - kb);
      fkb += BK;
      if (fkb >= depth) {
        fkb = 0;
        ++fr;
      }
      const CsdSrc src = csd_src(s_src, r);
      if (src.direct && (rem & 3) == 0) {
        csd_issue<C>(J, src, row0, col0, kb, rem, stage(s), wave, lane);
      } else {
        csd_stage_slow<C>(J, src, src.u, src.v, row0, col0, kb, rem, stage(s), wave, lane);
        last_slow = s;
      }
    };
    // wait for stage s with stages up to h issued
    auto wait = [&](int s, int h) {
      if (last_slow < s) d_wait_stages<C>(h - s);
    };
    auto prologue = [&]() {
      last_slow = -1;
      fr = 0;
      fkb = 0;
      const int n = min(NST, (p1 - p0) * nps_of());
#pragma unroll 1
      for (int s = 0; s < n; ++s) fill(s);
      return n;
    };
    const bool main_on = p1 > p0 && depth > 0;
    int pre = main_on ? prologue() : 0;

    // This position's tile of every C block lives in registers.  It is read
    // after the first stream's first stages are issued, and one wait covers
    // both, so the C loads overlap the operand loads.
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
    // s_waitcnt vmcnt(0), visible to the compiler's own wait tracking: the
    // C tiles and the stages issued before them are in place
    __builtin_amdgcn_s_waitcnt(0x0F70);

    for (int phase = 0; phase < 2; ++phase) {
      if (phase == 1) {
        p0 = J.nmain;
        p1 = J.nprod;
        depth = J.kr;
        if (!(p1 > p0 && depth > 0)) break;
        pre = prologue();
        wait(0, pre - 1);
      } else if (!main_on) {
        continue;
      }
      __syncthreads();
      const int nps = nps_of();
      const int ns_tot = (p1 - p0) * nps;

      f64x4 acc[FM][FN];
#pragma unroll
      for (int mi = 0; mi < FM; ++mi)
#pragma unroll
        for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};

      int cr = p0, ckb = 0;   // product and depth of stage s
      for (int s = 0; s < ns_tot; ++s) {
        const int rem = min(BK, depth - ckb);
        double al[FM], bl[FN];
        // a product's partial last stage masks its fragments; whole stages
        // run unmasked (the masks cost the 32 x 16 wave tiles 10%)
        const bool part = rem < BK;
        if (part) csd_stage_head<C, true>(stage(s), rem, aofs, bcol, lk, acc, al, bl);
        else csd_stage_head<C, false>(stage(s), rem, aofs, bcol, lk, acc, al, bl);
        // every fragment of stage s is read: publish stage s+1, refill s's buffer
        if (s + 1 < ns_tot) {
          wait(s + 1, min(s + NST - 1, ns_tot - 1));
          __syncthreads();
          if (s + NST < ns_tot) fill(s + NST);
        }
        if (part) csd_mfma<C, true>(al, bl, 4 * (C::KS - 1) + lk < rem, acc);
        else csd_mfma<C, false>(al, bl, true, acc);
        ckb += BK;
        if (ckb >= depth) {
          // product complete: C_j = fma(w_j, M, C_j) for its W terms in table
          // order; direct products fold their operand sign into w
          const CsdSrc src = csd_src(s_src, cr);
          ckb = 0;
          ++cr;
          for (int t = 0; t < src.nc; ++t) {
            const CsdCt& ce = s_ct[src.c0 + t];
            const int jt = __builtin_amdgcn_readfirstlane(ce.jt);
            const double wv = src.sg * unid(ce.w);
#pragma unroll
            for (int j = 0; j < NCB; ++j) {
              if (j == jt) {
#pragma unroll
                for (int mi = 0; mi < FM; ++mi)
#pragma unroll
                  for (int ni = 0; ni < FN; ++ni)
#pragma unroll
                    for (int e = 0; e < 4; ++e)
                      cacc[j][mi][ni][e] = __builtin_fma(wv, acc[mi][ni][e], cacc[j][mi][ni][e]);
              }
            }
          }
#pragma unroll
          for (int mi = 0; mi < FM; ++mi)
#pragma unroll
            for (int ni = 0; ni < FN; ++ni) acc[mi][ni] = f64x4{0.0, 0.0, 0.0, 0.0};
        }
      }
      __syncthreads();   // the next stream's first loads overwrite stages read here
    }

    // fresh C addresses for the stores (opaque to the compiler, so that the
    // load addresses are not kept live across the products to be reused)
    long long ldc_st = J.ldc;
    double* c_st = J.C;
    __asm__ volatile("" : "+s"(ldc_st), "+s"(c_st));
#pragma unroll
    for (int j = 0; j < NCB; ++j) {
      if (j < ncb) {
        gdouble* cb = gptr(c_st) + (long long)(j / J.nx) * J.ms + (long long)(j % J.nx) * J.ns * ldc_st;
#pragma unroll
        for (int mi = 0; mi < FM; ++mi)
#pragma unroll
          for (int ni = 0; ni < FN; ++ni)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
              const int row = crow + mi * 16, col = ccol + ni * 16 + 4 * e;
              if (row < J.ms && col < J.ns) cb[row + (long long)col * ldc_st] = cacc[j][mi][ni][e];
            }
      }
    }
  }
}

}  // namespace fmm
