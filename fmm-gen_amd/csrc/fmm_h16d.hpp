// fmm_h16d.hpp -- single-term BF16 / FP16 kernel with direct-to-LDS operand
// loads (gfx950).
//
// Runs the launches whose every product has one A and one B term with
// coefficient +-1 (plain GEMMs, Naive products over the pre-sums, k-fringe
// products): the same jobs, units, product order and C hand-over as
// abc_lowp_kernel, with a different contraction.
//
//   * Operands go from HBM to LDS without passing through registers
//     (global_load_lds_dwordx4: each lane of a wave moves 16 bytes to
//     consecutive 16-byte slots of a 1 KiB LDS run), into a ring of NST
//     stages BK deep; a stage is waited for with a precise vmcnt and one
//     barrier, so NST - 1 stages of loads stay in flight.
//   * A stays in its column-major orientation in LDS (As[k][i], i
//     contiguous); its MFMA operand, 8 consecutive k of one row per lane, is
//     read with two ds_read_b64_tr_b16 (transposing reads: within a group of
//     16 lanes, lane L receives element L%4 of the 4 values addressed by lane
//     4r + L/4, r = 0..3 -- layout measured on gfx950).  B (k contiguous) is
//     stored Bs[j][k] and read with one 16-byte read per operand.
//   * 16-byte chunks are XOR-swizzled within their row so the transposing A
//     reads (4 k rows x 32 bytes per 16 lanes) and the B reads (16 columns,
//     one k group) fall on distinct banks; the DMA writes are contiguous, the
//     swizzle is applied by choosing which global chunk each lane loads.
//   * A coefficient -1 cannot be applied on the way to LDS; it is applied to
//     the accumulators instead (negation is exact, the products are the same
//     up to sign, so the result equals the register-staged kernel's bit for
//     bit).
//   * Rows / columns past the leaf block in edge tiles load the last valid
//     chunk / column (their results are never stored); a partial last k
//     stage, and jobs without 16-byte operand access, are staged through
//     registers with zero fill into the same LDS layout.
//
// Work per unit and the C update are those of abc_lowp_kernel (l_scatter).
#pragma once
#include "fmm_lowp.hpp"

namespace fmm {

typedef __bf16 bf16x4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) bf16x4v lds_bf16x4v;
typedef __attribute__((address_space(3))) void lds_void;
typedef __attribute__((address_space(1))) const void gcvoid;

template <int IN_, int BM_, int BN_, int WM_, int WN_, int BK_, int NST_, bool SPLIT_ = false>
struct DCfg {
  static constexpr int IN = IN_;
  static constexpr int OP = IN_ == kInBF16 ? kOpBF16 : kOpF16;
  // split-BF16: A and B arrive as hi and lo BF16 planes (the lo plane at
  // the job's alo / blo); a product is al*bh + ah*bl + ah*bh per 32-deep
  // k-step, in the order of the register-staged split kernel
  static constexpr bool SPLIT = SPLIT_;
  static constexpr int NPL = SPLIT ? 2 : 1;
  static constexpr int BM = BM_, BN = BN_, BK = BK_;
  static constexpr int WM = WM_, WN = WN_, NST = NST_;
  static constexpr int NT = 64 * WM * WN;
  static constexpr int NW = NT / 64;
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int FM = WTM / 16, FN = WTN / 16;
  static constexpr int KS = BK / 32;            // MFMA k-steps per stage
  static constexpr int A_ROW = BM * 2;          // bytes of one k row of the A stage
  static constexpr int A_CH = BM / 8;           // 16-byte chunks per A row
  static constexpr int A_RPI = 1024 / A_ROW;    // A rows per wave load (1 KiB)
  static constexpr int B_ROW = BK * 2;          // bytes of one column of the B stage
  static constexpr int B_CH = BK / 8;           // chunks per B column
  static constexpr int B_CPI = 1024 / B_ROW;    // B columns per wave load
  static constexpr int A_BYTES = BK * A_ROW;
  static constexpr int B_BYTES = BN * B_ROW;
  static constexpr int PLANE = A_BYTES + B_BYTES;    // one operand plane of a stage: A, then B
  static constexpr int STAGE = NPL * PLANE;
  static constexpr int LDS_BYTES = NST * STAGE;
  static constexpr int A_LD = A_BYTES / 1024 / NW;   // wave loads per stage, plane and wave
  static constexpr int B_LD = B_BYTES / 1024 / NW;
  static constexpr int LPT = NPL * (A_LD + B_LD);    // loads per thread per stage
  static_assert(IN == kInBF16 || IN == kInF16, "16-bit inputs");
  static_assert(!SPLIT || IN == kInBF16, "split planes are BF16");
  static_assert(BK == 32 || BK == 64, "B swizzle written for 4 or 8 chunks per column");
  static_assert(A_CH >= 16 && A_ROW <= 1024 && 1024 % A_ROW == 0, "A stage rows");
  static_assert(A_LD * NW * 1024 == A_BYTES && B_LD * NW * 1024 == B_BYTES, "whole wave loads");
  static_assert(WTM % 16 == 0 && WTN % 16 == 0, "wave tile");
  static_assert(NST >= 2 && (NST - 1) * LPT <= 63, "vmcnt range");
  // A blocks whose MFMAs run after the barrier that ends a stage (their A
  // operands are read before it), hiding the barrier and the next stage's
  // first operand reads behind them
#ifndef FMM_H16D_TAIL
#define FMM_H16D_TAIL 2
#endif
  static constexpr int TAILM = FMM_H16D_TAIL < FM ? FMM_H16D_TAIL : FM - 1;
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");
};

// Physical 16-byte chunk of logical chunk c of A row k / of logical k group
// kg of B column j.
// A: rows k..k+3 of one 16-lane transposing read land 32 bytes apart, rows 8
// apart in the other half of 256 bytes.  B (NCH chunks per column): 8 (16)
// consecutive columns at one k group hit distinct 16-byte bank groups of
// 128 (256) bytes.  Hardware counters read zero bank conflicts for both
// layouts (SQ_LDS_BANK_CONFLICT); the 8-chunk alternative (j ^ j >> 3) & 7
// read 25% conflicted cycles and ran 2-4% slower.
__device__ __forceinline__ int d_swz_a(int k) { return ((k & 3) << 1) ^ (((k >> 3) & 1) << 3); }
template <int NCH>
__device__ __forceinline__ int d_swz_b(int j) {
  if constexpr (NCH == 4) return ((j >> 1) ^ (j >> 2)) & 3;
  else return j & 7;
}

template <int N>
__device__ __forceinline__ void d_wait_vm() {
  __asm__ volatile("s_waitcnt vmcnt(%0)" ::"n"(N) : "memory");
}

// Wait until at most n stages of this thread's loads are outstanding.
template <class C, int N = C::NST - 1>
__device__ __forceinline__ void d_wait_stages(int n) {
  if constexpr (N > 0) {
    if (n >= N) {
      d_wait_vm<N * C::LPT>();
      return;
    }
    d_wait_stages<C, N - 1>(n);
  } else {
    d_wait_vm<0>();
  }
}

// Which chunk of a stage a lane moves, for wave load t of A (B): the A row
// k within the stage and the logical chunk c (8 rows) of that row; the B
// column jj within the tile and the logical k group kg.
template <class C>
__device__ __forceinline__ void d_a_chunk(int wave, int lane, int t, int* k, int* c) {
  *k = (wave + C::NW * t) * C::A_RPI + lane / C::A_CH;
  *c = (lane % C::A_CH) ^ d_swz_a(*k);
}
template <class C>
__device__ __forceinline__ void d_b_chunk(int wave, int lane, int t, int* jj, int* kg) {
  *jj = (wave + C::NW * t) * C::B_CPI + lane / C::B_CH;
  *kg = (lane % C::B_CH) ^ d_swz_b<C::B_CH>(*jj);
}

// Job fields the contraction uses, held in (scalar) registers.
struct DDims {
  long long lda, ldb;
  int ms, ns;
};

// Byte offsets of this lane's direct loads from the stage's A base
// (A block + row0 + kb * lda) and B base (B block + col0 * ldb + kb).  Rows /
// columns past the leaf block load the last valid chunk / column.
template <class C>
struct DGeo {
  unsigned a[C::A_LD];
  unsigned b[C::B_LD];
};

template <class C>
__device__ __forceinline__ void d_geo(const DDims& J, int row0, int col0, int wave, int lane, DGeo<C>& g) {
  const int rlast = (J.ms - row0) / 8 - 1;   // last chunk with valid rows (direct path: ms % 8 == 0)
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) {
    int k, c;
    d_a_chunk<C>(wave, lane, t, &k, &c);
    g.a[t] = (unsigned)(((long long)k * J.lda + 8 * min(c, rlast)) * 2);
  }
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) {
    int jj, kg;
    d_b_chunk<C>(wave, lane, t, &jj, &kg);
    g.b[t] = (unsigned)(((long long)(min(col0 + jj, J.ns - 1) - col0) * J.ldb + 8 * kg) * 2);
  }
}

// 16 bytes from global base + voff to LDS at lds_addr + 16 * lane.  Written
// in assembly so that the compiler's wait-counter bookkeeping does not treat
// every LDS read as dependent on every load in flight (with the builtin it
// waited for all of them before each stage's loads); the waits for these
// loads are explicit (d_wait_stages).  No register is written.
__device__ __forceinline__ void d_dma16(unsigned lds_addr, unsigned voff, const void* base) {
  __asm__ volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, %2" ::"s"(lds_addr), "v"(voff), "s"(base)
                   : "memory", "m0");
}

__device__ __forceinline__ unsigned lds_u32(const char* p) {
  return (unsigned)(size_t)(__attribute__((address_space(3))) const char*)p;
}

// Direct loads of stage kb into LDS stage buffer st.
template <class C>
__device__ __forceinline__ void d_issue(const typename InT<C::IN>::T* a, const typename InT<C::IN>::T* b,
                                        long long lda, long long alo, long long blo, int kb, const DGeo<C>& g,
                                        char* st, int wave) {
  const unsigned l0 = __builtin_amdgcn_readfirstlane(lds_u32(st) + wave * 1024);
#pragma unroll
  for (int pl = 0; pl < C::NPL; ++pl) {
    const void* ab = unip(a + (long long)kb * lda + pl * alo);
    const void* bb = unip(b + kb + pl * blo);
    const unsigned lp = l0 + pl * C::PLANE;
#pragma unroll
    for (int t = 0; t < C::A_LD; ++t) d_dma16(lp + t * C::NW * 1024, g.a[t], ab);
#pragma unroll
    for (int t = 0; t < C::B_LD; ++t) d_dma16(lp + C::A_BYTES + t * C::NW * 1024, g.b[t], bb);
  }
}

// Register-staged loads of stage kb (rem valid k, rows / columns past the
// leaf block and k past rem read as zero) into the same LDS layout.
// a, b: the A block at row0, the B block at column col0.
template <class C>
__device__ __forceinline__ void d_stage_slow(const DDims& J, const typename InT<C::IN>::T* a0,
                                             const typename InT<C::IN>::T* b0, long long alo, long long blo,
                                             int row0, int col0, int kb, int rem, char* st0, int wave, int lane) {
  typedef typename InT<C::IN>::T T;
#pragma unroll
  for (int pl = 0; pl < C::NPL; ++pl) {
  const T* a = a0 + pl * alo;
  const T* b = b0 + pl * blo;
  char* st = st0 + pl * C::PLANE;
#pragma unroll
  for (int t = 0; t < C::A_LD; ++t) {
    int k, c;
    d_a_chunk<C>(wave, lane, t, &k, &c);
    unsigned int w[4] = {0u, 0u, 0u, 0u};
    const bool kok = k < rem;
    const T* p = a + (long long)(kb + k) * J.lda + 8 * c;
#pragma unroll
    for (int e = 0; e < 8; ++e) put_elem<C::IN>(w, e, kok && row0 + 8 * c + e < J.ms, p + e);
    *reinterpret_cast<u32x4*>(st + (wave + C::NW * t) * 1024 + lane * 16) = u32x4{w[0], w[1], w[2], w[3]};
  }
#pragma unroll
  for (int t = 0; t < C::B_LD; ++t) {
    int jj, kg;
    d_b_chunk<C>(wave, lane, t, &jj, &kg);
    unsigned int w[4] = {0u, 0u, 0u, 0u};
    const bool cok = col0 + jj < J.ns;
    const T* p = b + (long long)jj * J.ldb + kb + 8 * kg;
#pragma unroll
    for (int e = 0; e < 8; ++e) put_elem<C::IN>(w, e, cok && 8 * kg + e < rem, p + e);
    *reinterpret_cast<u32x4*>(st + C::A_BYTES + (wave + C::NW * t) * 1024 + lane * 16) = u32x4{w[0], w[1], w[2], w[3]};
  }
  }
}

// A operand of MFMA block mi: two transposing reads (k 0..3 and 4..7 of
// the lane's k group).
template <class C>
__device__ __forceinline__ u32x4 d_afrag(const char* pa) {
  const bf16x4v lo = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4v*)pa);
  const bf16x4v hi = __builtin_amdgcn_ds_read_tr16_b64_v4bf16((lds_bf16x4v*)(pa + 4 * C::A_ROW));
  const u32x2 l2 = __builtin_bit_cast(u32x2, lo);
  const u32x2 h2 = __builtin_bit_cast(u32x2, hi);
  return u32x4{l2.x, l2.y, h2.x, h2.y};
}

template <class C>
__device__ __forceinline__ void d_mfma1(const u32x4& af, const u32x4& bf, f32x4& acc) {
  if constexpr (C::OP == kOpBF16)
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, af), __builtin_bit_cast(bf16x8, bf), acc,
                                                  0, 0, 0);
  else
    acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, af), __builtin_bit_cast(f16x8, bf), acc, 0,
                                                 0, 0);
}

// B operands of one k-step (hi and lo planes when split).
template <class C>
struct DFrag {
  u32x4 b[C::NPL][C::FN];
};

template <class C>
__device__ __forceinline__ void d_bfrags(const char* st, const int (&bofs)[C::KS][C::FN], int kk, DFrag<C>& f) {
#pragma unroll
  for (int pl = 0; pl < C::NPL; ++pl)
#pragma unroll
    for (int ni = 0; ni < C::FN; ++ni)
      f.b[pl][ni] = *reinterpret_cast<const u32x4*>(st + pl * C::PLANE + C::A_BYTES + bofs[kk][ni]);
}

template <class C>
__device__ __forceinline__ void d_afrags(const char* st, const int (&aofs)[C::FM], int kk, int mi,
                                         u32x4 (&a)[C::NPL]) {
#pragma unroll
  for (int pl = 0; pl < C::NPL; ++pl) a[pl] = d_afrag<C>(st + pl * C::PLANE + kk * 32 * C::A_ROW + aofs[mi]);
}

// The MFMAs of one row of blocks (split: al*bh, ah*bl, ah*bh, the order of
// the register-staged split kernel).
template <class C>
__device__ __forceinline__ void d_mma(const u32x4 (&a)[C::NPL], const DFrag<C>& f, f32x4 (&acc)[C::FN]) {
#pragma unroll
  for (int ni = 0; ni < C::FN; ++ni) {
    if constexpr (C::SPLIT) {
      d_mfma1<C>(a[1], f.b[0][ni], acc[ni]);
      d_mfma1<C>(a[0], f.b[1][ni], acc[ni]);
    }
    d_mfma1<C>(a[0], f.b[0][ni], acc[ni]);
  }
}

// MFMAs of one stage (KS k-steps of 32); the A operands of block mi+1 are
// read while the MFMAs of block mi issue.  `hook` runs after every LDS read
// of the stage has been issued and before the MFMAs of the last TAILM blocks
// of the last k-step.
template <class C, class Hook>
__device__ __forceinline__ void d_stage(const char* st, const int (&aofs)[C::FM], const int (&bofs)[C::KS][C::FN],
                                        f32x4 (&acc)[C::FM][C::FN], Hook&& hook) {
#ifndef FMM_H16D_PRIO
#define FMM_H16D_PRIO 1
#endif
  if (FMM_H16D_PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int kk = 0; kk < C::KS; ++kk) {
    constexpr int T0 = C::FM - C::TAILM;
    const bool last = kk == C::KS - 1;
    DFrag<C> f;
    d_bfrags<C>(st, bofs, kk, f);
    u32x4 a[2][C::NPL];
    d_afrags<C>(st, aofs, kk, 0, a[0]);
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi) {
      if (last && mi == T0) break;
      if (mi + 1 < C::FM && !(last && mi + 1 == T0)) d_afrags<C>(st, aofs, kk, mi + 1, a[(mi + 1) & 1]);
      d_mma<C>(a[mi & 1], f, acc[mi]);
    }
    if (last) {
      u32x4 t[C::TAILM][C::NPL];
#pragma unroll
      for (int i = 0; i < C::TAILM; ++i) d_afrags<C>(st, aofs, kk, T0 + i, t[i]);
      if (FMM_H16D_PRIO) __builtin_amdgcn_s_setprio(0);
      hook();
      if (FMM_H16D_PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int i = 0; i < C::TAILM; ++i) d_mma<C>(t[i], f, acc[T0 + i]);
    }
  }
  if (FMM_H16D_PRIO) __builtin_amdgcn_s_setprio(0);
}

// Contraction of one single-term product over `depth`.
template <class C>
__device__ __forceinline__ void d_contract(const Job& J, const Prod& P, int row0, int col0, int depth, char* lds,
                                           int wave, int lane, int wm, int wn, f32x4 (&acc)[C::FM][C::FN]) {
  typedef typename InT<C::IN>::T T;
  if (!(P.na > 0 && P.nb > 0 && depth > 0)) return;
  const Term ta = term_at(J, P.a0), tb = term_at(J, P.b0);
  const T* a = unip(static_cast<const T*>(J.A) + (long long)ta.bi * J.ms + (long long)ta.bj * J.ks * J.lda);
  const T* b = unip(static_cast<const T*>(J.B) + (long long)tb.bi * J.ks + (long long)tb.bj * J.ns * J.ldb);
  const double sgn = unid(ta.coef * tb.coef);
  DDims D;
  D.lda = uni64(J.lda);
  D.ldb = uni64(J.ldb);
  D.ms = __builtin_amdgcn_readfirstlane(J.ms);
  D.ns = __builtin_amdgcn_readfirstlane(J.ns);
  // direct path: 16-byte access to A and B (vec_flags), whole 8-row chunks of
  // A, and 32-bit lane offsets
  const bool dma = (__builtin_amdgcn_readfirstlane(J.flags) & 3) == 3 && (D.ms & 7) == 0 && D.lda < (1ll << 25) &&
                   D.ldb < (1ll << 23);
  const long long alo = C::SPLIT ? uni64(J.alo) : 0, blo = C::SPLIT ? uni64(J.blo) : 0;
  const T* at = unip(a + row0);                     // A block at the tile's first row
  const T* bt = unip(b + (long long)col0 * D.ldb);  // B block at the tile's first column
  DGeo<C> g;
  if (dma) d_geo<C>(D, row0, col0, wave, lane, g);
  // fragment offsets within a stage
  const int l16 = lane & 15, lg = lane >> 4;
  int aofs[C::FM], bofs[C::KS][C::FN];
  {
    const int r = l16 >> 2, q = l16 & 3;
    const int x = d_swz_a(8 * lg + r);
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi) {
      const int c = ((wm * C::WTM + mi * 16) >> 3) + (q >> 1);
      aofs[mi] = (8 * lg + r) * C::A_ROW + ((c ^ x) << 4) + ((q & 1) << 3);
    }
#pragma unroll
    for (int kk = 0; kk < C::KS; ++kk)
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) {
        const int j = wn * C::WTN + ni * 16 + l16;
        bofs[kk][ni] = j * C::B_ROW + (((4 * kk + lg) ^ d_swz_b<C::B_CH>(j)) << 4);
      }
  }
  const int nfull = depth / C::BK;
  const int rem = depth - nfull * C::BK;
  auto stage = [&](int s) { return lds + (s % C::NST) * C::STAGE; };
  auto fill = [&](int s) {
    if (dma) d_issue<C>(at, bt, D.lda, alo, blo, s * C::BK, g, stage(s), wave);
    else d_stage_slow<C>(D, at, bt, alo, blo, row0, col0, s * C::BK, C::BK, stage(s), wave, lane);
  };
  // Every stage buffer is filled up front; the barrier that publishes stage
  // s+1 sits inside stage s (after all of stage s's LDS reads, before its
  // last MFMAs) and frees stage s's buffer for stage s + NST.
#pragma unroll
  for (int s = 0; s < C::NST; ++s)
    if (s < nfull) fill(s);
  if (nfull > 0) {
    d_wait_stages<C>(min(C::NST - 1, nfull - 1));
    __syncthreads();
  }
  for (int s = 0; s < nfull; ++s) {
    d_stage<C>(stage(s), aofs, bofs, acc, [&]() {
      if (s + 1 < nfull) {
        d_wait_stages<C>(min(C::NST - 2, nfull - 2 - s));
        __syncthreads();
        if (s + C::NST < nfull) fill(s + C::NST);
      }
    });
  }
  if (rem > 0) {
    __syncthreads();
    d_stage_slow<C>(D, at, bt, alo, blo, row0, col0, nfull * C::BK, rem, stage(nfull), wave, lane);
    __syncthreads();
    d_stage<C>(stage(nfull), aofs, bofs, acc, []() {});
  }
  __syncthreads();   // the next unit's first loads overwrite stages read here
  if (sgn < 0.0) {
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) acc[mi][ni] = -acc[mi][ni];
  }
}

// One unit (product group q of tile position p of job J).
template <class C>
__device__ __forceinline__ void d_unit(const Launch& L, const Job& J, int q, int p, char* lds, int tid) {
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave % C::WM, wn = wave / C::WM;
  int tmi, tni;
  tile_coords<4>(J.tiles_m, J.tiles_n, p, &tmi, &tni);
  const int row0 = tmi * C::BM, col0 = tni * C::BN;
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
    f32x4 acc[C::FM][C::FN];
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};
    d_contract<C>(J, P, row0, col0, depth, lds, wave, lane, wm, wn, acc);
    if (ordered && q > 0 && qq == q0) wait_turn(ctr, q, L);
    l_scatter<C>(J, P, row0, col0, lane, wm, wn, acc);
  }
  if (ordered) pass_turn(ctr, q + 1, L);
}

// The unit ticket is broadcast through the first word of the last stage
// buffer: it is read (and a barrier passed) before a unit's first loads, which
// fill stages 0..NST-2, and the last stage is first written one barrier later.
template <class C>
__global__ void __launch_bounds__(C::NT) h16d_kernel(Launch L) {
  __shared__ __align__(16) char lds[C::LDS_BYTES];
  const int tid = threadIdx.x;
  run_units(L, reinterpret_cast<int*>(lds + (C::NST - 1) * C::STAGE),
            [&](int j, int q, int p) { d_unit<C>(L, job_at(L, j), q, p, lds, tid); });
}

// C tile, wave grid, stage depth and ring depth of the direct kernels.
// Measured (BF16 GEMM, TFLOP/s at 8192^3 / 9600^3 / 16384^3): 32-deep
// stages in a ring of 4: 1000-1090 / 1030-1070 / 1050-1075; 64-deep stages,
// ring of 2: 1070-1095 / 1035-1040 / 1135-1150 (the default); 3 stages or
// 64 x 128 wave tiles or a 256 x 128 C tile: 2-10% slower.  Moving the
// stage barrier inside the stage (TAILM) is neutral here and +7% for the
// split kernel.
#ifndef FMM_H16D_CFG
#define FMM_H16D_CFG 256, 256, 2, 4, 64, 2
#endif
typedef DCfg<kInBF16, FMM_H16D_CFG> CfgBF16D;
typedef DCfg<kInF16, FMM_H16D_CFG> CfgF16D;
// split-BF16: two planes per stage (64 KiB), so two stages
#ifndef FMM_SPLITD_CFG
#define FMM_SPLITD_CFG 256, 256, 2, 4, 32, 2
#endif
typedef DCfg<kInBF16, FMM_SPLITD_CFG, true> CfgSplitD;

}  // namespace fmm
