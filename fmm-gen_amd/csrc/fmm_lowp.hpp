// fmm_lowp.hpp -- low-precision members of the ABC kernel family (gfx950).
//
// Same jobs, units, product order and C hand-over as the FP64 kernel of
// fmm_core.hpp; what changes is the arithmetic of a unit:
//   * A and B hold FP32, BF16 or FP16 elements; C is FP32.
//   * Operand sums  S = sum_t u_t A_t,  T = sum_t v_t B_t  are formed in FP32
//     from the loaded elements (s = u0*x0, then s = fma(u_t, x_t, s), the
//     order of the reference's packing, meta/abc_gen_mix.py:229-251), then
//     rounded once to the MFMA operand type (round to nearest even).
//   * Products run on MFMA with FP32 accumulation over the whole leaf depth:
//       FP32 operands   v_mfma_f32_16x16x4_f32
//       BF16 operands   v_mfma_f32_16x16x32_bf16
//       FP16 operands   v_mfma_f32_16x16x32_f16
//   * Split-BF16 (compensated) products: FP32 inputs, each operand sum split
//     as S = S_hi + S_lo with S_hi = bf16(S), S_lo = bf16(S - S_hi); a
//     product is S_hi T_hi + S_hi T_lo + S_lo T_hi (three BF16 MFMAs, FP32
//     accumulation), about 16 significant bits per operand instead of 8.
//   * W scatter: C_j = fma(w_j, acc, C_j) in FP32, 16-byte accesses.
//
// MFMA register layout (16x16 output, FP32 accumulators): lane l supplies
// row l%16 of the left operand and column l%16 of the right operand; its
// k values are k = l/16 (x4 f32) or k = 8*(l/16) + e, e = 0..7 (x32 16-bit);
// accumulator element v of lane l is D[4*(l/16) + v][l%16].  The left operand
// is A (rows of C), the right one B, so each lane owns 4 consecutive rows of
// one C column: one 16-byte read-modify-write of column-major C.
//
// LDS layouts.  16-bit operands: As[i][k] and Bs[j][k], k contiguous, row
// stride BK+8 elements, so a lane's 8 k values are one 16-byte read and the
// 16 rows read together fall on distinct banks.  A arrives column-major (i
// contiguous), so the packing step transposes it: a thread loads an AR x 8
// block (8 loads of AR consecutive rows), and writes AR rows of 8 k.  FP32
// operands: As[k][i] (row stride BM+16 floats, i contiguous as loaded) and
// Bs[j][k] (row stride BK+2), each MFMA operand one 4-byte read.
#pragma once
#include <type_traits>

#include "fmm_core.hpp"

namespace fmm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef unsigned short u16;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) const float gcfloat;
typedef __attribute__((address_space(1))) float gfloat;
typedef __attribute__((address_space(1))) const u16 gcu16;
typedef __attribute__((address_space(1))) const f32x4 gcf32x4;
typedef __attribute__((address_space(1))) f32x4 gf32x4;
typedef __attribute__((address_space(1))) const u32x2 gcuint2;
typedef __attribute__((address_space(1))) const u32x4 gcuint4;
typedef __attribute__((address_space(1))) const f32x2 gcf32x2;

// Input element kinds and MFMA operand kinds.
enum : int { kInF32 = 0, kInBF16 = 1, kInF16 = 2 };
enum : int { kOpF32 = 0, kOpBF16 = 1, kOpF16 = 2 };

template <int IN>
struct InT {
  typedef u16 T;   // 16-bit inputs are carried as raw bits
};
template <>
struct InT<kInF32> {
  typedef float T;
};

__device__ __forceinline__ float bf16_bits_to_f(unsigned int b) { return __uint_as_float(b << 16); }
__device__ __forceinline__ float f16_bits_to_f(unsigned int b) {
  return (float)__builtin_bit_cast(_Float16, (unsigned short)(b & 0xffffu));
}

template <int IN>
__device__ __forceinline__ float in_to_f(typename InT<IN>::T x) {
  if constexpr (IN == kInF32) return x;
  else if constexpr (IN == kInBF16) return bf16_bits_to_f(x);
  else return f16_bits_to_f(x);
}

// Unpack a 32-bit word of two 16-bit inputs (element 0 in the low half).
template <int IN>
__device__ __forceinline__ void unpack2(unsigned int w, float& lo, float& hi) {
  if constexpr (IN == kInBF16) {
    lo = __uint_as_float(w << 16);
    hi = __uint_as_float(w & 0xffff0000u);
  } else {
    lo = f16_bits_to_f(w);
    hi = f16_bits_to_f(w >> 16);
  }
}

// Kernel configuration.
template <int IN_, int OP_, bool SPLIT_, int BM_, int BN_, int BK_, int WM_, int WN_, int CH_, int PF_ = 1,
          bool ONE_ = false>
struct LCfg {
  static constexpr int IN = IN_, OP = OP_;
  static constexpr bool SPLIT = SPLIT_;
  static constexpr int BM = BM_, BN = BN_, BK = BK_;
  static constexpr int WM = WM_, WN = WN_;            // waves along m and n
  static constexpr int CH = CH_;                      // operand terms per prefetched chunk
  static constexpr int PF = PF_;                      // slices of loads in flight (register ring depth)
  // Every product of the launch has one term per operand with coefficient
  // +-1 (plain GEMMs, Naive products): no operand-sum code at all.
  static constexpr bool ONE = ONE_;
  static constexpr int NT = 64 * WM * WN;
  static constexpr int WTM = BM / WM, WTN = BN / WN;  // wave tile
  static constexpr int FM = WTM / 16, FN = WTN / 16;  // MFMA blocks per wave
  static constexpr bool H16 = OP != kOpF32;           // 16-bit MFMA operands
  static constexpr int KM = H16 ? 32 : 4;             // k per MFMA
  static constexpr int NPL = SPLIT ? 2 : 1;           // operand planes (hi, lo)
  // A items: AR rows x AK k per item.  B items: BKV consecutive k of one column.
  static constexpr int AR = H16 ? ((BM * BK) / (8 * NT) >= 4 ? 4 : 2) : 4;
  static constexpr int AK = H16 ? ((BM * BK) / (AR * NT) >= 8 ? 8 : 4) : 1;
  static constexpr int BKV = H16 ? 8 : 4;
  static constexpr int NA = (BM * BK) / (AR * AK * NT);   // A items per thread
  static constexpr int NB = (BN * BK) / (BKV * NT);       // B items per thread
  // LDS strides (elements of the operand type).  16-bit rows are unpadded
  // and their 8-element k groups XOR-swizzled (swz()).
  static constexpr int SA = H16 ? BK : BM + 16;
  static constexpr int SB = H16 ? BK : BK + 2;
  static constexpr int NKG = BK / 8;                  // k groups per 16-bit row
  // Single-term +-1 operands go to LDS without the FP32 round trip when the
  // input type is the operand type.
  static constexpr bool RAWCOPY = !SPLIT && ((IN == kInF32 && OP == kOpF32) || (IN == kInBF16 && OP == kOpBF16) ||
                                             (IN == kInF16 && OP == kOpF16));
  static constexpr int A_ELEMS = H16 ? BM * SA : BK * SA;
  static constexpr int B_ELEMS = BN * SB;
  static constexpr int OPSZ = H16 ? 2 : 4;
  static constexpr int LDS_BYTES = 2 * NPL * (A_ELEMS + B_ELEMS) * OPSZ;
  static_assert(NA >= 1 && NA * AR * AK * NT == BM * BK, "A item layout");
  static_assert(NB >= 1 && NB * BKV * NT == BN * BK, "B item layout");
  static_assert(BK % KM == 0 && WTM % 16 == 0 && WTN % 16 == 0, "tile shape");
  static_assert(!SPLIT || (IN == kInF32 && OP == kOpBF16), "split products take FP32 inputs on BF16 MFMA");
  static_assert(IN != kInF32 || !H16 || AR == 2 || AR == 4, "FP32 A item");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS");
  static_assert(!H16 || NKG == 4 || NKG == 8, "16-bit slice depth 32 or 64");
  static_assert(PF == 1 || PF == 2, "prefetch depth");
  static_assert(!ONE || RAWCOPY, "single-term variant copies raw operands");
};

// Physical k group of logical k group `kg` in 16-bit LDS row (or column) r.
// With 8 groups (128-byte rows) the 8 rows of one 128-byte LDS pass, and the
// two 128-byte halves of a 256-byte pass, all fall on distinct banks for the
// MFMA reads (16 consecutive rows, one group) and for the transposing A
// writes (rows AR apart); likewise with 4 groups (64-byte rows).
template <class C>
__device__ __forceinline__ int swz(int r, int kg) {
  if constexpr (C::NKG == 8) return kg ^ ((r ^ (r >> 3)) & 7);
  else return kg ^ (((r >> 1) ^ (r >> 3)) & 3);
}

// Prefetched slice data: FP32 values, or two 16-bit inputs per 32-bit word
// (element 2w in the low half), widened only when summed.
template <class C>
struct LRaw {
  static constexpr int PW = C::IN == kInF32 ? 1 : 2;   // elements per word
  typedef typename std::conditional<C::IN == kInF32, float, unsigned int>::type W;
  W a[C::CH][C::NA][C::AK][C::AR / PW];
  W b[C::CH][C::NB][C::BKV / PW];
};

template <class C>
struct LSum {
  float a[C::NA][C::AK][C::AR];
  float b[C::NB][C::BKV];
};

template <class C>
struct LGeo {
  int row0, col0, depth;
  bool interior;
  long long aoff[C::NA];   // element offset of A item i at k = 0 within a block
  long long boff[C::NB];
  int ar[C::NA], ak[C::NA];   // item row / first k within the tile slice
  int bj[C::NB], bk[C::NB];   // item column / first k
};

template <class C>
struct LSrcs {
  const typename InT<C::IN>::T* a[C::CH];
  const typename InT<C::IN>::T* b[C::CH];
  float u[C::CH];
  float v[C::CH];
  int na, nb;
};

template <class C>
__device__ __forceinline__ void l_make_srcs(const Job& J, int ta0, int na, int tb0, int nb, LSrcs<C>& S) {
  typedef typename InT<C::IN>::T T;
  S.na = na < C::CH ? na : C::CH;
  S.nb = nb < C::CH ? nb : C::CH;
#pragma unroll
  for (int t = 0; t < C::CH; ++t) {
    if (t < S.na) {
      const Term tm = term_at(J, ta0 + t);
      const T* p = static_cast<const T*>(J.A) + (long long)tm.bi * J.ms + (long long)tm.bj * J.ks * J.lda;
      S.a[t] = unip(p);
      S.u[t] = (float)unid(tm.coef);
    } else {
      S.a[t] = nullptr;
      S.u[t] = 0.f;
    }
    if (t < S.nb) {
      const Term tm = term_at(J, tb0 + t);
      const T* p = static_cast<const T*>(J.B) + (long long)tm.bi * J.ks + (long long)tm.bj * J.ns * J.ldb;
      S.b[t] = unip(p);
      S.v[t] = (float)unid(tm.coef);
    } else {
      S.b[t] = nullptr;
      S.v[t] = 0.f;
    }
  }
}

// Load N consecutive inputs into raw words (N*size in {8, 16, 32} bytes,
// aligned to min(16, N*size)).
template <int IN, int N, class W>
__device__ __forceinline__ void load_words(const typename InT<IN>::T* p, W* out) {
  if constexpr (IN == kInF32) {
    static_assert(N == 2 || N % 4 == 0, "FP32 vector width");
    if constexpr (N == 2) {
      const f32x2 v = *(gcf32x2*)p;
      out[0] = v.x;
      out[1] = v.y;
    } else {
#pragma unroll
      for (int q = 0; q < N / 4; ++q) {
        const f32x4 v = *(gcf32x4*)(p + 4 * q);
        out[4 * q] = v.x;
        out[4 * q + 1] = v.y;
        out[4 * q + 2] = v.z;
        out[4 * q + 3] = v.w;
      }
    }
  } else {
    static_assert(N == 2 || N == 4 || N % 8 == 0, "16-bit vector width");
    if constexpr (N == 2) {
      out[0] = *(const __attribute__((address_space(1))) unsigned int*)p;
    } else if constexpr (N == 4) {
      const u32x2 v = *(gcuint2*)p;
      out[0] = v.x;
      out[1] = v.y;
    } else {
#pragma unroll
      for (int q = 0; q < N / 8; ++q) {
        const u32x4 v = *(gcuint4*)(p + 8 * q);
        out[4 * q] = v.x;
        out[4 * q + 1] = v.y;
        out[4 * q + 2] = v.z;
        out[4 * q + 3] = v.w;
      }
    }
  }
}

// Element-wise load of one input (zero when !ok) into word slot `e` of `w`.
template <int IN, class W>
__device__ __forceinline__ void put_elem(W* w, int e, bool ok, const typename InT<IN>::T* p) {
  typedef typename InT<IN>::T T;
  if constexpr (IN == kInF32) {
    w[e] = ok ? *(const __attribute__((address_space(1))) T*)p : 0.f;
  } else {
    const unsigned int v = ok ? (unsigned int)*(const __attribute__((address_space(1))) T*)p : 0u;
    if (e & 1) w[e >> 1] |= v << 16;
    else w[e >> 1] = v;
  }
}

// Global loads of slice kb (zero outside the leaf block and the depth).
template <class C>
__device__ __forceinline__ void l_load_raw(const Job& J, const LSrcs<C>& S, int kb, const LGeo<C>& g, LRaw<C>& r) {
  typedef typename InT<C::IN>::T T;
  const bool full = g.interior && kb + C::BK <= g.depth;
  if (full) {
    const long long ka = (long long)kb * J.lda;
#pragma unroll
    for (int t = 0; t < C::CH; ++t) {
      if (t < S.na) {
#pragma unroll
        for (int i = 0; i < C::NA; ++i)
#pragma unroll
          for (int e = 0; e < C::AK; ++e)
            load_words<C::IN, C::AR>(S.a[t] + ka + g.aoff[i] + (long long)e * J.lda, r.a[t][i][e]);
      }
      if (t < S.nb) {
#pragma unroll
        for (int i = 0; i < C::NB; ++i) load_words<C::IN, C::BKV>(S.b[t] + kb + g.boff[i], r.b[t][i]);
      }
    }
    return;
  }
#pragma unroll
  for (int t = 0; t < C::CH; ++t) {
    if (t < S.na) {
#pragma unroll
      for (int i = 0; i < C::NA; ++i)
#pragma unroll
        for (int e = 0; e < C::AK; ++e) {
          const bool kok = kb + g.ak[i] + e < g.depth;
#pragma unroll
          for (int q = 0; q < C::AR; ++q) {
            const bool ok = kok && g.row0 + g.ar[i] + q < J.ms;
            put_elem<C::IN>(r.a[t][i][e], q, ok, S.a[t] + g.aoff[i] + (long long)(kb + e) * J.lda + q);
          }
        }
    }
    if (t < S.nb) {
#pragma unroll
      for (int i = 0; i < C::NB; ++i) {
        const bool cok = g.col0 + g.bj[i] < J.ns;
        const int krem = g.depth - (kb + g.bk[i]);
#pragma unroll
        for (int e = 0; e < C::BKV; ++e) put_elem<C::IN>(r.b[t][i], e, cok && e < krem, S.b[t] + kb + g.boff[i] + e);
      }
    }
  }
}

// Widen raw element e of a word array.
template <int IN, class W>
__device__ __forceinline__ float raw_at(const W* w, int e) {
  if constexpr (IN == kInF32) return w[e];
  else if constexpr (IN == kInBF16) return (e & 1) ? __uint_as_float(w[e >> 1] & 0xffff0000u) : __uint_as_float(w[e >> 1] << 16);
  else return f16_bits_to_f((e & 1) ? (w[e >> 1] >> 16) : w[e >> 1]);
}

template <class C>
__device__ __forceinline__ void l_accumulate(const LSrcs<C>& S, bool first, const LRaw<C>& r, LSum<C>& s, bool da,
                                             bool db) {
#pragma unroll
  for (int t = 0; t < C::CH; ++t) {
    if (da && t < S.na) {
#pragma unroll
      for (int i = 0; i < C::NA; ++i)
#pragma unroll
        for (int e = 0; e < C::AK; ++e)
#pragma unroll
          for (int q = 0; q < C::AR; ++q)
            s.a[i][e][q] = (first && t == 0) ? S.u[t] * raw_at<C::IN>(r.a[t][i][e], q)
                                             : __builtin_fmaf(S.u[t], raw_at<C::IN>(r.a[t][i][e], q), s.a[i][e][q]);
    }
    if (db && t < S.nb) {
#pragma unroll
      for (int i = 0; i < C::NB; ++i)
#pragma unroll
        for (int e = 0; e < C::BKV; ++e)
          s.b[i][e] = (first && t == 0) ? S.v[t] * raw_at<C::IN>(r.b[t][i], e)
                                        : __builtin_fmaf(S.v[t], raw_at<C::IN>(r.b[t][i], e), s.b[i][e]);
    }
  }
}

// Round N FP32 values (N = 4 or 8) to the 16-bit operand type, two per word
// (hi plane) and, for split products, the remainders (lo plane).
template <int OP>
__device__ __forceinline__ unsigned int pack2(float x0, float x1) {
  if constexpr (OP == kOpBF16) {
    const __bf16 a = (__bf16)x0, b = (__bf16)x1;
    return (unsigned int)__builtin_bit_cast(u16, a) | ((unsigned int)__builtin_bit_cast(u16, b) << 16);
  } else {
    const _Float16 a = (_Float16)x0, b = (_Float16)x1;
    return (unsigned int)__builtin_bit_cast(u16, a) | ((unsigned int)__builtin_bit_cast(u16, b) << 16);
  }
}

template <int OP, int N>
__device__ __forceinline__ void packN(const float (&x)[N], unsigned int (&w)[N / 2]) {
#pragma unroll
  for (int e = 0; e < N / 2; ++e) w[e] = pack2<OP>(x[2 * e], x[2 * e + 1]);
}

template <int N>
__device__ __forceinline__ void splitN(const float (&x)[N], unsigned int (&hi)[N / 2], unsigned int (&lo)[N / 2]) {
#pragma unroll
  for (int e = 0; e < N / 2; ++e) {
    const __bf16 h0 = (__bf16)x[2 * e], h1 = (__bf16)x[2 * e + 1];
    hi[e] = (unsigned int)__builtin_bit_cast(u16, h0) | ((unsigned int)__builtin_bit_cast(u16, h1) << 16);
    lo[e] = pack2<kOpBF16>(x[2 * e] - (float)h0, x[2 * e + 1] - (float)h1);
  }
}

// Store N/2 words (N = 4: 8 bytes, N = 8: 16 bytes) to LDS.
template <int N>
__device__ __forceinline__ void st_words(u16* p, const unsigned int (&w)[N / 2]) {
  if constexpr (N == 8) *reinterpret_cast<u32x4*>(p) = u32x4{w[0], w[1], w[2], w[3]};
  else *reinterpret_cast<u32x2*>(p) = u32x2{w[0], w[1]};
}

// Lo or hi 16-bit half of each of two words, packed (lo-word half first).
__device__ __forceinline__ unsigned int pick16(unsigned int w0, unsigned int w1, int hi) {
  return __builtin_amdgcn_perm(w1, w0, hi ? 0x07060302u : 0x05040100u);
}

// Operand sums of slice kb into LDS buffer (As, Bs).  `r` holds chunk 0 of
// the slice (issued earlier); products with more than CH terms load their
// remaining chunks here.  fa / fb: the A / B operand is one term with
// coefficient +-1 and goes to LDS straight from the raw words (sign flipped
// for -1); otherwise sums are formed in FP32 and rounded.
template <class C>
__device__ __forceinline__ void l_pack(const Job& J, const Prod& P, const LSrcs<C>& S0, int kb, const LGeo<C>& g,
                                       LRaw<C>& r, char* As, char* Bs, bool fa, bool fb) {
  LSum<C> s;
  if (!(fa && fb)) {
    l_accumulate<C>(S0, true, r, s, !fa, !fb);
    const int nmax = P.na > P.nb ? P.na : P.nb;
    for (int c = C::CH; c < nmax; c += C::CH) {
      LSrcs<C> S;
      l_make_srcs<C>(J, P.a0 + c, P.na - c, P.b0 + c, P.nb - c, S);
      l_load_raw<C>(J, S, kb, g, r);
      l_accumulate<C>(S, false, r, s, !fa, !fb);
    }
  }
  if constexpr (C::H16) {
    u16* a = reinterpret_cast<u16*>(As);
    u16* b = reinterpret_cast<u16*>(Bs);
#pragma unroll
    for (int i = 0; i < C::NA; ++i)
#pragma unroll
      for (int q = 0; q < C::AR; ++q) {
        constexpr int AK = C::AK;
        const int row = g.ar[i] + q;
        const int off = row * C::SA + swz<C>(row, g.ak[i] >> 3) * 8 + (g.ak[i] & 7);
        if constexpr (C::RAWCOPY) {
          if (fa) {
            const unsigned int sg = S0.u[0] < 0.f ? 0x80008000u : 0u;
            unsigned int w[AK / 2];
#pragma unroll
            for (int e = 0; e < AK / 2; ++e) w[e] = pick16(r.a[0][i][2 * e][q >> 1], r.a[0][i][2 * e + 1][q >> 1], q & 1) ^ sg;
            st_words<AK>(a + off, w);
            continue;
          }
        }
        float x[AK];
#pragma unroll
        for (int e = 0; e < AK; ++e) x[e] = s.a[i][e][q];
        if constexpr (C::SPLIT) {
          unsigned int hi[AK / 2], lo[AK / 2];
          splitN<AK>(x, hi, lo);
          st_words<AK>(a + off, hi);
          st_words<AK>(a + C::A_ELEMS + off, lo);
        } else {
          unsigned int w[AK / 2];
          packN<C::OP, AK>(x, w);
          st_words<AK>(a + off, w);
        }
      }
#pragma unroll
    for (int i = 0; i < C::NB; ++i) {
      const int off = g.bj[i] * C::SB + swz<C>(g.bj[i], g.bk[i] >> 3) * 8;
      if constexpr (C::RAWCOPY) {
        if (fb) {
          const unsigned int sg = S0.v[0] < 0.f ? 0x80008000u : 0u;
          u32x4 v;
          v.x = r.b[0][i][0] ^ sg;
          v.y = r.b[0][i][1] ^ sg;
          v.z = r.b[0][i][2] ^ sg;
          v.w = r.b[0][i][3] ^ sg;
          *reinterpret_cast<u32x4*>(b + off) = v;
          continue;
        }
      }
      float x[8];
#pragma unroll
      for (int e = 0; e < 8; ++e) x[e] = s.b[i][e];
      if constexpr (C::SPLIT) {
        unsigned int hi[4], lo[4];
        splitN<8>(x, hi, lo);
        st_words<8>(b + off, hi);
        st_words<8>(b + C::B_ELEMS + off, lo);
      } else {
        unsigned int w[4];
        packN<C::OP, 8>(x, w);
        st_words<8>(b + off, w);
      }
    }
  } else {
    float* a = reinterpret_cast<float*>(As);
    float* b = reinterpret_cast<float*>(Bs);
#pragma unroll
    for (int i = 0; i < C::NA; ++i) {
      f32x4 v;
      if (C::RAWCOPY && fa) {
        const float u = S0.u[0];
        v = f32x4{u * (float)r.a[0][i][0][0], u * (float)r.a[0][i][0][1], u * (float)r.a[0][i][0][2],
                  u * (float)r.a[0][i][0][3]};
      } else {
        v = f32x4{s.a[i][0][0], s.a[i][0][1], s.a[i][0][2], s.a[i][0][3]};
      }
      *reinterpret_cast<f32x4*>(a + g.ak[i] * C::SA + g.ar[i]) = v;
    }
#pragma unroll
    for (int i = 0; i < C::NB; ++i) {
      float* p = b + g.bj[i] * C::SB + g.bk[i];
      f32x2 v0, v1;
      if (C::RAWCOPY && fb) {
        const float v = S0.v[0];
        v0 = f32x2{v * (float)r.b[0][i][0], v * (float)r.b[0][i][1]};
        v1 = f32x2{v * (float)r.b[0][i][2], v * (float)r.b[0][i][3]};
      } else {
        v0 = f32x2{s.b[i][0], s.b[i][1]};
        v1 = f32x2{s.b[i][2], s.b[i][3]};
      }
      *reinterpret_cast<f32x2*>(p) = v0;
      *reinterpret_cast<f32x2*>(p + 2) = v1;
    }
  }
}

// MFMAs over k-steps [K0, K1) of the slice in LDS buffer (As, Bs).
template <class C, int K0, int K1>
__device__ __forceinline__ void l_mfma(const char* As, const char* Bs, int lane, int wm, int wn,
                                       f32x4 (&acc)[C::FM][C::FN]) {
  const int l16 = lane & 15, lk = lane >> 4;
  __builtin_amdgcn_s_setprio(1);
  if constexpr (C::H16) {
    const u16* a = reinterpret_cast<const u16*>(As);
    const u16* b = reinterpret_cast<const u16*>(Bs);
#pragma unroll
    for (int kk = K0; kk < K1; ++kk) {
      const int kg = kk * 4 + lk;
      u32x4 bh[C::FN], bl[C::FN];
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) {
        const int col = wn * C::WTN + ni * 16 + l16;
        const int off = col * C::SB + swz<C>(col, kg) * 8;
        bh[ni] = *reinterpret_cast<const u32x4*>(b + off);
        if constexpr (C::SPLIT) bl[ni] = *reinterpret_cast<const u32x4*>(b + C::B_ELEMS + off);
      }
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi) {
        const int row = wm * C::WTM + mi * 16 + l16;
        const int off = row * C::SA + swz<C>(row, kg) * 8;
        const u32x4 ah = *reinterpret_cast<const u32x4*>(a + off);
        if constexpr (C::SPLIT) {
          const u32x4 al = *reinterpret_cast<const u32x4*>(a + C::A_ELEMS + off);
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni) {
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, al),
                                                                  __builtin_bit_cast(bf16x8, bh[ni]), acc[mi][ni], 0, 0, 0);
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah),
                                                                  __builtin_bit_cast(bf16x8, bl[ni]), acc[mi][ni], 0, 0, 0);
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah),
                                                                  __builtin_bit_cast(bf16x8, bh[ni]), acc[mi][ni], 0, 0, 0);
          }
        } else if constexpr (C::OP == kOpBF16) {
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, ah),
                                                                  __builtin_bit_cast(bf16x8, bh[ni]), acc[mi][ni], 0, 0, 0);
        } else {
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni)
            acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, ah),
                                                                 __builtin_bit_cast(f16x8, bh[ni]), acc[mi][ni], 0, 0, 0);
        }
      }
    }
  } else {
    const float* a = reinterpret_cast<const float*>(As);
    const float* b = reinterpret_cast<const float*>(Bs);
#pragma unroll
    for (int kk = K0; kk < K1; ++kk) {
      const int k = kk * 4 + lk;
      float bv[C::FN];
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) bv[ni] = b[(wn * C::WTN + ni * 16 + l16) * C::SB + k];
#pragma unroll
      for (int mi = 0; mi < C::FM; ++mi) {
        const float av = a[k * C::SA + wm * C::WTM + mi * 16 + l16];
#pragma unroll
        for (int ni = 0; ni < C::FN; ++ni)
          acc[mi][ni] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv[ni], acc[mi][ni], 0, 0, 0);
      }
    }
  }
  __builtin_amdgcn_s_setprio(0);
}

// One slice of the pipelined contraction: MFMAs on LDS buffer `cur` while
// the sums of slice st+1 (its loads were issued PF slices earlier into `r`)
// go to the other buffer and the loads of slice st+1+PF are issued into `r`.
// The two halves of the workgroup's waves run the phases in opposite order,
// so one half's packing overlaps the other half's MFMAs on the same SIMD.
template <class C>
__device__ __forceinline__ void l_step(const Job& J, const Prod& P, const LSrcs<C>& S, int st, int nsteps,
                                       const LGeo<C>& g, LRaw<C>& r, char* lds, int lane, int wm, int wn, int grp,
                                       bool fa, bool fb, f32x4 (&acc)[C::FM][C::FN]) {
  constexpr int KS = C::BK / C::KM;   // MFMA k-steps per slice
  constexpr int BUF = C::NPL * (C::A_ELEMS + C::B_ELEMS) * C::OPSZ;
  constexpr int BOFF = C::NPL * C::A_ELEMS * C::OPSZ;
  char* cur = lds + (st & 1) * BUF;
  char* nxt = lds + ((st & 1) ^ 1) * BUF;
  auto sums = [&]() {
    if (st + 1 < nsteps) {
      l_pack<C>(J, P, S, (st + 1) * C::BK, g, r, nxt, nxt + BOFF, fa, fb);
      if (st + 1 + C::PF < nsteps) l_load_raw<C>(J, S, (st + 1 + C::PF) * C::BK, g, r);
    }
  };
  if (grp == 0) {
    l_mfma<C, 0, KS / 2>(cur, cur + BOFF, lane, wm, wn, acc);
    sums();
    l_mfma<C, KS / 2, KS>(cur, cur + BOFF, lane, wm, wn, acc);
  } else {
    sums();
    l_mfma<C, 0, KS>(cur, cur + BOFF, lane, wm, wn, acc);
  }
  __syncthreads();
}

// Contraction of one unit: double-buffered LDS slices, one barrier per slice,
// PF register buffers of raw loads in flight (slice s uses buffer s % PF
// after the first; the loop is unrolled by PF so the buffer choice is
// static).
template <class C>
__device__ __forceinline__ void l_contract(const Job& J, const Prod& P, const LGeo<C>& g, char* lds, int lane, int wm,
                                           int wn, int grp, f32x4 (&acc)[C::FM][C::FN]) {
  constexpr int BOFF = C::NPL * C::A_ELEMS * C::OPSZ;
  const int nsteps = (g.depth + C::BK - 1) / C::BK;
  if (!(P.na > 0 && P.nb > 0 && nsteps > 0)) return;
  LSrcs<C> S;
  l_make_srcs<C>(J, P.a0, P.na, P.b0, P.nb, S);
  const bool fa = C::ONE || (C::RAWCOPY && P.na == 1 && __builtin_fabsf(S.u[0]) == 1.f);
  const bool fb = C::ONE || (C::RAWCOPY && P.nb == 1 && __builtin_fabsf(S.v[0]) == 1.f);
  if constexpr (C::PF == 1) {
    LRaw<C> r;
    l_load_raw<C>(J, S, 0, g, r);
    l_pack<C>(J, P, S, 0, g, r, lds, lds + BOFF, fa, fb);
    if (nsteps > 1) l_load_raw<C>(J, S, C::BK, g, r);
    __syncthreads();
    for (int st = 0; st < nsteps; ++st) l_step<C>(J, P, S, st, nsteps, g, r, lds, lane, wm, wn, grp, fa, fb, acc);
  } else {
    // slice 0 via ra; then odd slices in ra, even slices in rb
    LRaw<C> ra, rb;
    l_load_raw<C>(J, S, 0, g, ra);
    l_pack<C>(J, P, S, 0, g, ra, lds, lds + BOFF, fa, fb);
    if (nsteps > 1) l_load_raw<C>(J, S, C::BK, g, ra);
    if (nsteps > 2) l_load_raw<C>(J, S, 2 * C::BK, g, rb);
    __syncthreads();
    for (int st = 0; st < nsteps; st += 2) {
      l_step<C>(J, P, S, st, nsteps, g, ra, lds, lane, wm, wn, grp, fa, fb, acc);
      if (st + 1 < nsteps) l_step<C>(J, P, S, st + 1, nsteps, g, rb, lds, lane, wm, wn, grp, fa, fb, acc);
    }
  }
}

// C_j = C_j + w_j * acc (C_j = w_j * acc for a store job) in FP32: lane l
// holds rows 4*(l/16)+0..3 of column l%16 of each 16x16 block.  In a tile
// wholly inside the leaf block every read of a term is issued before its
// first write (with a run-time ldc the compiler cannot prove two elements
// distinct and would otherwise wait out each read-modify-write in turn).
template <class C>
__device__ __forceinline__ void l_scatter(const Job& J, const Prod& P, int row0, int col0, int lane, int wm, int wn,
                                          const f32x4 (&acc)[C::FM][C::FN]) {
  const int l16 = lane & 15, lk = lane >> 4;
  const bool store = (J.flags & kJobStore) != 0;
  const bool cvec = (J.flags & 4) != 0;
  const bool inside = cvec && row0 + C::BM <= J.ms && col0 + C::BN <= J.ns;
  for (int t = 0; t < P.nc; ++t) {
    const Term tm = term_at(J, P.c0 + t);
    const float w = (float)tm.coef;
    gfloat* cb = (gfloat*)static_cast<float*>(J.C) + (long long)tm.bi * J.ms + (long long)tm.bj * J.ns * J.ldc;
    if (inside) {
      // RG block rows per round: all their reads, then all their writes
      constexpr int RG = C::FM % 2 == 0 ? 2 : 1;
#pragma unroll
      for (int m0 = 0; m0 < C::FM; m0 += RG) {
        f32x4 old[RG][C::FN];
        if (!store) {
#pragma unroll
          for (int mi = 0; mi < RG; ++mi)
#pragma unroll
            for (int ni = 0; ni < C::FN; ++ni)
              old[mi][ni] = *(gf32x4*)(cb + row0 + wm * C::WTM + (m0 + mi) * 16 + 4 * lk +
                                       (long long)(col0 + wn * C::WTN + ni * 16 + l16) * J.ldc);
        }
#pragma unroll
        for (int mi = 0; mi < RG; ++mi)
#pragma unroll
          for (int ni = 0; ni < C::FN; ++ni) {
            const f32x4 v = acc[m0 + mi][ni];
            f32x4 o;
            if (store) {
              o = f32x4{w * v.x, w * v.y, w * v.z, w * v.w};
            } else {
              const f32x4 x = old[mi][ni];
              o.x = __builtin_fmaf(w, v.x, x.x);
              o.y = __builtin_fmaf(w, v.y, x.y);
              o.z = __builtin_fmaf(w, v.z, x.z);
              o.w = __builtin_fmaf(w, v.w, x.w);
            }
            *(gf32x4*)(cb + row0 + wm * C::WTM + (m0 + mi) * 16 + 4 * lk +
                       (long long)(col0 + wn * C::WTN + ni * 16 + l16) * J.ldc) = o;
          }
      }
      continue;
    }
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi) {
      const int row = row0 + wm * C::WTM + mi * 16 + 4 * lk;
      if (row >= J.ms) continue;
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) {
        const int col = col0 + wn * C::WTN + ni * 16 + l16;
        if (col >= J.ns) continue;
        gfloat* c = cb + row + (long long)col * J.ldc;
        const f32x4 v = acc[mi][ni];
        if (cvec && row + 4 <= J.ms) {
          f32x4 o;
          if (store) {
            o = f32x4{w * v.x, w * v.y, w * v.z, w * v.w};
          } else {
            const f32x4 x = *(gf32x4*)c;
            o.x = __builtin_fmaf(w, v.x, x.x);
            o.y = __builtin_fmaf(w, v.y, x.y);
            o.z = __builtin_fmaf(w, v.z, x.z);
            o.w = __builtin_fmaf(w, v.w, x.w);
          }
          *(gf32x4*)c = o;
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (row + e < J.ms) c[e] = store ? w * v[e] : __builtin_fmaf(w, v[e], c[e]);
        }
      }
    }
  }
}

// One unit (product group q of tile position p of job J) with configuration C.
template <class C>
__device__ __forceinline__ void l_unit(const Launch& L, const Job& J, int q, int p, char* lds, int tid) {
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave % C::WM, wn = wave / C::WM;
  int tmi, tni;
  tile_coords<4>(J.tiles_m, J.tiles_n, p, &tmi, &tni);
  LGeo<C> g;
  g.row0 = tmi * C::BM;
  g.col0 = tni * C::BN;
  g.interior = (J.flags & 3) == 3 && g.row0 + C::BM <= J.ms && g.col0 + C::BN <= J.ns;
#pragma unroll
  for (int it = 0; it < C::NA; ++it) {
    const int v = tid + C::NT * it;
    constexpr int RG = C::BM / C::AR;   // row groups per slice
    g.ar[it] = (v % RG) * C::AR;
    g.ak[it] = (v / RG) * C::AK;
    g.aoff[it] = g.row0 + g.ar[it] + (long long)g.ak[it] * J.lda;
  }
#pragma unroll
  for (int it = 0; it < C::NB; ++it) {
    const int v = tid + C::NT * it;
    constexpr int KG = C::BK / C::BKV;   // k groups per column
    g.bk[it] = (v % KG) * C::BKV;
    g.bj[it] = v / KG;
    g.boff[it] = (long long)(g.col0 + g.bj[it]) * J.ldb + g.bk[it];
  }

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

    f32x4 acc[C::FM][C::FN];
#pragma unroll
    for (int mi = 0; mi < C::FM; ++mi)
#pragma unroll
      for (int ni = 0; ni < C::FN; ++ni) acc[mi][ni] = f32x4{0.f, 0.f, 0.f, 0.f};

    l_contract<C>(J, P, g, lds, lane, wm, wn, wave / (C::WM * C::WN / 2), acc);

    if (ordered && q > 0 && qq == q0) wait_turn(ctr, q, L);
    l_scatter<C>(J, P, g.row0, g.col0, lane, wm, wn, acc);
  }
  if (ordered) pass_turn(ctr, q + 1, L);
}

// The persistent kernel: C1 runs jobs with the default tile shape, C2 the
// jobs flagged kJobTileT (both: same element types and thread count).
template <class C1, class C2>
__global__ void __launch_bounds__(C1::NT) abc_lowp_kernel(Launch L) {
  static_assert(C1::NT == C2::NT && C1::IN == C2::IN && C1::OP == C2::OP && C1::SPLIT == C2::SPLIT, "kernel pair");
  constexpr int LDSB = C1::LDS_BYTES > C2::LDS_BYTES ? C1::LDS_BYTES : C2::LDS_BYTES;
  __shared__ __align__(16) char lds[LDSB];
  __shared__ int s_unit;
  const int tid = threadIdx.x;
  run_units(L, &s_unit, [&](int j, int q, int p) {
    const Job& J = job_at(L, j);
    const int tt = __builtin_amdgcn_readfirstlane(J.flags & kJobTileT);
    if (tt && !__is_same(C1, C2)) l_unit<C2>(L, J, q, p, lds, tid);
    else l_unit<C1>(L, J, q, p, lds, tid);
  });
}

// --- AB and Naive pack types with low-precision inputs --------------------
//
// Same structure as presum_kernel / madd_kernel (fmm_core.hpp): one leaf
// block per grid.y, kEwRows x kEwCols tiles, two consecutive rows per lane.
// The operand sums are formed in FP32 (s = u0*x0, s = fma(u_t, x_t, s)) and
// stored rounded to the input type -- the rounding the ABC kernel applies when
// it writes the sum to LDS -- so a product of the stored sums equals the ABC
// product bit for bit; M and C are FP32.
struct SumJobL {
  const void* X;
  void* Y;
  long long ldx, ldy;
  int rows, cols;
  int ny;
  int side;
  int vec;                  // 1: X aligned to two elements, even ldx and rows (VEC instance)
  int pad1;
  const Prod* prods;
  const Term* terms;
  long long plane;          // split output: element offset of the lo plane
};

template <int IN>
__device__ __forceinline__ void store_in(typename InT<IN>::T* p, float v) {
  if constexpr (IN == kInF32) {
    *(__attribute__((address_space(1))) float*)p = v;
  } else if constexpr (IN == kInBF16) {
    const __bf16 h = (__bf16)v;
    *(__attribute__((address_space(1))) u16*)p = __builtin_bit_cast(u16, h);
  } else {
    const _Float16 h = (_Float16)v;
    *(__attribute__((address_space(1))) u16*)p = __builtin_bit_cast(u16, h);
  }
}

// SPLIT (FP32 inputs): the sum is stored as two BF16 planes, hi = bf16(s)
// at Y and lo = bf16(s - hi) at Y + plane -- the split the split-BF16
// kernels apply to an FP32 operand in registers.
template <int IN, bool SPLIT = false, bool VEC = false>
__global__ void __launch_bounds__(kEwThreads) presum_lowp_kernel(SumJobL S) {
  typedef typename InT<IN>::T T;
  static_assert(!SPLIT || IN == kInF32, "split planes from FP32 sums");
  const int r = (int)blockIdx.y;
  int r0, c0;
  ew_tile(S.rows, S.cols, &r0, &c0);
  if (c0 >= S.cols) return;
  // product and term records as scalar loads; the wave index made uniform
  // so that the column loop, and the record reads in it, are too
  const auto* prods = (const __attribute__((address_space(4))) Prod*)S.prods;
  const auto* terms = (const __attribute__((address_space(4))) Term*)S.terms;
  const int t0 = S.side ? prods[r].b0 : prods[r].a0;
  const int nt = S.side ? prods[r].nb : prods[r].na;
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = r0 + 2 * lane;
  T* y = static_cast<T*>(S.Y) + (long long)(r % S.ny) * S.rows + (long long)(r / S.ny) * S.cols * S.ldy;
  for (int c = c0 + wave; c < min(c0 + kEwCols, S.cols); c += kEwThreads / 64) {
    float s0 = 0.f, s1 = 0.f;
    // up to 4 terms' loads in flight, summed in term order
    for (int tb = 0; tb < nt; tb += 4) {
      float x0[4], x1[4], u[4];
      // every load is issued (terms past nt and rows past the block read a
      // clamped valid address, masked below), so the waits count loads
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int tq = min(tb + q, nt - 1);
        const int bi = terms[t0 + tq].bi, bj = terms[t0 + tq].bj;
        u[q] = (float)terms[t0 + tq].coef;
        const T* x = static_cast<const T*>(S.X) + (long long)bi * S.rows + (long long)bj * S.cols * S.ldx +
                     (long long)c * S.ldx;
        if constexpr (VEC) {
          const int ip = min(i, S.rows - 2);
          if constexpr (IN == kInF32) {
            const unsigned long long v = *(const __attribute__((address_space(1))) unsigned long long*)(x + ip);
            x0[q] = __uint_as_float((unsigned)v);
            x1[q] = __uint_as_float((unsigned)(v >> 32));
          } else {
            const unsigned w = *(const __attribute__((address_space(1))) unsigned*)(x + ip);
            x0[q] = in_to_f<IN>((T)(w & 0xffffu));
            x1[q] = in_to_f<IN>((T)(w >> 16));
          }
        } else {
          x0[q] = in_to_f<IN>(*(const __attribute__((address_space(1))) T*)(x + min(i, S.rows - 1)));
          x1[q] = in_to_f<IN>(*(const __attribute__((address_space(1))) T*)(x + min(i + 1, S.rows - 1)));
        }
        if (i >= S.rows) x0[q] = 0.f;
        if (i + 1 >= S.rows) x1[q] = 0.f;
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (tb + q < nt) {
          s0 = tb + q == 0 ? u[q] * x0[q] : __builtin_fmaf(u[q], x0[q], s0);
          s1 = tb + q == 0 ? u[q] * x1[q] : __builtin_fmaf(u[q], x1[q], s1);
        }
      }
    }
    if constexpr (SPLIT) {
      const long long off = (long long)(r % S.ny) * S.rows + (long long)(r / S.ny) * S.cols * S.ldy +
                            (long long)c * S.ldy;
      __attribute__((address_space(1))) u16* hc = (__attribute__((address_space(1))) u16*)S.Y + off;
      auto put = [&](int e, float v) {
        const __bf16 h = (__bf16)v;
        const __bf16 l = (__bf16)(v - (float)h);
        hc[e] = __builtin_bit_cast(u16, h);
        hc[e + S.plane] = __builtin_bit_cast(u16, l);
      };
      if (VEC) {
        // rows even, i even: both rows or neither; Y columns stay aligned
        if (i < S.rows) {
          const __bf16 h0 = (__bf16)s0, h1 = (__bf16)s1;
          const __bf16 l0 = (__bf16)(s0 - (float)h0), l1 = (__bf16)(s1 - (float)h1);
          *(__attribute__((address_space(1))) unsigned*)(hc + i) =
              __builtin_bit_cast(u16, h0) | ((unsigned)__builtin_bit_cast(u16, h1) << 16);
          *(__attribute__((address_space(1))) unsigned*)(hc + i + S.plane) =
              __builtin_bit_cast(u16, l0) | ((unsigned)__builtin_bit_cast(u16, l1) << 16);
        }
      } else {
        if (i < S.rows) put(i, s0);
        if (i + 1 < S.rows) put(i + 1, s1);
      }
    } else {
      T* yc = y + (long long)c * S.ldy;
      if (VEC && i < S.rows) {
        if constexpr (IN == kInF32) {
          *(__attribute__((address_space(1))) unsigned long long*)(yc + i) =
              (unsigned long long)__float_as_uint(s0) | ((unsigned long long)__float_as_uint(s1) << 32);
        } else if constexpr (IN == kInBF16) {
          const __bf16 h0 = (__bf16)s0, h1 = (__bf16)s1;
          *(__attribute__((address_space(1))) unsigned*)(yc + i) =
              __builtin_bit_cast(u16, h0) | ((unsigned)__builtin_bit_cast(u16, h1) << 16);
        } else {
          const _Float16 h0 = (_Float16)s0, h1 = (_Float16)s1;
          *(__attribute__((address_space(1))) unsigned*)(yc + i) =
              __builtin_bit_cast(u16, h0) | ((unsigned)__builtin_bit_cast(u16, h1) << 16);
        }
      } else if (!VEC) {
        if (i < S.rows) store_in<IN>(yc + i, s0);
        if (i + 1 < S.rows) store_in<IN>(yc + i + 1, s1);
      }
    }
  }
}

struct AddJobL {
  const float* M;
  float* C;
  long long ldm, ldc;
  int rows, cols;
  int mx, nx;
  int nlive;
  int vec;
  const int* cptr;
  const int* cprod;
  const double* cw;
};

__global__ void __launch_bounds__(kEwThreads) madd_lowp_kernel(AddJobL A) {
  const int b = (int)blockIdx.y;
  int r0, c0;
  ew_tile(A.rows, A.cols, &r0, &c0);
  if (c0 >= A.cols) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = r0 + 2 * lane;
  const int l0 = A.cptr[b], l1 = A.cptr[b + 1];
  gfloat* cb = (gfloat*)A.C + (long long)(b / A.nx) * A.rows + (long long)(b % A.nx) * A.cols * A.ldc;
  for (int c = c0 + wave; c < min(c0 + kEwCols, A.cols); c += kEwThreads / 64) {
    gfloat* cc = cb + (long long)c * A.ldc;
    const bool v2 = A.vec && i + 1 < A.rows;
    float c0v = 0.f, c1v = 0.f;
    if (v2) {
      const f32x2 v = *(gcf32x2*)(cc + i);
      c0v = v.x;
      c1v = v.y;
    } else {
      if (i < A.rows) c0v = cc[i];
      if (i + 1 < A.rows) c1v = cc[i + 1];
    }
    for (int l = l0; l < l1; ++l) {
      const int r = A.cprod[l];
      if (r >= A.nlive) continue;
      const float w = (float)A.cw[l];
      gcfloat* m = (gcfloat*)A.M + (long long)(r % A.mx) * A.rows + (long long)(r / A.mx) * A.cols * A.ldm +
                   (long long)c * A.ldm;
      float m0 = 0.f, m1 = 0.f;
      if (v2) {
        const f32x2 v = *(gcf32x2*)(m + i);
        m0 = v.x;
        m1 = v.y;
      } else {
        if (i < A.rows) m0 = m[i];
        if (i + 1 < A.rows) m1 = m[i + 1];
      }
      c0v = __builtin_fmaf(w, m0, c0v);
      c1v = __builtin_fmaf(w, m1, c1v);
    }
    if (v2) {
      *(__attribute__((address_space(1))) f32x2*)(cc + i) = f32x2{c0v, c1v};
    } else {
      if (i < A.rows) cc[i] = c0v;
      if (i + 1 < A.rows) cc[i + 1] = c1v;
    }
  }
}

// The configurations built: 8 waves of 64x64 per workgroup, C tiles of
// 256x128 (default) or 128x256 (jobs flagged kJobTileT).  Slice depth,
// prefetched term chunk and prefetch depth were chosen by timing 8192^3
// GEMMs and FMMs on the device among the variants that fit the 256 VGPRs of
// two waves per SIMD (a few spilled registers where that measured faster):
// deeper slices and two slices of loads in flight hide the L2 / HBM latency
// that a 32-deep BF16 slice (about 500 MFMA cycles per SIMD) cannot.  Operand
// sums with more terms than a chunk load the rest synchronously.  Split-BF16
// has the 256x128 shape only (its hi and lo planes leave no LDS for the
// deeper 128x256 slices).
// (slice depth BK, term chunk CH, prefetch depth PF) per configuration
#ifndef FMM_F32_CFG
#define FMM_F32_CFG 16, 2, 2
#endif
#ifndef FMM_F32T_CFG
#define FMM_F32T_CFG 16, 2, 1
#endif
#ifndef FMM_H16_CFG
#define FMM_H16_CFG 32, 2, 2
#endif
#ifndef FMM_H16T_CFG
#define FMM_H16T_CFG 32, 1, 1
#endif
#ifndef FMM_SPLIT_CFG
#define FMM_SPLIT_CFG 32, 1, 1
#endif
#ifndef FMM_F32_ONE_CFG
#define FMM_F32_ONE_CFG 32, 1, 2
#endif
#ifndef FMM_H16_ONE_CFG
#define FMM_H16_ONE_CFG 64, 1, 2
#endif
template <int IN, int OP, bool SPLIT, int BM, int BN, int WM, int WN, int BK, int CH, int PF>
using LC = LCfg<IN, OP, SPLIT, BM, BN, BK, WM, WN, CH, PF>;
template <int IN, int OP, int BM, int BN, int WM, int WN, int BK, int CH, int PF>
using LC1 = LCfg<IN, OP, false, BM, BN, BK, WM, WN, CH, PF, true>;
typedef LC<kInF32, kOpF32, false, 256, 128, 4, 2, FMM_F32_CFG> CfgF32;
typedef LC<kInF32, kOpF32, false, 128, 256, 2, 4, FMM_F32T_CFG> CfgF32T;
typedef LC<kInBF16, kOpBF16, false, 256, 128, 4, 2, FMM_H16_CFG> CfgBF16;
typedef LC<kInBF16, kOpBF16, false, 128, 256, 2, 4, FMM_H16T_CFG> CfgBF16T;
typedef LC<kInF16, kOpF16, false, 256, 128, 4, 2, FMM_H16_CFG> CfgF16;
typedef LC<kInF16, kOpF16, false, 128, 256, 2, 4, FMM_H16T_CFG> CfgF16T;
typedef LC<kInF32, kOpBF16, true, 256, 128, 4, 2, FMM_SPLIT_CFG> CfgBF16x3;
// single-term (+-1) launches: plain GEMMs and Naive product launches
// C tile and wave grid (BM, BN, WM, WN) of the single-term kernels
#ifndef FMM_F32_ONE_SHAPE
#define FMM_F32_ONE_SHAPE 256, 128, 4, 2
#endif
#ifndef FMM_H16_ONE_SHAPE
#define FMM_H16_ONE_SHAPE 256, 128, 4, 2
#endif
typedef LC1<kInF32, kOpF32, FMM_F32_ONE_SHAPE, FMM_F32_ONE_CFG> CfgF32One;
typedef LC1<kInBF16, kOpBF16, FMM_H16_ONE_SHAPE, FMM_H16_ONE_CFG> CfgBF16One;
typedef LC1<kInF16, kOpF16, FMM_H16_ONE_SHAPE, FMM_H16_ONE_CFG> CfgF16One;

}  // namespace fmm
