// fmm_refine.hpp -- FP64 C += A*B through low-precision products and one
// residual correction.
//
// With A1 = fl_T(A), B1 = fl_T(B) (T the main product's input type) and the
// FP32 remainders dA = fl32(A - A1), dB = fl32(B - B1):
//     A*B = A1*B1 + (dA*B1 + A1*dB) + dA*dB
// The main product C~ = A1*B1 runs on the fast path (an FMM or GEMM of type
// T, FP32 result); the correction E = dA*B1 + A1*dB -- the residual A*B - C~
// up to the dropped dA*dB and the main product's own arithmetic error -- is
// one product with depth 2k, [dA | A1] * [B1 ; dB] (FP32 inputs with FP32 or
// split-BF16 arithmetic, or -- when the main product is BF16 -- BF16 inputs:
// E is 2^-8 of C, so BF16 rounding of dA, dB costs ~2^-17 of C); then
// C += C~ + E in FP64.  FP64 appears only in the
// split (A - A1 is formed in FP64) and in this final update.
#pragma once
#include "fmm_host.hpp"

namespace fmm {

constexpr int kRefThreads = 256;

template <int IN>
__device__ __forceinline__ void round_in(double x, typename InT<IN>::T* out, float* as_f32) {
  if constexpr (IN == kInF32) {
    const float f = (float)x;
    *out = f;
    *as_f32 = f;
  } else if constexpr (IN == kInBF16) {
    const __bf16 h = (__bf16)(float)x;
    *out = __builtin_bit_cast(u16, h);
    *as_f32 = (float)h;
  } else {
    const _Float16 h = (_Float16)(float)x;
    *out = __builtin_bit_cast(u16, h);
    *as_f32 = (float)h;
  }
}

// A (m x k, FP64) -> A1 (m x k, type IN, ld la1) and AA = [dA | A1] (m x 2k,
// type CI = the correction's input type, FP32 or BF16; ld laa).  A1 is
// exact in CI (CI is BF16 only when IN is).  One thread per element.
template <int IN, int CI>
__global__ void __launch_bounds__(kRefThreads) split_a_kernel(int m, int k, const double* A, long long lda,
                                                              typename InT<IN>::T* A1, long long la1,
                                                              typename InT<CI>::T* AA, long long laa) {
  const long long total = (long long)m * k;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % m), p = (int)(t / m);
    const double a = A[i + p * lda];
    float a1f, unused;
    round_in<IN>(a, A1 + i + p * la1, &a1f);
    round_in<CI>(a - (double)a1f, AA + i + p * laa, &unused);
    round_in<CI>((double)a1f, AA + i + (p + (long long)k) * laa, &unused);
  }
}

// B (k x n, FP64) -> B1 (k x n, type IN, ld lb1) and BB = [B1 ; dB] (2k x n,
// type CI, ld lbb).
template <int IN, int CI>
__global__ void __launch_bounds__(kRefThreads) split_b_kernel(int k, int n, const double* B, long long ldb,
                                                              typename InT<IN>::T* B1, long long lb1,
                                                              typename InT<CI>::T* BB, long long lbb) {
  const long long total = (long long)k * n;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(t % k), j = (int)(t / k);
    const double b = B[p + j * ldb];
    float b1f, unused;
    round_in<IN>(b, B1 + p + j * lb1, &b1f);
    round_in<CI>((double)b1f, BB + p + j * lbb, &unused);
    round_in<CI>(b - (double)b1f, BB + p + (long long)k + j * lbb, &unused);
  }
}

__global__ void __launch_bounds__(kRefThreads) zero_f32_kernel(float* X, long long n) {
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    X[t] = 0.f;
}

// C += Ct + Et in FP64 (Ct, Et: m x n FP32 with leading dimension lt).
__global__ void __launch_bounds__(kRefThreads) combine_kernel(int m, int n, const float* Ct, const float* Et,
                                                              long long lt, double* C, long long ldc) {
  const long long total = (long long)m * n;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < total;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % m), j = (int)(t / m);
    C[i + j * ldc] = C[i + j * ldc] + ((double)Ct[i + j * lt] + (double)Et[i + j * lt]);
  }
}

inline int ew_blocks(long long n) { return (int)std::max(1LL, std::min((n + kRefThreads - 1) / kRefThreads, 16384LL)); }

inline long long round8(long long x) { return (x + 7) & ~7LL; }

// Workspace layout of one refined product.
struct RefLayout {
  long long la1, lb1, laa, lbb, lt;
  size_t hdr, a1, b1, aa, bb, ct, et, wsm, wsc, bytes;
};

inline int in_size(int dtype) { return (dtype == kBF16 || dtype == kF16) ? 2 : 4; }

inline RefLayout ref_layout(int main_dtype, int corr_dtype, int m, int n, int k, size_t ws_main, size_t ws_corr) {
  RefLayout L{};
  const int s = in_size(main_dtype);
  const int sc = in_size(corr_dtype);
  L.la1 = round8(std::max(m, 1));
  L.lb1 = round8(std::max(k, 1));
  L.laa = round8(std::max(m, 1));
  L.lbb = round8(std::max(2 * k, 1));
  L.lt = round8(std::max(m, 1));
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += round256(std::max<size_t>(bytes, 1));
    return o;
  };
  // status header at offset 0: its error word (where an ABC workspace has
  // it) collects the flags of both sub-calls
  L.hdr = take(sizeof(int) * (size_t)kWsCounters);
  L.a1 = take((size_t)L.la1 * k * s);
  L.b1 = take((size_t)L.lb1 * n * s);
  L.aa = take((size_t)L.laa * 2 * k * sc);
  L.bb = take((size_t)L.lbb * n * sc);
  L.ct = take((size_t)L.lt * n * 4);
  L.et = take((size_t)L.lt * n * 4);
  L.wsm = take(ws_main);
  L.wsc = take(ws_corr);
  L.bytes = off;
  return L;
}

}  // namespace fmm
