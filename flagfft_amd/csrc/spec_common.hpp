// spec_common.hpp — compile-time radix sequences shared by the specialised
// single-pass (leaf.hip) and super-stage (stage.hip) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"

namespace flagfft {
namespace dev {

template <int... Rs>
struct RadixSeq {
  static constexpr int kS = sizeof...(Rs);
  static constexpr int kR[kS] = {Rs...};
  static constexpr int kN = (Rs * ... * 1);
  static constexpr int prior(int s) {
    int p = 1;
    for (int i = 0; i < s; ++i) p *= kR[i];
    return p;
  }
  // Butterflies per thread of stage s when a transform is spread over
  // ceil(kN / E) threads with E elements each.
  template <int E>
  static constexpr int butterflies(int s) {
    return (E + kR[s] - 1) / kR[s];
  }
  // Offset of stage s's twiddles in a per-thread array holding every
  // inner-stage twiddle (stages 1..kS-1, (R - 1) per butterfly).
  template <int E>
  static constexpr int tw_offset(int s) {
    int o = 0;
    for (int i = 1; i < s; ++i) o += butterflies<E>(i) * (kR[i] - 1);
    return o;
  }
  template <int E>
  static constexpr int tw_total() {
    return tw_offset<E>(kS) > 0 ? tw_offset<E>(kS) : 1;
  }
  // Offset of stage s's base twiddles w^k (one per butterfly) when inner
  // twiddles are formed by binary powers.
  template <int E>
  static constexpr int pow_offset(int s) {
    int o = 0;
    for (int i = 1; i < s; ++i) o += butterflies<E>(i);
    return o;
  }
  template <int E>
  static constexpr int pow_total() {
    return pow_offset<E>(kS) > 0 ? pow_offset<E>(kS) : 1;
  }
  // Prefetch every inner-stage twiddle when they fit in 96 VGPRs; otherwise
  // each stage loads its own (FP64 with radix-16 stages).
  template <typename T, int E>
  static constexpr bool prefetch() {
    return tw_offset<E>(kS) * static_cast<int>(sizeof(T)) / 2 <= 96;
  }
};

// Every inner-stage twiddle w_N^(j k N/(q R)) a thread will need, loaded at
// kernel start so their latency overlaps the data loads instead of
// serialising one table round trip per stage.
template <typename T, int E, int TPU, typename RS>
__device__ __forceinline__ void prefetch_twiddles(V2<T> *tw_reg, const GlbC<T> *tw, int t) {
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    if constexpr (s > 0) {
      constexpr int RR = RS::kR[s];
      constexpr int q = RS::prior(s);
      constexpr int NB = RS::kN / RR;
      constexpr int MB = RS::template butterflies<E>(s);
      constexpr int off = RS::template tw_offset<E>(s);
      constexpr int ts = RS::kN / (q * RR);
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        const int k = (b < NB) ? b % q : 0;
#pragma unroll
        for (int j = 1; j < RR; ++j) tw_reg[off + m * (RR - 1) + (j - 1)] = ldg(tw + j * k * ts);
      }
    }
  });
}

constexpr int leaf_lds_stride(int n) {
  const int ext = (n - 1) + ((n - 1) >> 5) + 1;
  return (ext % 2 == 0) ? ext + 1 : ext;
}

// The base twiddle w_N^(k N/(q R)) of every inner-stage butterfly of a thread
// (binary-powering form: one value per butterfly), loaded at kernel start.
template <typename T, int E, int TPU, typename RS>
__device__ __forceinline__ void prefetch_base_twiddles(V2<T> *tw_reg, const GlbC<T> *tw, int t) {
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    if constexpr (s > 0) {
      constexpr int RR = RS::kR[s];
      constexpr int q = RS::prior(s);
      constexpr int NB = RS::kN / RR;
      constexpr int MB = RS::template butterflies<E>(s);
      constexpr int off = RS::template pow_offset<E>(s);
      constexpr int ts = RS::kN / (q * RR);
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        const int k = (b < NB) ? b % q : 0;
        tw_reg[off + m] = ldg(tw + k * ts);
      }
    }
  });
}

}  // namespace dev
}  // namespace flagfft
