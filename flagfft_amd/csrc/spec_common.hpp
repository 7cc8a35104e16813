// spec_common.hpp — compile-time radix sequences shared by the specialised
// single-pass (leaf.hip) and super-stage (stage.hip) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"

namespace flagfft {
namespace dev {

// Row kinds of the single-pass kernels: complex rows, real rows in (R2C:
// split-spectrum unpack fused), Hermitian rows in (C2R: pack fused).
enum LeafMode : int { kLeafComplex = 0, kLeafR2C = 1, kLeafC2R = 2 };

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

// w^j for a compile-time j from the binary powers p[b] = w^(2^b): at most
// popcount(j) - 1 products, so a few roundings instead of one table load per j.
template <typename T, int J>
__device__ __forceinline__ V2<T> tw_power(const V2<T> *p) {
  V2<T> r{1, 0};
  bool first = true;
  sfor<4>([&](auto B) __attribute__((always_inline)) {
    constexpr int b = decltype(B)::value;
    if constexpr ((J >> b) & 1) {
      r = first ? p[b] : cmul(r, p[b]);
      first = false;
    }
  });
  return r;
}

// x[j] *= w^j, j = 1..R-1, with w^j formed by binary powers of w.
template <typename T, int R>
__device__ __forceinline__ void twiddle_powers(V2<T> *x, V2<T> w) {
  V2<T> pw[4];
  pw[0] = w;
  pw[1] = cmul(pw[0], pw[0]);
  pw[2] = cmul(pw[1], pw[1]);
  pw[3] = cmul(pw[2], pw[2]);
  sfor<R>([&](auto J) __attribute__((always_inline)) {
    constexpr int jj = decltype(J)::value;
    if constexpr (jj > 0) x[jj] = cmul(x[jj], tw_power<T, jj>(pw));
  });
}

// Radix stages of one R-point transform spread over TPU lanes (lane t).
// LDS index e is padded by one slot every 2^PS elements.
// Stage 0 takes element e from load0(e); inner stages exchange through LDS
// (buf); the last stage hands natural-order element idx to sink(idx, v).
//   kLdsIn   load0 reads LDS too, so the first LDS write waits for every lane
//   sink_lds the sink writes LDS: wait for every lane's last-stage reads first
// Every lane of the workgroup must call this (it contains barriers).
// kTwPow: inner-stage twiddles w^(j k) by binary powers of the butterfly's
// base twiddle w^k, which twr holds (prefetch_base_twiddles), instead of R - 1
// table loads; for register-bound kernels.
// No extra work on the first stage's inputs.
struct NoPre0 {
  template <typename C>
  __device__ __forceinline__ void operator()(int, C *) const {}
};

// pre0(b, x) runs on the r inputs of first-stage butterfly b after load0
// (the inter-pass twiddles of a super-stage).
template <typename T, int E, int TPU, typename RS, bool kLdsIn, int PS = 5, bool kTwPow = false,
          bool kPreBase = false, typename Load, typename Sink, typename Pre0 = NoPre0>
__device__ __forceinline__ void run_stages(LdsC<T> *buf, int t, const V2<T> *twr, const GlbC<T> *tws, bool sink_lds,
                                           Load load0, Sink sink, Pre0 pre0 = Pre0{}) {
  auto lpad = [](int e) __attribute__((always_inline)) { return e + (e >> PS); };
  constexpr int R = RS::kN;
  constexpr bool kPre = !kTwPow && RS::template prefetch<T, E>();
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    constexpr int RR = RS::kR[s];
    constexpr int q = RS::prior(s);
    constexpr int NB = R / RR;
    constexpr int MB = RS::template butterflies<E>(s);
    V2<T> x[MB * RR];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int b = t + m * TPU;
      if (b >= NB) continue;
#pragma unroll
      for (int j = 0; j < RR; ++j) {
        const int e = b + j * NB;
        if constexpr (s == 0)
          x[m * RR + j] = load0(e);
        else
          x[m * RR + j] = ldc(buf + lpad(e));
      }
      if constexpr (s == 0) pre0(b, &x[m * RR]);
      if constexpr (q > 1) {
        if constexpr (kPre) {
          constexpr int off = RS::template tw_offset<E>(s);
#pragma unroll
          for (int j = 1; j < RR; ++j) x[m * RR + j] = cmul(x[m * RR + j], twr[off + m * (RR - 1) + (j - 1)]);
        } else if constexpr (kTwPow) {
          V2<T> pw[4];
          if constexpr (kPreBase) {
            pw[0] = twr[RS::template pow_offset<E>(s) + m];  // prefetched base twiddle w^k
          } else {
            constexpr int ts = R / (q * RR);
            pw[0] = ldt(tws, static_cast<uint32_t>(b % q) * ts);
          }
          pw[1] = cmul(pw[0], pw[0]);
          pw[2] = cmul(pw[1], pw[1]);
          pw[3] = cmul(pw[2], pw[2]);
          sfor<RR>([&](auto J) __attribute__((always_inline)) {
            constexpr int j = decltype(J)::value;
            if constexpr (j > 0) x[m * RR + j] = cmul(x[m * RR + j], tw_power<T, j>(pw));
          });
        } else {
          constexpr int ts = R / (q * RR);
          const uint32_t kk = static_cast<uint32_t>(b % q);
#pragma unroll
          for (int j = 1; j < RR; ++j) x[m * RR + j] = cmul(x[m * RR + j], ldt(tws, j * kk * ts));
        }
      }
      dft<RR>(&x[m * RR]);
    }
    if constexpr (s + 1 < RS::kS) {
      if constexpr (s > 0 || kLdsIn) __syncthreads();
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        if (b >= NB) continue;
        const int kk = (q > 1) ? b % q : 0;
        const int base = (b - kk) * RR + kk;
#pragma unroll
        for (int j = 0; j < RR; ++j) stc(buf + lpad(base + j * q), x[m * RR + j]);
      }
      __syncthreads();
    } else {
      if constexpr (RS::kS > 1 || kLdsIn) {
        if (sink_lds) __syncthreads();
      }
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        if (b >= NB) continue;
#pragma unroll
        for (int j = 0; j < RR; ++j) sink(b + j * q, x[m * RR + j]);
      }
    }
  });
}

}  // namespace dev
}  // namespace flagfft
