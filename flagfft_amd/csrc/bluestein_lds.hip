// bluestein_lds.hip — Bluestein with the whole length-M convolution in the
// LDS of one workgroup (one launch per batch of sequences, FP32), for
// lengths whose M exceeds the single-pass limit.
#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"
#include "fft_args.hpp"
#include "kernels.hpp"
#include "spec_common.hpp"

namespace flagfft {
namespace dev {

// Bluestein with the whole length-M convolution in LDS (one launch, one
// sequence per workgroup), for M above the default single-pass limit:
//   a = chirp * x (zero past n), A = FFT_M(a) -> LDS, W = FFT_M(conj(A * bhat)),
//   y[k] = conj(W[k]) * chirp[k], k < n           (src/exec/raw_nodes.cpp:153-204)
// One pad slot every 64 elements so M = 20160 fits the 160 KiB of LDS; up to
// 1024 lanes per workgroup keep the per-lane element count (and registers) low.
template <typename T, int E, bool kPreBase, int... Rs>
__global__ __launch_bounds__(1024) void k_bluestein_lds(const BluesteinLdsArgs a) {
  using RS = RadixSeq<Rs...>;
  constexpr int M = RS::kN;
  constexpr int TPU = (M + E - 1) / E;
  constexpr int PS = 6;
  LdsC<T> *buf = lds_base<T>();
  const int t = threadIdx.x;
  const uint32_t q = blockIdx.x;
  const uint32_t b = fdiv(q, a.fd_ncol);
  const uint32_t c = q - b * static_cast<uint32_t>(a.ncol);
  const int64_t ibase = static_cast<int64_t>(b) * a.gi.bs + static_cast<int64_t>(c) * a.gi.cs;
  const int64_t obase = static_cast<int64_t>(b) * a.go.bs + static_cast<int64_t>(c) * a.go.cs;
  const GlbC<T> *tws = gc<T>(a.tw);
  const GlbC<T> *bhat = gc<T>(a.bhat);
  const T scale = static_cast<T>(a.scale);
  // kPreBase: one base twiddle per inner-stage butterfly loaded up front and
  // shared by both transforms (more registers), else one load per stage
  V2<T> twr[kPreBase ? RS::template pow_total<E>() : 1];
  if constexpr (kPreBase) prefetch_base_twiddles<T, E, TPU, RS>(twr, tws, t);
  auto pad = [](int e) __attribute__((always_inline)) { return e + (e >> PS); };
  auto to_lds = [&](int idx, V2<T> v) __attribute__((always_inline)) { stc(buf + pad(idx), v); };
  run_stages<T, E, TPU, RS, false, PS, true, kPreBase>(
      buf, t, twr, tws, true,
      [&](int e) __attribute__((always_inline)) {
        return src_elem<T>(a.in, ibase, a.gi.es, a.src_mode | kSrcChirpFlag, a.n, a.chirp, a.conj_in, e);
      },
      to_lds);
  __syncthreads();
  run_stages<T, E, TPU, RS, true, PS, true, kPreBase>(
      buf, t, twr, tws, false,
      [&](int e) __attribute__((always_inline)) { return cconj(cmul(ldc(buf + pad(e)), ldg(bhat + e))); },
      [&](int idx, V2<T> v) __attribute__((always_inline)) {
        dst_elem<T>(a.out, obase, a.go.es, a.dst_mode | kDstChirpFlag, a.n, a.chirp, nullptr, a.conj_out, scale,
                    idx, v);
      });
}

// In-place form of the LDS-resident Bluestein convolution: each stage's
// butterflies read and write back their own r slots, so a stage needs one
// barrier and a lane holds one butterfly at a time.
//   forward  decimation in frequency: stage s (span L = ML / (r_0..r_{s-1}),
//            m = L / r) transforms slots blk L + k + j m (j < r) and scales
//            output jj by w_L^(k jj); the spectrum ends in digit-reversed
//            order (slot P(k) = sum_s digit_s(k) m_s, digits least
//            significant first);
//   middle   the last forward stage (m = 1), the product with bhat stored in
//            the same digit-reversed order, the conjugation and the first
//            stage of the second transform share one butterfly;
//   second   decimation in time, stages in reverse, input j scaled by
//            w_L^(k j) before the butterfly, which takes digit-reversed input
//            to natural order.
namespace ip {
template <typename RS, int s>
struct Stage {
  static constexpr int r = RS::kR[s];
  static constexpr int L = RS::kN / RS::prior(s);
  static constexpr int m = L / r;
  static constexpr int NB = RS::kN / r;
};
template <typename RS, int TPU>
constexpr int per_lane(int s) {
  return (RS::kN / RS::kR[s] + TPU - 1) / TPU;
}
}  // namespace ip

// x[j] *= w^j (j = 1..R-1) for the butterfly's base twiddle w = tws[kp]:
// kTwLoad reads each w^j = tws[j kp] from the table, else forms it by binary
// powers of w (fewer loads, more VALU work).
template <typename T, int R, bool kTwLoad>
__device__ __forceinline__ void twiddle_run(V2<T> *x, const GlbC<T> *tws, uint32_t kp) {
  if constexpr (kTwLoad) {
#pragma unroll
    for (int j = 1; j < R; ++j) x[j] = cmul(x[j], ldt(tws, static_cast<uint32_t>(j) * kp));
  } else {
    V2<T> pw[4];
    pw[0] = ldt(tws, kp);
    pw[1] = cmul(pw[0], pw[0]);
    pw[2] = cmul(pw[1], pw[1]);
    pw[3] = cmul(pw[2], pw[2]);
    sfor<R>([&](auto J) __attribute__((always_inline)) {
      constexpr int j = decltype(J)::value;
      if constexpr (j > 0) x[j] = cmul(x[j], tw_power<T, j>(pw));
    });
  }
}

// One cyclic convolution of ML = RS::kN points through LDS (in place):
// forward DIF transform of load(e), product with bhat (digit-reversed order),
// conjugation, DIT transform; store(e, v) receives output e in natural order.
// Only inputs j < J0 of each first-stage butterfly can be nonzero and only its
// outputs j < J1 are stored (compile-time pruning of zero padding / unused
// outputs).  Twiddles come from tws = w_{ML TS}^i (every TS-th entry used).
// Every lane of the workgroup must call this (it contains barriers); the
// caller separates two calls on the same LDS with a barrier.
template <typename T, int TPU, bool kTwLoad, int TS, int J0, int J1, typename RS, typename Load, typename Store>
__device__ __forceinline__ void conv_lds(LdsC<T> *buf, int t, const GlbC<T> *tws, const GlbC<T> *bhat, Load load,
                                         Store store) {
  constexpr int S = RS::kS;
  static_assert(S >= 2, "at least two radix stages");
  constexpr int PS = 6;
  auto pad = [](int e) __attribute__((always_inline)) { return e + (e >> PS); };
  // first transform, stage 0 (input from load)
  {
    using St = ip::Stage<RS, 0>;
#pragma unroll 1
    for (int i = 0; i < ip::per_lane<RS, TPU>(0); ++i) {
      const int k = t + i * TPU;
      if (k >= St::NB) break;
      V2<T> x[St::r];
#pragma unroll
      for (int j = 0; j < St::r; ++j) x[j] = (j < J0) ? load(k + j * St::m) : V2<T>{0, 0};
      dft<St::r>(x);
      twiddle_run<T, St::r, kTwLoad>(x, tws, static_cast<uint32_t>(k * TS));
#pragma unroll
      for (int j = 0; j < St::r; ++j) stc(buf + pad(k + j * St::m), x[j]);
    }
  }
  __syncthreads();
  // first transform, stages 1..S-2
  sfor<S - 2>([&](auto SS) __attribute__((always_inline)) {
    constexpr int s = decltype(SS)::value + 1;
    using St = ip::Stage<RS, s>;
#pragma unroll
    for (int i = 0; i < ip::per_lane<RS, TPU>(s); ++i) {
      const int beta = t + i * TPU;
      if (beta >= St::NB) continue;
      const int blk = beta / St::m;
      const int k = beta - blk * St::m;
      const int base = blk * St::L + k;
      V2<T> x[St::r];
#pragma unroll
      for (int j = 0; j < St::r; ++j) x[j] = ldc(buf + pad(base + j * St::m));
      dft<St::r>(x);
      twiddle_run<T, St::r, kTwLoad>(x, tws, static_cast<uint32_t>(k * RS::prior(s) * TS));
#pragma unroll
      for (int j = 0; j < St::r; ++j) stc(buf + pad(base + j * St::m), x[j]);
    }
    __syncthreads();
  });
  // last stage of the first transform, product with bhat, conjugate, first
  // stage of the second transform (no twiddles: m = 1)
  {
    using St = ip::Stage<RS, S - 1>;
    static_assert(St::m == 1, "last stage spans one butterfly");
#pragma unroll
    for (int i = 0; i < ip::per_lane<RS, TPU>(S - 1); ++i) {
      const int beta = t + i * TPU;
      if (beta >= St::NB) continue;
      const int base = beta * St::r;
      V2<T> x[St::r];
#pragma unroll
      for (int j = 0; j < St::r; ++j) x[j] = ldc(buf + pad(base + j));
      dft<St::r>(x);
#pragma unroll
      for (int j = 0; j < St::r; ++j) x[j] = cconj(cmul(x[j], ldt(bhat, static_cast<uint32_t>(base + j))));
      dft<St::r>(x);
#pragma unroll
      for (int j = 0; j < St::r; ++j) stc(buf + pad(base + j), x[j]);
    }
  }
  __syncthreads();
  // second transform, stages S-2..1
  sfor<S - 2>([&](auto SS) __attribute__((always_inline)) {
    constexpr int s = S - 2 - decltype(SS)::value;
    using St = ip::Stage<RS, s>;
#pragma unroll
    for (int i = 0; i < ip::per_lane<RS, TPU>(s); ++i) {
      const int beta = t + i * TPU;
      if (beta >= St::NB) continue;
      const int blk = beta / St::m;
      const int k = beta - blk * St::m;
      const int base = blk * St::L + k;
      V2<T> x[St::r];
#pragma unroll
      for (int j = 0; j < St::r; ++j) x[j] = ldc(buf + pad(base + j * St::m));
      twiddle_run<T, St::r, kTwLoad>(x, tws, static_cast<uint32_t>(k * RS::prior(s) * TS));
      dft<St::r>(x);
#pragma unroll
      for (int j = 0; j < St::r; ++j) stc(buf + pad(base + j * St::m), x[j]);
    }
    __syncthreads();
  });
  // second transform, stage 0 (natural order, to store)
  {
    using St = ip::Stage<RS, 0>;
#pragma unroll 1
    for (int i = 0; i < ip::per_lane<RS, TPU>(0); ++i) {
      const int k = t + i * TPU;
      if (k >= St::NB) break;
      V2<T> x[St::r];
#pragma unroll
      for (int j = 0; j < St::r; ++j) x[j] = ldc(buf + pad(k + j * St::m));
      twiddle_run<T, St::r, kTwLoad>(x, tws, static_cast<uint32_t>(k * TS));
      dft<St::r>(x);
#pragma unroll
      for (int j = 0; j < J1; ++j) store(k + j * St::m, x[j]);
    }
  }
}

// Sequence accessors of the LDS Bluestein kernels.  kFast: contiguous
// complex input and output (element offsets e < n fit 32 bits from a uniform
// sequence base); otherwise the general source / sink forms.
template <typename T, bool kFast>
struct BlueIo {
  const BluesteinLdsArgs &a;
  int64_t ibase, obase;
  const GlbC<T> *chirp;
  T scale;
  int n;
  __device__ __forceinline__ BlueIo(const BluesteinLdsArgs &args, uint32_t q) : a(args) {
    const uint32_t b = fdiv(q, a.fd_ncol);
    const uint32_t c = q - b * static_cast<uint32_t>(a.ncol);
    ibase = static_cast<int64_t>(b) * a.gi.bs + static_cast<int64_t>(c) * a.gi.cs;
    obase = static_cast<int64_t>(b) * a.go.bs + static_cast<int64_t>(c) * a.go.cs;
    chirp = gc<T>(a.chirp);
    scale = static_cast<T>(a.scale);
    n = static_cast<int>(a.n);
  }
  // a[e] = x[e] chirp[e], e < n
  __device__ __forceinline__ V2<T> in(int e) const {
    if constexpr (kFast) {
      V2<T> v = ldt(gc<T>(a.in) + ibase, static_cast<uint32_t>(e));
      if (a.conj_in) v.y = -v.y;
      return cmul(v, ldt(chirp, static_cast<uint32_t>(e)));
    } else {
      return src_elem<T>(a.in, ibase, a.gi.es, a.src_mode | kSrcChirpFlag, a.n, a.chirp, a.conj_in, e);
    }
  }
  // y[e] = conj(W[e]) chirp[e] (then conjugated / scaled as asked), e < n
  __device__ __forceinline__ void out(int e, V2<T> v) const {
    if constexpr (kFast) {
      v = cmul(cconj(v), ldt(chirp, static_cast<uint32_t>(e)));
      if (a.conj_out) v.y = -v.y;
      v.x *= scale;
      v.y *= scale;
      GlbC<T> *o = gc<T>(a.out) + obase;
      stg(reinterpret_cast<GlbC<T> *>(reinterpret_cast<FLAGFFT_GLB char *>(o) +
                                      static_cast<uint32_t>(e) * static_cast<uint32_t>(sizeof(V2<T>))),
          v);
    } else {
      dst_elem<T>(a.out, obase, a.go.es, a.dst_mode | kDstChirpFlag, a.n, a.chirp, nullptr, a.conj_out, scale, e, v);
    }
  }
};

// Whole convolution (M = RS::kN) in LDS, one workgroup per sequence.
// Requires M even and n <= M / 2 (Bluestein's M >= 2n - 1): with an even
// first radix r_0 the inputs j >= r_0 / 2 of every first-stage butterfly are
// zero padding and its outputs j >= r_0 / 2 are past n.
template <typename T, int TPU, bool kTwLoad, bool kFast, typename RS>
__device__ __forceinline__ void bluestein_ip_body(const BluesteinLdsArgs &a) {
  constexpr int R0 = RS::kR[0];
  constexpr int J0 = (R0 % 2 == 0) ? R0 / 2 : R0;
  const BlueIo<T, kFast> io(a, blockIdx.x);
  const int n = io.n;
  conv_lds<T, TPU, kTwLoad, 1, J0, J0, RS>(
      lds_base<T>(), threadIdx.x, gc<T>(a.tw), gc<T>(a.bhat),
      [&](int e) __attribute__((always_inline)) { return e < n ? io.in(e) : V2<T>{0, 0}; },
      [&](int e, V2<T> v) __attribute__((always_inline)) {
        if (e < n) io.out(e, v);
      });
}

template <typename T, int TPU, bool kTwLoad, int... Rs>
__global__ __launch_bounds__(TPU) void k_bluestein_ip(const BluesteinLdsArgs a) {
  using RS = RadixSeq<Rs...>;
  static_assert(RS::kN % 2 == 0, "even convolution length");
  if (a.gi.es == 1 && a.go.es == 1 && a.src_mode == kSrcComplex && a.dst_mode == kDstComplex)
    bluestein_ip_body<T, TPU, kTwLoad, true, RS>(a);
  else
    bluestein_ip_body<T, TPU, kTwLoad, false, RS>(a);
}

// Convolution of M = 2 ML points (ML = RS::kN in LDS) as two halves.  The
// radix-2 first stage of a DIF transform of a (a[e] = 0 for e >= n, n <= ML)
// gives the independent blocks a[e] and a[e] w_M^e (e < ML); each is
// convolved in LDS against its half of bhat (digit-reversed over [2, RS...]),
// and the radix-2 last stage of the DIT transform combines them:
// W[e] = Z0[e] + w_M^e Z1[e], e < n.  Z0 waits in the park workspace (n
// elements per sequence, written and read back by the same lane).
// tw = w_M^i, i < M.
template <typename T, int TPU, bool kFast, typename RS>
__device__ __forceinline__ void bluestein_halves_body(const BluesteinLdsArgs &a) {
  constexpr int ML = RS::kN;
  constexpr int R0 = RS::kR[0];
  const BlueIo<T, kFast> io(a, blockIdx.x);
  const int n = io.n;
  const GlbC<T> *tws = gc<T>(a.tw);
  const GlbC<T> *bhat = gc<T>(a.bhat);
  GlbC<T> *park = gc<T>(a.park) + static_cast<int64_t>(blockIdx.x) * n;
  LdsC<T> *buf = lds_base<T>();
  const int t = threadIdx.x;
  conv_lds<T, TPU, false, 2, R0, R0, RS>(
      buf, t, tws, bhat, [&](int e) __attribute__((always_inline)) { return e < n ? io.in(e) : V2<T>{0, 0}; },
      [&](int e, V2<T> v) __attribute__((always_inline)) {
        if (e < n) stg(park + e, v);
      });
  __syncthreads();
  conv_lds<T, TPU, false, 2, R0, R0, RS>(
      buf, t, tws, bhat + ML,
      [&](int e) __attribute__((always_inline)) {
        return e < n ? cmul(io.in(e), ldt(tws, static_cast<uint32_t>(e))) : V2<T>{0, 0};
      },
      [&](int e, V2<T> v) __attribute__((always_inline)) {
        if (e < n) io.out(e, cadd(ldg(park + e), cmul(v, ldt(tws, static_cast<uint32_t>(e)))));
      });
}

template <typename T, int TPU, int... Rs>
__global__ __launch_bounds__(TPU) void k_bluestein_halves(const BluesteinLdsArgs a) {
  using RS = RadixSeq<Rs...>;
  if (a.gi.es == 1 && a.go.es == 1 && a.src_mode == kSrcComplex && a.dst_mode == kDstComplex)
    bluestein_halves_body<T, TPU, true, RS>(a);
  else
    bluestein_halves_body<T, TPU, false, RS>(a);
}

// One block of a convolution split over P workgroups (M = P ML, ML = RS::kN
// in LDS).  The radix-P first stage of a DIF transform of a (a[e] = 0 for
// e >= n, n <= M / 2) gives P independent blocks
//   b_h[e] = sum_{j < P/2} a[e + j ML] w_M^(h (e + j ML)),  e < ML,
// workgroup (q, h) convolves b_h against block h of bhat (digit-reversed over
// [P, RS...]) and parks Z_h[e], e < min(ML, n); k_bluestein_combine applies
// the radix-P last stage of the DIT transform.  tw = w_M^i, i < M.
template <typename T, int TPU, int P, bool kFast, typename RS>
__device__ __forceinline__ void bluestein_split_body(const BluesteinLdsArgs &a) {
  constexpr int ML = RS::kN;
  constexpr int M = ML * P;
  static_assert(P == 2 || P == 4, "split into 2 or 4 blocks");
  // workgroups go round-robin over the 8 XCDs: keep the P blocks of a
  // sequence on one XCD (its L2 serves their repeated input reads) when the
  // sequence count allows
  uint32_t q, slot;
  const uint32_t nseq = gridDim.x / P;
  if ((nseq & 7u) == 0) {
    const uint32_t xcd = blockIdx.x & 7u, local = blockIdx.x >> 3;
    q = (local / P) * 8u + xcd;
    slot = q * P + local % P;
  } else {
    q = blockIdx.x / P;
    slot = blockIdx.x;
  }
  const int h = static_cast<int>(slot % P);
  const BlueIo<T, kFast> io(a, q);
  const int n = io.n;
  const GlbC<T> *tws = gc<T>(a.tw);
  GlbC<T> *park = gc<T>(a.park) + static_cast<int64_t>(slot) * ML;
  const GlbC<T> *cws = gc<T>(a.chirp_split) + static_cast<int64_t>(h) * n;
  conv_lds<T, TPU, false, P, RS::kR[0], RS::kR[0], RS>(
      lds_base<T>(), threadIdx.x, tws, gc<T>(a.bhat) + h * ML,
      [&](int e) __attribute__((always_inline)) {
        V2<T> v{0, 0};
#pragma unroll
        for (int j = 0; j < P / 2; ++j) {
          const int e2 = e + j * ML;
          if (e2 < n) {
            if constexpr (kFast) {
              // x[e2] chirp[e2] w_M^(h e2) from one FP64-generated table
              V2<T> xv = ldt(gc<T>(a.in) + io.ibase, static_cast<uint32_t>(e2));
              if (a.conj_in) xv.y = -xv.y;
              v = cadd(v, cmul(xv, ldt(cws, static_cast<uint32_t>(e2))));
            } else {
              v = cadd(v, cmul(io.in(e2), ldt(tws, static_cast<uint32_t>((h * e2) % M))));
            }
          }
        }
        return v;
      },
      [&](int e, V2<T> v) __attribute__((always_inline)) {
        if (e < n) stg(park + e, v);
      });
}

template <typename T, int TPU, int P, int... Rs>
__global__ __launch_bounds__(TPU) void k_bluestein_split(const BluesteinLdsArgs a) {
  using RS = RadixSeq<Rs...>;
  if (a.gi.es == 1 && a.src_mode == kSrcComplex)
    bluestein_split_body<T, TPU, P, true, RS>(a);
  else
    bluestein_split_body<T, TPU, P, false, RS>(a);
}

// W[e] = sum_h Z_h[e mod ML] w_M^(h e mod M), y[e] = conj(W[e]) chirp[e]
// (then conjugated / scaled and stored in the destination layout), e < n.
template <typename T, int P>
__global__ __launch_bounds__(256) void k_bluestein_combine(const BluesteinCombineArgs a) {
  // same XCD as the split kernel's blocks of sequence q (see there)
  uint32_t q, j;
  const uint32_t nseq = gridDim.x / static_cast<uint32_t>(a.nb);
  if ((nseq & 7u) == 0) {
    const uint32_t local = blockIdx.x >> 3;
    const uint32_t ql = fdiv(local, a.fd_nb);
    q = ql * 8u + (blockIdx.x & 7u);
    j = local - ql * static_cast<uint32_t>(a.nb);
  } else {
    q = fdiv(blockIdx.x, a.fd_nb);
    j = blockIdx.x - q * static_cast<uint32_t>(a.nb);
  }
  const int e = static_cast<int>(j) * 256 + static_cast<int>(threadIdx.x);
  if (e >= a.n) return;
  const int ml = a.ml;
  const int m = P * ml;
  const int k = e >= ml ? e - ml : e;  // e < n <= M / 2 <= 2 ml
  const GlbC<T> *z = gc<T>(a.in) + static_cast<int64_t>(q) * P * ml;
  const GlbC<T> *tws = gc<T>(a.tw);
  V2<T> w = ldg(z + k);
#pragma unroll
  for (int h = 1; h < P; ++h) {
    int i = h * e;  // < P M / 2
    while (i >= m) i -= m;
    w = cadd(w, cmul(ldg(z + h * ml + k), ldt(tws, static_cast<uint32_t>(i))));
  }
  const uint32_t b = fdiv(q, a.fd_ncol);
  const uint32_t c = q - b * static_cast<uint32_t>(a.ncol);
  const int64_t obase = static_cast<int64_t>(b) * a.go.bs + static_cast<int64_t>(c) * a.go.cs;
  dst_elem<T>(a.out, obase, a.go.es, a.dst_mode | kDstChirpFlag, a.n, a.chirp, nullptr, a.conj_out,
              static_cast<T>(a.scale), e, w);
}

}  // namespace dev

namespace {

struct BluesteinLdsEntry {
  bool dbl, halves;  // halves: the LDS holds M / 2, both halves in one workgroup
  int split;         // > 0: one block of M / split per workgroup (k_bluestein_split)
  int nrad;
  int rad[4];  // radices of the LDS part
  int tpu;     // lanes of the in-place kernels
  int e32;     // elements per lane of the Stockham kernels (0: none)
  // variants: 0 Stockham stages, 1 Stockham with base twiddles prefetched at
  // kernel start, 2 in-place stages with table twiddles, 3 in-place stages
  // with binary-powered twiddles; entries without Stockham kernels hold
  // their one kernel in every slot
  const void *f[4];
};

#define FLAGFFT_BLDS_F32(E, R0, R1, R2, R3)                                                           \
  {                                                                                                \
    false, false, 0, 4, {R0, R1, R2, R3}, 1024, E, {                                               \
      (const void *)dev::k_bluestein_lds<float, E, false, R0, R1, R2, R3>,                         \
          (const void *)dev::k_bluestein_lds<float, E, true, R0, R1, R2, R3>,                      \
          (const void *)dev::k_bluestein_ip<float, 1024, true, R0, R1, R2, R3>,                    \
          (const void *)dev::k_bluestein_ip<float, 1024, false, R0, R1, R2, R3>                    \
    }                                                                                              \
  }
#define FLAGFFT_BLDS_ONE(DBL, HALVES, SPLIT, TPU, KERNEL, NR, ...)                                   \
  {                                                                                                \
    DBL, HALVES, SPLIT, NR, {__VA_ARGS__}, TPU, 0, {                                               \
      (const void *)KERNEL, (const void *)KERNEL, (const void *)KERNEL, (const void *)KERNEL       \
    }                                                                                              \
  }
#define FLAGFFT_BLDS_IP64(TPU, R0, R1, R2, R3) \
  FLAGFFT_BLDS_ONE(true, false, 0, TPU, (dev::k_bluestein_ip<double, TPU, false, R0, R1, R2, R3>), 4, R0, R1, R2, R3)
#define FLAGFFT_BLDS_HALVES(T, DBL, TPU, R0, R1, R2, R3) \
  FLAGFFT_BLDS_ONE(DBL, true, 0, TPU, (dev::k_bluestein_halves<T, TPU, R0, R1, R2, R3>), 4, R0, R1, R2, R3)
#define FLAGFFT_BLDS_SPLIT4(T, DBL, P, TPU, R0, R1, R2, R3) \
  FLAGFFT_BLDS_ONE(DBL, false, P, TPU, (dev::k_bluestein_split<T, TPU, P, R0, R1, R2, R3>), 4, R0, R1, R2, R3)
#define FLAGFFT_BLDS_SPLIT3(T, DBL, P, TPU, R0, R1, R2) \
  FLAGFFT_BLDS_ONE(DBL, false, P, TPU, (dev::k_bluestein_split<T, TPU, P, R0, R1, R2>), 3, R0, R1, R2, 0)

// M in LDS: 16384 / 20160 complex FP32 and 8192 / 10080 complex FP64 (one
// pad slot per 64 elements); the halves kernels reach twice those M, the
// split kernels spread M over 2 or 4 workgroups.  The last radix of the FP32
// orders is odd: the fused middle stage walks LDS with a stride of that
// radix, and an odd stride keeps the 8-byte accesses free of bank conflicts
// (M = 20160: 14x12x12x10 -> 16x12x15x7 ran 17% faster).
const BluesteinLdsEntry kBluesteinLdsTable[] = {
    FLAGFFT_BLDS_F32(16, 16, 16, 16, 4),                     // M = 16384
    FLAGFFT_BLDS_F32(20, 16, 12, 15, 7),                     // M = 20160
    FLAGFFT_BLDS_HALVES(float, false, 1024, 8, 16, 16, 8),   // M = 32768
    FLAGFFT_BLDS_HALVES(float, false, 1024, 10, 12, 12, 14), // M = 40320 (16x12x15x7 spills here)
    FLAGFFT_BLDS_IP64(512, 16, 16, 8, 4),                    // M = 8192
    FLAGFFT_BLDS_IP64(1024, 14, 12, 10, 6),                  // M = 10080
    FLAGFFT_BLDS_HALVES(double, true, 512, 16, 16, 8, 4),    // M = 16384
    FLAGFFT_BLDS_HALVES(double, true, 512, 14, 12, 10, 6),   // M = 20160
    // split forms (few sequences): FP32
    FLAGFFT_BLDS_SPLIT4(float, false, 2, 512, 16, 16, 8, 4),    // M = 16384
    FLAGFFT_BLDS_SPLIT3(float, false, 4, 256, 16, 16, 16),      // M = 16384
    FLAGFFT_BLDS_SPLIT4(float, false, 2, 512, 16, 9, 10, 7),    // M = 20160
    FLAGFFT_BLDS_SPLIT4(float, false, 4, 512, 14, 12, 10, 3),   // M = 20160
    FLAGFFT_BLDS_SPLIT4(float, false, 2, 1024, 16, 16, 16, 4),  // M = 32768
    FLAGFFT_BLDS_SPLIT4(float, false, 4, 512, 16, 16, 8, 4),    // M = 32768
    FLAGFFT_BLDS_SPLIT4(float, false, 2, 1024, 16, 12, 15, 7),  // M = 40320
    FLAGFFT_BLDS_SPLIT4(float, false, 4, 512, 16, 9, 10, 7),    // M = 40320
    // FP64
    FLAGFFT_BLDS_SPLIT3(double, true, 2, 256, 16, 16, 16),      // M = 8192
    FLAGFFT_BLDS_SPLIT3(double, true, 4, 256, 16, 16, 8),       // M = 8192
    FLAGFFT_BLDS_SPLIT4(double, true, 2, 512, 14, 12, 10, 3),   // M = 10080
    FLAGFFT_BLDS_SPLIT3(double, true, 4, 256, 14, 12, 15),      // M = 10080
    FLAGFFT_BLDS_SPLIT4(double, true, 2, 512, 16, 16, 8, 4),    // M = 16384
    FLAGFFT_BLDS_SPLIT3(double, true, 4, 256, 16, 16, 16),      // M = 16384
    FLAGFFT_BLDS_SPLIT4(double, true, 2, 768, 14, 12, 10, 6),   // M = 20160 (768 lanes: 3 waves per SIMD, no spills; 11% faster than 512)
    FLAGFFT_BLDS_SPLIT4(double, true, 4, 768, 14, 12, 10, 3),   // M = 20160
};
constexpr int kBluesteinLdsCount = static_cast<int>(sizeof(kBluesteinLdsTable) / sizeof(kBluesteinLdsTable[0]));

int lds_part(const BluesteinLdsEntry &e) {
  int m = 1;
  for (int j = 0; j < e.nrad; ++j) m *= e.rad[j];
  return m;
}

}  // namespace

int bluestein_lds_kernel_id(const std::vector<int> &radices, bool dbl, bool halves) {
  for (int i = 0; i < kBluesteinLdsCount; ++i) {
    const BluesteinLdsEntry &e = kBluesteinLdsTable[i];
    if (e.dbl != dbl || e.halves != halves || e.split != 0 || static_cast<int>(radices.size()) != e.nrad) continue;
    bool same = true;
    for (int j = 0; j < e.nrad; ++j) same = same && radices[static_cast<size_t>(j)] == e.rad[j];
    if (same) return i;
  }
  return -1;
}

int bluestein_lds_ept(int id) { return kBluesteinLdsTable[id].e32; }
bool bluestein_lds_has_stockham(int id) { return kBluesteinLdsTable[id].e32 > 0; }
const void *bluestein_lds_kernel(int id, int variant) { return kBluesteinLdsTable[id].f[variant & 3]; }
int bluestein_lds_threads(int id, int variant) {
  const BluesteinLdsEntry &e = kBluesteinLdsTable[id];
  if (variant >= 2 || e.e32 == 0) return e.tpu;
  return (lds_part(e) + e.e32 - 1) / e.e32;
}
int bluestein_split_kernel_id(int64_t m, bool dbl, int split) {
  for (int i = 0; i < kBluesteinLdsCount; ++i) {
    const BluesteinLdsEntry &e = kBluesteinLdsTable[i];
    if (e.dbl == dbl && e.split == split && static_cast<int64_t>(lds_part(e)) * split == m) return i;
  }
  return -1;
}

std::vector<int> bluestein_lds_kernel_radices(int id) {
  return std::vector<int>(kBluesteinLdsTable[id].rad, kBluesteinLdsTable[id].rad + kBluesteinLdsTable[id].nrad);
}

const void *bluestein_combine_kernel(bool dbl, int split) {
  if (split == 2) return dbl ? (const void *)dev::k_bluestein_combine<double, 2> : (const void *)dev::k_bluestein_combine<float, 2>;
  if (split == 4) return dbl ? (const void *)dev::k_bluestein_combine<double, 4> : (const void *)dev::k_bluestein_combine<float, 4>;
  return nullptr;
}

// Whole-sequence instances of one precision (the planner's candidates).
std::vector<BluesteinLdsShape> bluestein_lds_candidates(bool dbl) {
  std::vector<BluesteinLdsShape> out;
  for (const BluesteinLdsEntry &e : kBluesteinLdsTable) {
    if (e.dbl != dbl || e.split != 0) continue;
    BluesteinLdsShape sh;
    sh.radices.assign(e.rad, e.rad + e.nrad);
    sh.halves = e.halves;
    sh.m = static_cast<int64_t>(lds_part(e)) * (e.halves ? 2 : 1);
    out.push_back(sh);
  }
  return out;
}

}  // namespace flagfft
