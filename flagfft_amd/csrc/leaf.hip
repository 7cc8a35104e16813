// leaf.hip — compile-time specialised single-pass transforms (gfx950).
//
// For the lengths listed in kLeafTable the whole radix sequence is a template
// parameter: every index, prior product, twiddle stride and LDS offset folds to
// a constant, each thread keeps its E elements in registers across stages,
// stage 0 reads HBM directly and the last stage writes HBM directly.  Only
// contiguous complex rows use it (the generic k_pass covers every other layout
// and mode).  Same arithmetic as the reference's leaf kernel
// (python/flagfft_codegen/kernels.py:853-976): a mixed-radix Stockham
// transform in on-chip memory, natural order in and out.
#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"
#include "fft_args.hpp"
#include "kernels.hpp"
#include "spec_common.hpp"

namespace flagfft {
namespace dev {

// Row modes of the specialised single-pass kernel.
//   kLeafComplex  length-N complex rows
//   kLeafR2C      real rows of 2N: z_j = x_2j + i x_2j+1, Z = FFT_N(z) (N = n/2),
//                 X_k = E_k + w_n^k O_k, X_{N-k} = conj(E_k - w_n^k O_k) with
//                 E_k = (Z_k + conj Z_{N-k})/2, O_k = -i (Z_k - conj Z_{N-k})/2, k <= N/2
//   kLeafC2R      Hermitian rows of N+1 bins: Z'_k = (X_k + conj X_{N-k})
//                 + i conj(w_n^k) (X_k - conj X_{N-k}) (Im of X_0, X_N dropped),
//                 z = conj FFT_N(conj Z') -> (x_2j, x_2j+1)
//   (LeafArgs::dc_nyq_pair: Re X_N travels as Im X_0, bin N is not accessed)
// (python/flagfft_codegen/kernels.py:1098-1221 promotes real rows to full-length
// complex transforms instead; the split-spectrum identity halves the work.)

template <typename T, int E, int Mode, int... Rs>
__global__ __launch_bounds__(kMaxBlockThreads) void k_leaf(const LeafArgs a) {
  using RS = RadixSeq<Rs...>;
  constexpr int N = RS::kN;
  constexpr int TPU = (N + E - 1) / E;  // threads per transform
  constexpr int LS = leaf_lds_stride(N);
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x;
  const int sl = tid / TPU;
  const int t = tid - sl * TPU;
  const int64_t seq = static_cast<int64_t>(blockIdx.x) * a.C + sl;
  const bool active = seq < a.nseq;
  LdsC<T> *buf = lds + sl * LS;
  const GlbC<T> *src = gc<T>(a.in) + (active ? seq * a.idist : 0);
  GlbC<T> *dst = gc<T>(a.out) + (active ? seq * a.odist : 0);
  const T scale = static_cast<T>(a.scale);
  const GlbC<T> *wn = gc<T>(a.aux);
  const GlbC<T> *tw = gc<T>(a.tw);
  // split-spectrum passes: thread t handles k = t + i TPU, k <= N/2; every
  // table (and, for C2R, input) load of them is issued before any is used,
  // so their latencies overlap instead of one round trip per k
  constexpr int NK = (N / 2 + TPU) / TPU;
  V2<T> wk[(Mode == kLeafComplex) ? 1 : NK];
  if constexpr (Mode != kLeafComplex) {
#pragma unroll
    for (int i = 0; i < NK; ++i) {
      const int k = t + i * TPU;
      wk[i] = (k <= N / 2) ? ldg(wn + k) : V2<T>{1, 0};
    }
  }
  if constexpr (Mode == kLeafC2R) {
    // fused pre-process into LDS (conjugated: the inverse runs as a forward transform)
    V2<T> xks[NK], xhs[NK];
#pragma unroll
    for (int i = 0; i < NK; ++i) {
      const int k = t + i * TPU;
      const bool ok = active && k <= N / 2;
      xks[i] = ok ? ldg(src + k) : V2<T>{0, 0};
      // a paired row carries X_N's real part in the imaginary part of X_0
      xhs[i] = (ok && !(k == 0 && a.dc_nyq_pair)) ? ldg(src + (N - k)) : V2<T>{0, 0};
    }
#pragma unroll
    for (int i = 0; i < NK; ++i) {
      const int k = t + i * TPU;
      if (!active || k > N / 2) break;
      V2<T> xk = xks[i];
      V2<T> xh = xhs[i];
      if (k == 0) {
        if (a.dc_nyq_pair) xh.x = xk.y;
        xk.y = T(0);
        xh.y = T(0);
      }
      const V2<T> w = wk[i];
      {
        const V2<T> c = cconj(xh);
        const V2<T> z = cadd(cadd(xk, c), cmul_pi(cmul(csub(xk, c), cconj(w))));
        stc(buf + lpad(k), cconj(z));
      }
      if (k != 0 && 2 * k != N) {
        // partner N - k: roles of X_k and X_{N-k} swap; conj(w_n^{N-k}) = -w_n^k
        const V2<T> c = cconj(xk);
        const V2<T> z = cadd(cadd(xh, c), cmul_pi(cmul(csub(xh, c), V2<T>{-w.x, -w.y})));
        stc(buf + lpad(N - k), cconj(z));
      }
    }
    __syncthreads();
  }
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    constexpr int RR = RS::kR[s];
    constexpr int q = RS::prior(s);
    constexpr int NB = N / RR;
    constexpr int MB = RS::template butterflies<E>(s);
    V2<T> x[MB * RR];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int b = t + m * TPU;
      if (b < NB) {
#pragma unroll
        for (int j = 0; j < RR; ++j) {
          const int e = b + j * NB;
          V2<T> v;
          if constexpr (s == 0 && Mode != kLeafC2R) {
            v = active ? ldg(src + e) : V2<T>{0, 0};
            if (a.conj_in) v.y = -v.y;
          } else {
            v = ldc(buf + lpad(e));
          }
          x[m * RR + j] = v;
        }
        if constexpr (q > 1) {
          // w^(j k) by binary powers of the butterfly's base twiddle: one
          // table load per butterfly instead of RR - 1
          constexpr int ts = N / (q * RR);
          twiddle_powers<T, RR>(&x[m * RR], ldg(tw + (b % q) * ts));
        }
        dft<RR>(&x[m * RR]);
      }
    }
    if constexpr (s + 1 < RS::kS) {
      if constexpr (s > 0 || Mode == kLeafC2R) __syncthreads();
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        if (b < NB) {
          const int k = (q > 1) ? b % q : 0;
          const int base = (b - k) * RR + k;
#pragma unroll
          for (int j = 0; j < RR; ++j) stc(buf + lpad(base + j * q), x[m * RR + j]);
        }
      }
      __syncthreads();
    } else if constexpr (Mode == kLeafR2C) {
      // natural-order Z into LDS, then the fused split-spectrum post-process
      if constexpr (s > 0) __syncthreads();
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        if (b < NB) {
#pragma unroll
          for (int j = 0; j < RR; ++j) stc(buf + lpad(b + j * q), x[m * RR + j]);
        }
      }
      __syncthreads();
      const T half = T(0.5);
#pragma unroll
      for (int i = 0; i < NK; ++i) {
        const int k = t + i * TPU;
        if (!active || k > N / 2) break;
        const V2<T> zk = ldc(buf + lpad(k));
        const V2<T> zc = cconj(ldc(buf + lpad(k == 0 ? 0 : N - k)));
        const V2<T> ev{(zk.x + zc.x) * half, (zk.y + zc.y) * half};
        const V2<T> od = cmul_mi(V2<T>{(zk.x - zc.x) * half, (zk.y - zc.y) * half});
        const V2<T> wo = cmul(wk[i], od);
        V2<T> lo = cadd(ev, wo);
        V2<T> hi = cconj(csub(ev, wo));
        if (a.conj_out) {
          lo.y = -lo.y;
          hi.y = -hi.y;
        }
        if (k == 0 && a.dc_nyq_pair) {
          stg(dst, V2<T>{lo.x, hi.x});  // X_0 and X_N are real: one slot
        } else {
          stg(dst + k, lo);
          if (2 * k != N) stg(dst + (N - k), hi);
        }
      }
    } else if (active) {
      // last stage: q * RR == N, so k == b and the output index is b + j q
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        if (b < NB) {
#pragma unroll
          for (int j = 0; j < RR; ++j) {
            V2<T> v = x[m * RR + j];
            if (a.conj_out) v.y = -v.y;
            if (scale != T(1)) {
              v.x *= scale;
              v.y *= scale;
            }
            stg(dst + b + j * q, v);
          }
        }
      }
    }
  });
}

}  // namespace dev

namespace {

struct LeafEntry {
  int nrad;
  int rad[4];
  int e32, e64;        // elements per thread of the FP32 / FP64 instances
  const void *f[2][3];  // [FP64][LeafMode]
};

#define FLAGFFT_LEAF_FNS(E32, E64, ...)                                                            \
  {                                                                                                \
    {(const void *)dev::k_leaf<float, E32, dev::kLeafComplex, __VA_ARGS__>,                         \
     (const void *)dev::k_leaf<float, E32, dev::kLeafR2C, __VA_ARGS__>,                             \
     (const void *)dev::k_leaf<float, E32, dev::kLeafC2R, __VA_ARGS__>},                            \
    {                                                                                              \
      (const void *)dev::k_leaf<double, E64, dev::kLeafComplex, __VA_ARGS__>,                       \
          (const void *)dev::k_leaf<double, E64, dev::kLeafR2C, __VA_ARGS__>,                       \
          (const void *)dev::k_leaf<double, E64, dev::kLeafC2R, __VA_ARGS__>                        \
    }                                                                                              \
  }
#define FLAGFFT_LEAF1(E32, E64, R0) {1, {R0, 0, 0, 0}, E32, E64, FLAGFFT_LEAF_FNS(E32, E64, R0)}
#define FLAGFFT_LEAF2(E32, E64, R0, R1) {2, {R0, R1, 0, 0}, E32, E64, FLAGFFT_LEAF_FNS(E32, E64, R0, R1)}
#define FLAGFFT_LEAF3(E32, E64, R0, R1, R2) {3, {R0, R1, R2, 0}, E32, E64, FLAGFFT_LEAF_FNS(E32, E64, R0, R1, R2)}
#define FLAGFFT_LEAF4(E32, E64, R0, R1, R2, R3) \
  {4, {R0, R1, R2, R3}, E32, E64, FLAGFFT_LEAF_FNS(E32, E64, R0, R1, R2, R3)}

const LeafEntry kLeafTable[] = {
    FLAGFFT_LEAF1(2, 2, 2),
    FLAGFFT_LEAF1(4, 4, 4),
    FLAGFFT_LEAF1(8, 8, 8),
    FLAGFFT_LEAF1(16, 16, 16),
    FLAGFFT_LEAF2(16, 16, 16, 2),
    FLAGFFT_LEAF2(16, 16, 16, 4),
    FLAGFFT_LEAF2(16, 16, 16, 8),
    FLAGFFT_LEAF2(16, 16, 16, 16),
    FLAGFFT_LEAF3(16, 16, 16, 16, 2),
    FLAGFFT_LEAF3(16, 16, 16, 16, 4),
    FLAGFFT_LEAF3(16, 16, 16, 16, 8),
    FLAGFFT_LEAF3(16, 16, 16, 16, 16),
    FLAGFFT_LEAF4(16, 16, 16, 16, 16, 2),
    FLAGFFT_LEAF3(16, 16, 10, 10, 10),
    FLAGFFT_LEAF4(16, 16, 15, 10, 10, 2),
    FLAGFFT_LEAF3(16, 16, 16, 10, 10),
    FLAGFFT_LEAF3(16, 16, 9, 9, 9),
    FLAGFFT_LEAF4(16, 16, 15, 10, 10, 4),  // 6000 (and the R2C / C2R halves of 12000)
    // 3, 5, 7, 9 and 15 times a power of two, 1000 times one (the
    // run-time-radix single pass is about 3x slower)
    FLAGFFT_LEAF3(16, 16, 16, 12, 2),       // 384
    FLAGFFT_LEAF3(16, 16, 16, 12, 4),       // 768
    FLAGFFT_LEAF3(16, 16, 16, 12, 8),       // 1536
    FLAGFFT_LEAF3(16, 16, 16, 16, 12),      // 3072
    FLAGFFT_LEAF4(16, 16, 16, 16, 12, 2),   // 6144
    FLAGFFT_LEAF3(16, 16, 16, 10, 4),       // 640
    FLAGFFT_LEAF3(16, 16, 16, 10, 8),       // 1280
    FLAGFFT_LEAF3(16, 16, 16, 16, 10),      // 2560
    FLAGFFT_LEAF4(16, 16, 16, 16, 10, 2),   // 5120
    FLAGFFT_LEAF3(16, 16, 16, 15, 4),       // 960
    FLAGFFT_LEAF3(16, 16, 16, 15, 8),       // 1920
    FLAGFFT_LEAF3(16, 16, 16, 16, 15),      // 3840
    FLAGFFT_LEAF3(16, 16, 15, 10, 8),       // 1200
    FLAGFFT_LEAF3(16, 16, 16, 15, 10),      // 2400
    FLAGFFT_LEAF4(16, 16, 16, 15, 10, 2),   // 4800
    FLAGFFT_LEAF4(16, 16, 16, 10, 10, 2),   // 3200
    FLAGFFT_LEAF4(16, 16, 16, 10, 10, 4),   // 6400
    FLAGFFT_LEAF3(16, 16, 16, 9, 8),        // 1152
    FLAGFFT_LEAF3(16, 16, 16, 16, 9),       // 2304
    FLAGFFT_LEAF4(16, 16, 16, 16, 9, 2),    // 4608
    FLAGFFT_LEAF3(16, 16, 16, 14, 4),       // 896
    FLAGFFT_LEAF3(16, 16, 16, 14, 8),       // 1792
    FLAGFFT_LEAF3(16, 16, 16, 16, 14),      // 3584
    FLAGFFT_LEAF4(16, 16, 16, 16, 14, 2),   // 7168
    FLAGFFT_LEAF3(16, 16, 15, 12, 8),       // 1440
    FLAGFFT_LEAF3(16, 16, 16, 15, 12),      // 2880
    FLAGFFT_LEAF4(16, 16, 16, 15, 12, 2),   // 5760
    FLAGFFT_LEAF4(16, 16, 10, 10, 10, 2),   // 2000
    FLAGFFT_LEAF4(16, 16, 10, 10, 10, 4),   // 4000
    FLAGFFT_LEAF4(16, 16, 10, 10, 10, 8),   // 8000
    // whole-sequence-in-LDS lengths above the default single-pass limit
    FLAGFFT_LEAF4(22, 22, 15, 9, 9, 9),    // 10935 = 3^7 * 5
    FLAGFFT_LEAF4(20, 20, 10, 10, 10, 10),  // 10000 (FP32: 80 KB of LDS; FP64 runs two passes)
};

}  // namespace

int leaf_kernel_id(const std::vector<int> &radices) {
  const int n = static_cast<int>(sizeof(kLeafTable) / sizeof(kLeafTable[0]));
  for (int i = 0; i < n; ++i) {
    const LeafEntry &e = kLeafTable[i];
    if (static_cast<int>(radices.size()) != e.nrad) continue;
    bool same = true;
    for (int j = 0; j < e.nrad; ++j) same = same && radices[static_cast<size_t>(j)] == e.rad[j];
    if (same) return i;
  }
  return -1;
}

int leaf_kernel_ept(int id, bool dbl) { return dbl ? kLeafTable[id].e64 : kLeafTable[id].e32; }

const void *leaf_kernel(bool dbl, int id, int mode) { return kLeafTable[id].f[dbl ? 1 : 0][mode]; }

}  // namespace flagfft
