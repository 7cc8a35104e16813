// leaf_split.hip — single-pass transforms of few sequences spread over P
// workgroups per sequence (gfx950).
//
// Decimation in frequency by P: for N = P L,
//   X[P k + h] = FFT_L(y_h)[k],  y_h[n1] = w_N^(n1 h) sum_{n2 < P} x[n1 + L n2] w_P^(n2 h),
// so workgroup (q, h) reads the whole sequence q (the P workgroups of a
// sequence share one XCD, whose L2 serves the repeated reads), forms y_h on
// the fly in its first stage, runs the L-point Stockham stages in LDS and
// writes every P-th output.  One launch, P times the workgroups of the
// one-workgroup-per-sequence leaf; for batches smaller than the CU count.
#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"
#include "fft_args.hpp"
#include "kernels.hpp"
#include "spec_common.hpp"

namespace flagfft {
namespace dev {

template <typename T, int E, int P, int... Rs>
__global__ __launch_bounds__(kMaxBlockThreads) void k_leaf_split(const LeafSplitArgs a) {
  using RS = RadixSeq<Rs...>;
  constexpr int L = RS::kN;
  constexpr int TPU = (L + E - 1) / E;
  const int t = threadIdx.x;
  uint32_t q, h;
  const uint32_t nseq = static_cast<uint32_t>(a.nseq);
  if ((nseq & 7u) == 0) {
    const uint32_t local = blockIdx.x >> 3;
    q = (local / P) * 8u + (blockIdx.x & 7u);
    h = local % P;
  } else {
    q = blockIdx.x / P;
    h = blockIdx.x % P;
  }
  const GlbC<T> *src = gc<T>(a.in) + static_cast<int64_t>(q) * a.idist;
  GlbC<T> *dst = gc<T>(a.out) + static_cast<int64_t>(q) * a.odist;
  const GlbC<T> *twn = gc<T>(a.tw_n);
  const T scale = static_cast<T>(a.scale);
  // w_P^(n2 h) = w_N^((n2 h mod P) L)
  V2<T> wp[P];
#pragma unroll
  for (int n2 = 0; n2 < P; ++n2) wp[n2] = ldt(twn, static_cast<uint32_t>(((n2 * h) % P) * L));
  const V2<T> dummy[1] = {V2<T>{1, 0}};
  run_stages<T, E, TPU, RS, false, 5, true>(
      lds_base<T>(), t, dummy, gc<T>(a.tw_l), false,
      [&](int e) __attribute__((always_inline)) {
        V2<T> v{0, 0};
#pragma unroll
        for (int n2 = 0; n2 < P; ++n2) {
          V2<T> x = ldt(src, static_cast<uint32_t>(e + L * n2));
          if (a.conj_in) x.y = -x.y;
          v = n2 == 0 ? x : cadd(v, cmul(x, wp[n2]));
        }
        return cmul(v, ldt(twn, static_cast<uint32_t>(e) * h));
      },
      [&](int idx, V2<T> v) __attribute__((always_inline)) {
        if (a.conj_out) v.y = -v.y;
        v.x *= scale;
        v.y *= scale;
        stg(dst + (static_cast<int64_t>(P) * idx + h), v);
      });
}

}  // namespace dev

namespace {

struct LeafSplitEntry {
  int n, p, nrad;
  int rad[4];
  int e32, e64;
  const void *f[2];  // [FP64]
};

#define FLAGFFT_LSPLIT3(N, P, E32, E64, R0, R1, R2)                                                          \
  {                                                                                                          \
    N, P, 3, {R0, R1, R2, 0}, E32, E64, {                                                                    \
      (const void *)dev::k_leaf_split<float, E32, P, R0, R1, R2>,                                            \
          (const void *)dev::k_leaf_split<double, E64, P, R0, R1, R2>                                        \
    }                                                                                                        \
  }
#define FLAGFFT_LSPLIT4(N, P, E32, E64, R0, R1, R2, R3)                                                      \
  {                                                                                                          \
    N, P, 4, {R0, R1, R2, R3}, E32, E64, {                                                                   \
      (const void *)dev::k_leaf_split<float, E32, P, R0, R1, R2, R3>,                                        \
          (const void *)dev::k_leaf_split<double, E64, P, R0, R1, R2, R3>                                    \
    }                                                                                                        \
  }

const LeafSplitEntry kLeafSplitTable[] = {
    FLAGFFT_LSPLIT3(8192, 4, 8, 8, 16, 16, 8),      // L = 2048
    FLAGFFT_LSPLIT3(16384, 8, 8, 8, 16, 16, 8),     // L = 2048
    FLAGFFT_LSPLIT4(10935, 5, 9, 9, 9, 9, 9, 3),    // L = 2187 = 3^7
};

}  // namespace

int leaf_split_kernel_id(int64_t n) {
  const int cnt = static_cast<int>(sizeof(kLeafSplitTable) / sizeof(kLeafSplitTable[0]));
  for (int i = 0; i < cnt; ++i)
    if (kLeafSplitTable[i].n == n) return i;
  return -1;
}

void leaf_split_shape(int id, bool dbl, int &p, int &l, int &ept, std::vector<int> &radices) {
  const LeafSplitEntry &e = kLeafSplitTable[id];
  p = e.p;
  l = e.n / e.p;
  ept = dbl ? e.e64 : e.e32;
  radices.assign(e.rad, e.rad + e.nrad);
}

const void *leaf_split_kernel(int id, bool dbl) { return kLeafSplitTable[id].f[dbl ? 1 : 0]; }

}  // namespace flagfft
