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

// Mode (leaf.hip's LeafMode): kLeafComplex rows of N complex; kLeafR2C rows
// of 2N reals read as N complex pairs z, output N + 1 bins; kLeafC2R N + 1
// Hermitian bins in, 2N reals out.  The real modes use P = 2: the
// split-spectrum partner N - k of a bin k = 2k' + h is in the same block
// (N even), so pack and unpack stay fused.
//   R2C  X_k = E_k + w_2N^k O_k, X_{N-k} = conj(E_k - w_2N^k O_k),
//        E_k = (Z_k + conj Z_{N-k}) / 2, O_k = -i (Z_k - conj Z_{N-k}) / 2
//   C2R  c_k = conj((X_k + conj X_{N-k}) + i conj(w_2N^k) (X_k - conj X_{N-k}))
//        (Im of X_0, X_N dropped), z = conj FFT_N(c) -> (x_2j, x_2j+1)
template <typename T, int E, int P, int Mode, int... Rs>
__global__ __launch_bounds__(kMaxBlockThreads) void k_leaf_split(const LeafSplitArgs a) {
  using RS = RadixSeq<Rs...>;
  constexpr int L = RS::kN;
  constexpr int N = P * L;
  constexpr int TPU = (L + E - 1) / E;
  constexpr int PS = 5;
  static_assert(Mode == kLeafComplex || P == 2, "real modes split in two");
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
  const GlbC<T> *wn = gc<T>(a.aux);
  const T scale = static_cast<T>(a.scale);
  LdsC<T> *buf = lds_base<T>();
  auto lpad = [](int e) __attribute__((always_inline)) { return e + (e >> PS); };
  // w_P^(n2 h) = w_N^((n2 h mod P) L)
  V2<T> wp[P];
#pragma unroll
  for (int n2 = 0; n2 < P; ++n2) wp[n2] = ldt(twn, static_cast<uint32_t>(((n2 * h) % P) * L));
  // element m of the sequence the N-point core transforms
  auto elem = [&](int m) __attribute__((always_inline)) {
    if constexpr (Mode == kLeafC2R) {
      V2<T> xk = ldt(src, static_cast<uint32_t>(m));
      V2<T> xh = ldt(src, static_cast<uint32_t>(N - m));
      if (m == 0) {
        xk.y = T(0);
        xh.y = T(0);
      }
      const V2<T> c = cconj(xh);
      return cconj(cadd(cadd(xk, c), cmul_pi(cmul(csub(xk, c), cconj(ldt(wn, static_cast<uint32_t>(m)))))));
    } else {
      V2<T> x = ldt(src, static_cast<uint32_t>(m));
      if (a.conj_in) x.y = -x.y;
      return x;
    }
  };
  auto load0 = [&](int e) __attribute__((always_inline)) {
    V2<T> v{0, 0};
#pragma unroll
    for (int n2 = 0; n2 < P; ++n2) {
      const V2<T> x = elem(e + L * n2);
      v = n2 == 0 ? x : cadd(v, cmul(x, wp[n2]));
    }
    return cmul(v, ldt(twn, static_cast<uint32_t>(e) * h));
  };
  const V2<T> dummy[1] = {V2<T>{1, 0}};
  if constexpr (Mode == kLeafR2C) {
    // Z[2k' + h] into LDS (natural order k'), then the fused unpack
    run_stages<T, E, TPU, RS, false, PS, true>(buf, t, dummy, gc<T>(a.tw_l), true, load0,
                                               [&](int idx, V2<T> v) __attribute__((always_inline)) {
                                                 stc(buf + lpad(idx), v);
                                               });
    __syncthreads();
    const T half = T(0.5);
    for (int kp = t; kp < L; kp += TPU) {
      const int k = P * kp + static_cast<int>(h);
      if (2 * k > N) continue;
      const int partner = h == 0 ? (kp == 0 ? 0 : L - kp) : L - 1 - kp;  // slot of Z[N - k]
      const V2<T> zk = ldc(buf + lpad(kp));
      const V2<T> zc = cconj(ldc(buf + lpad(partner)));
      const V2<T> ev{(zk.x + zc.x) * half, (zk.y + zc.y) * half};
      const V2<T> od = cmul_mi(V2<T>{(zk.x - zc.x) * half, (zk.y - zc.y) * half});
      const V2<T> wo = cmul(ldt(wn, static_cast<uint32_t>(k)), od);
      V2<T> lo = cadd(ev, wo);
      V2<T> hi = cconj(csub(ev, wo));
      if (a.conj_out) {
        lo.y = -lo.y;
        hi.y = -hi.y;
      }
      stg(dst + k, lo);
      if (2 * k != N) stg(dst + (N - k), hi);
    }
  } else {
    run_stages<T, E, TPU, RS, false, PS, true>(buf, t, dummy, gc<T>(a.tw_l), false, load0,
                                               [&](int idx, V2<T> v) __attribute__((always_inline)) {
                                                 if (a.conj_out) v.y = -v.y;
                                                 v.x *= scale;
                                                 v.y *= scale;
                                                 stg(dst + (static_cast<int64_t>(P) * idx + h), v);
                                               });
  }
}

}  // namespace dev

namespace {

struct LeafSplitEntry {
  int n, p, nrad;
  int rad[4];
  int e32, e64;
  const void *f[2][3];  // [FP64][LeafMode]; real modes only for P = 2
};

#define FLAGFFT_LSPLIT_C(T, E, P, ...) (const void *)dev::k_leaf_split<T, E, P, dev::kLeafComplex, __VA_ARGS__>
#define FLAGFFT_LSPLIT_R(T, E, P, M, ...) (const void *)dev::k_leaf_split<T, E, P, M, __VA_ARGS__>
// complex only
#define FLAGFFT_LSPLIT3(N, P, E32, E64, R0, R1, R2)                                                  \
  {                                                                                                  \
    N, P, 3, {R0, R1, R2, 0}, E32, E64, {                                                            \
      {FLAGFFT_LSPLIT_C(float, E32, P, R0, R1, R2), nullptr, nullptr},                               \
      {FLAGFFT_LSPLIT_C(double, E64, P, R0, R1, R2), nullptr, nullptr}                               \
    }                                                                                                \
  }
#define FLAGFFT_LSPLIT4(N, P, E32, E64, R0, R1, R2, R3)                                              \
  {                                                                                                  \
    N, P, 4, {R0, R1, R2, R3}, E32, E64, {                                                           \
      {FLAGFFT_LSPLIT_C(float, E32, P, R0, R1, R2, R3), nullptr, nullptr},                           \
      {FLAGFFT_LSPLIT_C(double, E64, P, R0, R1, R2, R3), nullptr, nullptr}                           \
    }                                                                                                \
  }
// complex and both real modes, P = 2
#define FLAGFFT_LSPLIT3R(N, E32, E64, R0, R1, R2)                                                    \
  {                                                                                                  \
    N, 2, 3, {R0, R1, R2, 0}, E32, E64, {                                                            \
      {FLAGFFT_LSPLIT_C(float, E32, 2, R0, R1, R2), FLAGFFT_LSPLIT_R(float, E32, 2, dev::kLeafR2C, R0, R1, R2), \
       FLAGFFT_LSPLIT_R(float, E32, 2, dev::kLeafC2R, R0, R1, R2)},                                  \
      {FLAGFFT_LSPLIT_C(double, E64, 2, R0, R1, R2), FLAGFFT_LSPLIT_R(double, E64, 2, dev::kLeafR2C, R0, R1, R2), \
       FLAGFFT_LSPLIT_R(double, E64, 2, dev::kLeafC2R, R0, R1, R2)}                                  \
    }                                                                                                \
  }

const LeafSplitEntry kLeafSplitTable[] = {
    FLAGFFT_LSPLIT3(8192, 4, 8, 8, 16, 16, 8),      // L = 2048
    FLAGFFT_LSPLIT3(16384, 8, 8, 8, 16, 16, 8),     // L = 2048
    FLAGFFT_LSPLIT4(10935, 5, 9, 9, 9, 9, 9, 3),    // L = 2187 = 3^7
    FLAGFFT_LSPLIT3R(3000, 10, 10, 15, 10, 10),      // L = 1500; real rows of 6000
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

const void *leaf_split_kernel(int id, bool dbl, int mode) { return kLeafSplitTable[id].f[dbl ? 1 : 0][mode]; }

}  // namespace flagfft
