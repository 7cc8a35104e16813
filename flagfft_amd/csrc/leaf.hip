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
};

constexpr int leaf_lds_stride(int n) {
  const int ext = (n - 1) + ((n - 1) >> 5) + 1;
  return (ext % 2 == 0) ? ext + 1 : ext;
}

template <typename T, int E, int... Rs>
__global__ __launch_bounds__(kMaxBlockThreads) void k_leaf(const LeafArgs a) {
  using RS = RadixSeq<Rs...>;
  constexpr int N = RS::kN;
  constexpr int TPU = N / E;  // threads per transform
  constexpr int LS = leaf_lds_stride(N);
  static_assert(N % E == 0, "E must divide N");
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x;
  const int sl = tid / TPU;
  const int t = tid - sl * TPU;
  const int64_t seq = static_cast<int64_t>(blockIdx.x) * a.C + sl;
  const bool active = seq < a.nseq;
  LdsC<T> *buf = lds + sl * LS;
  const GlbC<T> *src = gc<T>(a.in) + (active ? seq * a.idist : 0);
  GlbC<T> *dst = gc<T>(a.out) + (active ? seq * a.odist : 0);
  const GlbC<T> *tw = gc<T>(a.tw);
  const T scale = static_cast<T>(a.scale);
  V2<T> x[E];
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    constexpr int RR = RS::kR[s];
    constexpr int q = RS::prior(s);
    constexpr int NB = N / RR;
    constexpr int MB = E / RR;
    static_assert(E % RR == 0, "every radix must divide E");
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int b = t + m * TPU;
#pragma unroll
      for (int j = 0; j < RR; ++j) {
        const int e = b + j * NB;
        V2<T> v;
        if constexpr (s == 0) {
          v = active ? ldg(src + e) : V2<T>{0, 0};
          if (a.conj_in) v.y = -v.y;
        } else {
          v = ldc(buf + lpad(e));
        }
        x[m * RR + j] = v;
      }
      if constexpr (q > 1) {
        constexpr int ts = N / (q * RR);
        const int k = b % q;
#pragma unroll
        for (int j = 1; j < RR; ++j) x[m * RR + j] = cmul(x[m * RR + j], ldg(tw + j * k * ts));
      }
      dft<RR>(&x[m * RR]);
    }
    if constexpr (s + 1 < RS::kS) {
      if constexpr (s > 0) __syncthreads();
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
        const int k = (q > 1) ? b % q : 0;
        const int base = (b - k) * RR + k;
#pragma unroll
        for (int j = 0; j < RR; ++j) stc(buf + lpad(base + j * q), x[m * RR + j]);
      }
      __syncthreads();
    } else if (active) {
      // last stage: q * RR == N, so k == b and the output index is b + j q
#pragma unroll
      for (int m = 0; m < MB; ++m) {
        const int b = t + m * TPU;
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
  });
}

}  // namespace dev

namespace {

struct LeafEntry {
  int nrad;
  int rad[4];
  int ept;
  const void *f32;
  const void *f64;
};

#define FLAGFFT_LEAF1(E, R0) \
  { 1, {R0, 0, 0, 0}, E, (const void *)dev::k_leaf<float, E, R0>, (const void *)dev::k_leaf<double, E, R0> }
#define FLAGFFT_LEAF2(E, R0, R1)                                                  \
  {                                                                               \
    2, {R0, R1, 0, 0}, E, (const void *)dev::k_leaf<float, E, R0, R1>,            \
        (const void *)dev::k_leaf<double, E, R0, R1>                              \
  }
#define FLAGFFT_LEAF3(E, R0, R1, R2)                                              \
  {                                                                               \
    3, {R0, R1, R2, 0}, E, (const void *)dev::k_leaf<float, E, R0, R1, R2>,       \
        (const void *)dev::k_leaf<double, E, R0, R1, R2>                          \
  }
#define FLAGFFT_LEAF4(E, R0, R1, R2, R3)                                          \
  {                                                                               \
    4, {R0, R1, R2, R3}, E, (const void *)dev::k_leaf<float, E, R0, R1, R2, R3>,  \
        (const void *)dev::k_leaf<double, E, R0, R1, R2, R3>                      \
  }

const LeafEntry kLeafTable[] = {
    FLAGFFT_LEAF1(2, 2),
    FLAGFFT_LEAF1(4, 4),
    FLAGFFT_LEAF1(8, 8),
    FLAGFFT_LEAF1(16, 16),
    FLAGFFT_LEAF2(16, 16, 2),
    FLAGFFT_LEAF2(16, 16, 4),
    FLAGFFT_LEAF2(16, 16, 8),
    FLAGFFT_LEAF2(16, 16, 16),
    FLAGFFT_LEAF3(16, 16, 16, 2),
    FLAGFFT_LEAF3(16, 16, 16, 4),
    FLAGFFT_LEAF3(16, 16, 16, 8),
    FLAGFFT_LEAF3(16, 16, 16, 16),
    FLAGFFT_LEAF4(16, 16, 16, 16, 2),
    FLAGFFT_LEAF3(10, 10, 10, 10),
    FLAGFFT_LEAF3(12, 12, 12, 12),
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

int leaf_kernel_ept(int id) { return kLeafTable[id].ept; }

hipError_t launch_leaf(const LeafArgs &a, bool dbl, int id, const LaunchShape &s, hipStream_t st) {
  const void *f = dbl ? kLeafTable[id].f64 : kLeafTable[id].f32;
  if (s.lds_bytes > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(s.lds_bytes));
    if (e != hipSuccess) return e;
  }
  LeafArgs copy = a;
  void *args[] = {&copy};
  hipError_t e = hipLaunchKernel(f, s.grid, s.block, args, s.lds_bytes, st);
  if (e != hipSuccess) return e;
  return hipGetLastError();
}

}  // namespace flagfft
