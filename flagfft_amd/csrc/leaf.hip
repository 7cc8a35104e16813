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
  const GlbC<T> *tw = gc<T>(a.tw);
  const T scale = static_cast<T>(a.scale);
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    constexpr int RR = RS::kR[s];
    constexpr int q = RS::prior(s);
    constexpr int NB = N / RR;
    constexpr int MB = (E + RR - 1) / RR;  // butterflies per thread in this stage
    V2<T> x[MB * RR];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int b = t + m * TPU;
      if (b < NB) {
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
    }
    if constexpr (s + 1 < RS::kS) {
      if constexpr (s > 0) __syncthreads();
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

// ---------------------------------------------------------------------------
// One Stockham super-stage of a multi-pass transform (compile-time radices,
// C adjacent butterflies per workgroup).  Unit u = b * tcnt + i of sequence b
// reads x[i + e*tcnt] (e < R), multiplies by w_{pR}^{e (i mod p)}, transforms
// and writes y[(i - k) R + k + e p].  Lane -> (slot = tid % C, t = tid / C) so
// a wavefront's HBM reads are C-element row segments.  With block_store the
// C units' outputs form one contiguous block (first super-stage of contiguous
// sequences) and leave through LDS as one coalesced copy; otherwise each slot
// stores directly.  Bluestein chains use the chirp source (first pass) and
// the pointwise / chirp sinks (last passes), as the generic pass does.
// ---------------------------------------------------------------------------
template <typename T, int E, int C, int... Rs>
__global__ __launch_bounds__(kMaxBlockThreads) void k_stage(const StageArgs a) {
  using RS = RadixSeq<Rs...>;
  constexpr int R = RS::kN;
  constexpr int TPU = (R + E - 1) / E;
  constexpr int LS = leaf_lds_stride(R);
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x;
  const int sl = tid % C;
  const int t = tid / C;
  const int blk = remap_block(blockIdx.x, a.nblocks, a.xcd_remap);
  const int64_t u0 = static_cast<int64_t>(blk) * C;
  const int64_t u = u0 + sl;
  const bool active = u < a.units;
  const uint32_t uu = active ? static_cast<uint32_t>(u) : 0u;
  const uint32_t bq = fdiv(uu, a.fd_tcnt);
  const uint32_t i = uu - bq * static_cast<uint32_t>(a.tcnt);
  const bool first = a.p == 1;
  const uint32_t k = first ? 0u : i - fdiv(i, a.fd_p) * static_cast<uint32_t>(a.p);
  const int64_t tcnt = a.tcnt;
  const GlbC<T> *src = gc<T>(a.in) + static_cast<int64_t>(bq) * a.bs_in;
  const GlbC<T> *tws = gc<T>(a.tw_small);
  const GlbC<T> *tw_lo = gc<T>(a.tw_lo);
  const GlbC<T> *tw_hi = gc<T>(a.tw_hi);
  const GlbC<T> *chirp = gc<T>(a.chirp);
  const uint32_t tw_mask = (1u << a.tw_bits) - 1u;
  const uint32_t kmul = k * static_cast<uint32_t>(a.twp_mult);
  const T scale = static_cast<T>(a.scale);
  const bool chirp_in = (a.src_mode & kSrcChirpFlag) != 0;
  LdsC<T> *buf = lds + sl * LS;
  sfor<RS::kS>([&](auto S) __attribute__((always_inline)) {
    constexpr int s = decltype(S)::value;
    constexpr int RR = RS::kR[s];
    constexpr int q = RS::prior(s);
    constexpr int NB = R / RR;
    constexpr int MB = (E + RR - 1) / RR;  // butterflies per thread in this stage
    V2<T> x[MB * RR];
#pragma unroll
    for (int m = 0; m < MB; ++m) {
      const int b = t + m * TPU;
      if (b >= NB) continue;
#pragma unroll
      for (int j = 0; j < RR; ++j) {
        const int e = b + j * NB;
        V2<T> v;
        if constexpr (s == 0) {
          const int64_t elog = static_cast<int64_t>(i) + e * tcnt;
          if (!active || (chirp_in && elog >= a.nlog)) {
            v = V2<T>{0, 0};
          } else {
            v = ldg(src + elog);
            if (a.conj_in) v.y = -v.y;
            if (chirp_in) v = cmul(v, ldg(chirp + elog));
          }
          if (!first) {
            const uint32_t mm = static_cast<uint32_t>(e) * kmul;
            v = cmul(v, cmul(ldg(tw_hi + (mm >> a.tw_bits)), ldg(tw_lo + (mm & tw_mask))));
          }
        } else {
          v = ldc(buf + lpad(e));
        }
        x[m * RR + j] = v;
      }
      if constexpr (q > 1) {
        constexpr int ts = R / (q * RR);
        const int kk = b % q;
#pragma unroll
        for (int j = 1; j < RR; ++j) x[m * RR + j] = cmul(x[m * RR + j], ldg(tws + j * kk * ts));
      }
      dft<RR>(&x[m * RR]);
    }
    if constexpr (s + 1 < RS::kS) {
      if constexpr (s > 0) __syncthreads();
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
      // last LDS stage: natural-order output element b + j q of this unit
      if (a.block_store) {
        if constexpr (s > 0) __syncthreads();
#pragma unroll
        for (int m = 0; m < MB; ++m) {
          const int b = t + m * TPU;
          if (b >= NB) continue;
#pragma unroll
          for (int j = 0; j < RR; ++j) stc(buf + lpad(b + j * q), x[m * RR + j]);
        }
      } else if (active) {
#pragma unroll
        for (int m = 0; m < MB; ++m) {
          const int b = t + m * TPU;
          if (b >= NB) continue;
#pragma unroll
          for (int j = 0; j < RR; ++j) {
            const int64_t oidx = (i - k) * static_cast<int64_t>(R) + k + static_cast<int64_t>(b + j * q) * a.p;
            dst_elem<T>(a.out, static_cast<int64_t>(bq) * a.bs_out, 1, a.dst_mode, a.nlog, a.chirp, a.bhat,
                        a.conj_out, scale, oidx, x[m * RR + j]);
          }
        }
      }
    }
  });
  if (a.block_store) {
    // the workgroup's C units own one contiguous block of C * R output elements
    __syncthreads();
    GlbC<T> *dst = gc<T>(a.out) + u0 * R;
    const int64_t total = a.units * static_cast<int64_t>(R);
    constexpr int NT = C * TPU;
    batched<T, 8>(
        C * R, tid, NT,
        [&](int f) __attribute__((always_inline)) {
          const int s2 = f / R, e2 = f - s2 * R;
          V2<T> v = ldc(lds + s2 * LS + lpad(e2));
          if (a.conj_out) v.y = -v.y;
          if (scale != T(1)) {
            v.x *= scale;
            v.y *= scale;
          }
          return v;
        },
        [&](int f, V2<T> v) __attribute__((always_inline)) {
          if (u0 * R + f < total) stg(dst + f, v);
        });
  }
}

}  // namespace dev

namespace {

struct StageEntry {
  int nrad;
  int rad[3];
  int c32, e32, c64, e64;
  const void *f32;
  const void *f64;
};

#define FLAGFFT_STAGE2(C32, E32, C64, E64, R0, R1)                                                \
  {                                                                                               \
    2, {R0, R1, 0}, C32, E32, C64, E64, (const void *)dev::k_stage<float, E32, C32, R0, R1>,      \
        (const void *)dev::k_stage<double, E64, C64, R0, R1>                                      \
  }
#define FLAGFFT_STAGE3(C32, E32, C64, E64, R0, R1, R2)                                            \
  {                                                                                               \
    3, {R0, R1, R2}, C32, E32, C64, E64, (const void *)dev::k_stage<float, E32, C32, R0, R1, R2>, \
        (const void *)dev::k_stage<double, E64, C64, R0, R1, R2>                                  \
  }

const StageEntry kStageTable[] = {
    FLAGFFT_STAGE2(16, 16, 8, 16, 16, 4),       // 64
    FLAGFFT_STAGE2(32, 16, 16, 16, 16, 5),      // 80
    FLAGFFT_STAGE2(32, 16, 16, 16, 16, 6),      // 96
    FLAGFFT_STAGE2(16, 16, 8, 16, 16, 8),       // 128
    FLAGFFT_STAGE2(16, 16, 8, 16, 16, 10),      // 160
    FLAGFFT_STAGE2(16, 16, 8, 16, 16, 16),      // 256
    FLAGFFT_STAGE3(16, 16, 8, 16, 16, 16, 2),   // 512
    FLAGFFT_STAGE3(8, 16, 4, 16, 16, 16, 4),    // 1024
    FLAGFFT_STAGE3(4, 16, 4, 16, 16, 16, 8),    // 2048
    FLAGFFT_STAGE3(4, 32, 2, 16, 16, 16, 16),   // 4096
    FLAGFFT_STAGE2(32, 16, 16, 16, 9, 9),       // 81
    FLAGFFT_STAGE2(16, 16, 8, 16, 15, 9),       // 135
    FLAGFFT_STAGE3(8, 16, 4, 16, 9, 9, 9),      // 729
};

}  // namespace

int stage_kernel_id(const std::vector<int> &radices) {
  const int n = static_cast<int>(sizeof(kStageTable) / sizeof(kStageTable[0]));
  for (int i = 0; i < n; ++i) {
    const StageEntry &e = kStageTable[i];
    if (static_cast<int>(radices.size()) != e.nrad) continue;
    bool same = true;
    for (int j = 0; j < e.nrad; ++j) same = same && radices[static_cast<size_t>(j)] == e.rad[j];
    if (same) return i;
  }
  return -1;
}

void stage_kernel_shape(int id, bool dbl, int &units_per_group, int &ept) {
  units_per_group = dbl ? kStageTable[id].c64 : kStageTable[id].c32;
  ept = dbl ? kStageTable[id].e64 : kStageTable[id].e32;
}

const void *stage_kernel(bool dbl, int id) { return dbl ? kStageTable[id].f64 : kStageTable[id].f32; }

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
    FLAGFFT_LEAF3(16, 10, 10, 10),
    FLAGFFT_LEAF4(16, 15, 10, 10, 2),
    FLAGFFT_LEAF3(16, 16, 10, 10),
    FLAGFFT_LEAF3(16, 9, 9, 9),
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

const void *leaf_kernel(bool dbl, int id) { return dbl ? kLeafTable[id].f64 : kLeafTable[id].f32; }

}  // namespace flagfft
