// stage.hip — compile-time specialised Stockham super-stages (gfx950).
//
// One launch is one super-stage of a multi-pass transform (the reference's
// four-step column / row passes, python/flagfft_codegen/kernels.py:710-742,
// restated as Stockham passes with fused twiddles).
#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"
#include "fft_args.hpp"
#include "kernels.hpp"
#include "spec_common.hpp"

namespace flagfft {
namespace dev {

// Super-stage modes.
//   kStagePass         one pass of a chain: unit u = b * tcnt + i of sequence b
//                      reads x[i + e tcnt] (e < R), multiplies by
//                      w_{pR}^{e (i mod p)}, transforms and writes
//                      y[(i - k) R + k + e p], k = i mod p; Bluestein chains
//                      use the chirp source (first pass) and the pointwise /
//                      chirp sinks.  Column sequences (ncol > 1) run the same
//                      index map over column-interleaved units.
//   kStageConvolution  middle of a two-pass Bluestein convolution M = R0 * R:
//                      the forward chain's second pass (p = tcnt = R0), the
//                      pointwise conj(X * bhat), and the inverse chain's first
//                      pass (its super-radices run in the opposite order, so it
//                      reads the same R elements i + e R0), written to
//                      y[i R + e] as one contiguous block per workgroup.
enum StageMode : int { kStagePass = 0, kStageConvolution = 1 };

// C adjacent units per workgroup; lane -> (slot = tid % C, t = tid / C) so a
// wavefront's HBM reads are C-element row segments.  With block_store the C
// units' outputs form one contiguous block (first pass of contiguous
// sequences, and every convolution pass) and leave through LDS as one
// coalesced copy; otherwise each slot stores directly.
template <typename T, int E, int C, int Mode, int... Rs>
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
  // first pass: p is one per column; k: position inside the prior product
  // (column-interleaved), kt = k / ncol its per-column twiddle exponent
  const bool first = a.p == a.ncol;
  const uint32_t k = i - fdiv(i, a.fd_p) * static_cast<uint32_t>(a.p);
  const uint32_t kt = first ? 0u : fdiv(k, a.fd_ncol);
  const int64_t tcnt = a.tcnt;
  const GlbC<T> *src = gc<T>(a.in) + static_cast<int64_t>(bq) * a.bs_in;
  const GlbC<T> *tws = gc<T>(a.tw_small);
  const GlbC<T> *tw_lo = gc<T>(a.tw_lo);
  const GlbC<T> *tw_hi = gc<T>(a.tw_hi);
  const GlbC<T> *chirp = gc<T>(a.chirp);
  const uint32_t tw_mask = (1u << a.tw_bits) - 1u;
  const uint32_t kmul = kt * static_cast<uint32_t>(a.twp_mult);
  const T scale = static_cast<T>(a.scale);
  const bool chirp_in = (a.src_mode & kSrcChirpFlag) != 0;
  LdsC<T> *buf = lds + sl * LS;
  // inner-stage twiddles by binary powers of one base twiddle per butterfly
  V2<T> twr[1];
  // inter-pass twiddle of element e = b + j NB of first-stage butterfly b:
  // w^(e kmul) = w^(b kmul) (w^(NB kmul))^j, two-level table lookups of the
  // base and the step, powers of the step by squaring
  constexpr int NB0 = R / RS::kR[0];
  auto tw2 = [&](uint32_t mm) __attribute__((always_inline)) {
    return cmul(ldg(tw_hi + (mm >> a.tw_bits)), ldg(tw_lo + (mm & tw_mask)));
  };
  const V2<T> step = first ? V2<T>{1, 0} : tw2(static_cast<uint32_t>(NB0) * kmul);
  auto pre0 = [&](int b, V2<T> *x) __attribute__((always_inline)) {
    if (first) return;
    constexpr int R0 = RS::kR[0];
    const V2<T> base = tw2(static_cast<uint32_t>(b) * kmul);
    V2<T> pw[4];
    pw[0] = step;
    pw[1] = cmul(pw[0], pw[0]);
    pw[2] = cmul(pw[1], pw[1]);
    pw[3] = cmul(pw[2], pw[2]);
    x[0] = cmul(x[0], base);
    sfor<R0>([&](auto J) __attribute__((always_inline)) {
      constexpr int jj = decltype(J)::value;
      if constexpr (jj > 0) x[jj] = cmul(x[jj], cmul(base, tw_power<T, jj>(pw)));
    });
  };

  auto load_global = [&](int e) __attribute__((always_inline)) {
    const int64_t elog = static_cast<int64_t>(i) + e * tcnt;
    V2<T> v{0, 0};
    if (active && !(chirp_in && elog >= a.nlog)) {
      v = ldg(src + elog);
      if (a.conj_in) v.y = -v.y;
      if (chirp_in) v = cmul(v, ldg(chirp + elog));
    }
    return v;
  };
  auto to_lds = [&](int idx, V2<T> v) __attribute__((always_inline)) { stc(buf + lpad(idx), v); };

  bool block_store = a.block_store != 0;
  if constexpr (Mode == kStageConvolution) {
    block_store = true;
    run_stages<T, E, TPU, RS, false, 5, true>(buf, t, twr, tws, true, load_global, to_lds, pre0);
    __syncthreads();
    const GlbC<T> *bhat = gc<T>(a.bhat);
    run_stages<T, E, TPU, RS, true, 5, true>(
        buf, t, twr, tws, true,
        [&](int e) __attribute__((always_inline)) {
          return cconj(cmul(ldc(buf + lpad(e)), ldg(bhat + (i + static_cast<int64_t>(e) * a.p))));
        },
        to_lds);
  } else if (block_store) {
    run_stages<T, E, TPU, RS, false, 5, true>(buf, t, twr, tws, true, load_global, to_lds, pre0);
  } else {
    run_stages<T, E, TPU, RS, false, 5, true>(
        buf, t, twr, tws, false, load_global,
        [&](int idx, V2<T> v) __attribute__((always_inline)) {
          if (!active) return;
          const int64_t oidx = (i - k) * static_cast<int64_t>(R) + k + static_cast<int64_t>(idx) * a.p;
          dst_elem<T>(a.out, static_cast<int64_t>(bq) * a.bs_out, 1, a.dst_mode, a.nlog, a.chirp, a.bhat,
                      a.conj_out, scale, oidx, v);
        },
        pre0);
  }
  if (block_store) {
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
  const void *f[2][2];  // [FP64][StageMode]
  bool fast;            // runs near the other specialised passes' rate (see stage_kernel_fast)
};

#define FLAGFFT_STAGE_FNS(C32, E32, C64, E64, ...)                                                     \
  {                                                                                                    \
    {(const void *)dev::k_stage<float, E32, C32, dev::kStagePass, __VA_ARGS__>,                        \
     (const void *)dev::k_stage<float, E32, C32, dev::kStageConvolution, __VA_ARGS__>},                \
    {                                                                                                  \
      (const void *)dev::k_stage<double, E64, C64, dev::kStagePass, __VA_ARGS__>,                      \
          (const void *)dev::k_stage<double, E64, C64, dev::kStageConvolution, __VA_ARGS__>            \
    }                                                                                                  \
  }
#define FLAGFFT_STAGE2(C32, E32, C64, E64, R0, R1) \
  {2, {R0, R1, 0}, C32, E32, C64, E64, FLAGFFT_STAGE_FNS(C32, E32, C64, E64, R0, R1), true}
#define FLAGFFT_STAGE3(C32, E32, C64, E64, R0, R1, R2) \
  {3, {R0, R1, R2}, C32, E32, C64, E64, FLAGFFT_STAGE_FNS(C32, E32, C64, E64, R0, R1, R2), true}
#define FLAGFFT_STAGE2_SLOW(C32, E32, C64, E64, R0, R1) \
  {2, {R0, R1, 0}, C32, E32, C64, E64, FLAGFFT_STAGE_FNS(C32, E32, C64, E64, R0, R1), false}
#define FLAGFFT_STAGE3_SLOW(C32, E32, C64, E64, R0, R1, R2) \
  {3, {R0, R1, R2}, C32, E32, C64, E64, FLAGFFT_STAGE_FNS(C32, E32, C64, E64, R0, R1, R2), false}

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
    // 5- and 10-smooth passes (10^k, 5^k lengths): the run-time-radix pass
    // reads them at 0.4-0.9 TB/s, these at 0.9-1.9 (MI355X, 8-32 MB
    // transforms), well below the power-of-two passes
    FLAGFFT_STAGE2_SLOW(64, 16, 32, 16, 5, 5),       // 25
    FLAGFFT_STAGE2_SLOW(32, 16, 16, 16, 10, 10),     // 100
    FLAGFFT_STAGE3_SLOW(32, 16, 16, 16, 5, 5, 5),    // 125
    FLAGFFT_STAGE3_SLOW(16, 32, 8, 16, 10, 10, 10),  // 1000
    // wider tiles (longer contiguous runs per strided read), picked by the
    // planner's units-per-workgroup choice
    FLAGFFT_STAGE2(64, 16, 32, 16, 16, 8),      // 128
    FLAGFFT_STAGE2(32, 16, 16, 16, 16, 10),     // 160
    FLAGFFT_STAGE2(32, 16, 16, 16, 16, 16),     // 256
    FLAGFFT_STAGE3(32, 32, 16, 16, 16, 16, 2),  // 512
    FLAGFFT_STAGE3(16, 32, 8, 16, 16, 16, 4),   // 1024
    FLAGFFT_STAGE3(8, 32, 4, 32, 16, 16, 8),    // 2048
};

}  // namespace

int stage_kernel_id(const std::vector<int> &radices, bool dbl, int units) {
  const int n = static_cast<int>(sizeof(kStageTable) / sizeof(kStageTable[0]));
  int first = -1;
  for (int i = 0; i < n; ++i) {
    const StageEntry &e = kStageTable[i];
    if (static_cast<int>(radices.size()) != e.nrad) continue;
    bool same = true;
    for (int j = 0; j < e.nrad; ++j) same = same && radices[static_cast<size_t>(j)] == e.rad[j];
    if (!same) continue;
    if (units <= 0 || (dbl ? e.c64 : e.c32) == units) return i;
    if (first < 0) first = i;
  }
  return units > 0 ? -1 : first;
}

int stage_kernel_id(const std::vector<int> &radices) { return stage_kernel_id(radices, false, 0); }

bool stage_kernel_fast(int id) { return kStageTable[id].fast; }

void stage_kernel_shape(int id, bool dbl, int &units_per_group, int &ept) {
  units_per_group = dbl ? kStageTable[id].c64 : kStageTable[id].c32;
  ept = dbl ? kStageTable[id].e64 : kStageTable[id].e32;
}

const void *stage_kernel(bool dbl, int id, int mode) { return kStageTable[id].f[dbl ? 1 : 0][mode]; }

}  // namespace flagfft
