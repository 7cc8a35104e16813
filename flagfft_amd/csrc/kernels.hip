// kernels.hip — gfx950 FFT kernels.
//
// Every transform is a sequence of LDS-resident Stockham transforms.  One
// workgroup holds C "units" (independent DFTs of length R) in LDS; a unit is a
// whole sequence (single-pass transforms) or one butterfly of a Stockham
// super-stage (multi-pass transforms of long sequences).  Global traffic goes
// through staged loads/stores whose thread order is chosen per pass so that a
// wavefront touches contiguous bytes; twiddles come from FP64-generated tables.
//
// Reference behaviour restated here (the reference emits Triton; no code is shared):
//   leaf FFT per program            python/flagfft_codegen/kernels.py:853-976
//   four-step column/row passes     python/flagfft_codegen/kernels.py:710-742
//   Bluestein prepare/pointwise/finalize  python/flagfft_codegen/jit_source.py:125-228
//   Rader prepare/pointwise/finalize      python/flagfft_codegen/jit_source.py:270-388
//   R2C/C2R expand/pack             python/flagfft_codegen/kernels.py:1098-1221
#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "colpass_body.hpp"
#include "dev_common.hpp"
#include "fft_args.hpp"
#include "kernels.hpp"

namespace flagfft {
namespace dev {

constexpr int kBatch = 8;

// Where the first stage reads and the last stage writes when they talk to HBM directly.
struct SrcDesc {
  const void *in;
  int64_t base, es, nlog;
  const void *chirp;
  int mode, conj, active;
};

template <typename T>
struct DstDesc {
  void *out;
  int64_t base, es, nlog;
  const void *chirp, *bhat;
  int mode, conj, active;
  T scale;
};

#define FLAGFFT_SRC_PARAMS \
  const void *s_in, int64_t s_base, int64_t s_es, int64_t s_nlog, const void *s_chirp, int s_mode, int s_conj, int s_active
#define FLAGFFT_DST_PARAMS                                                                                        \
  void *d_out, int64_t d_base, int64_t d_es, int64_t d_nlog, const void *d_chirp, const void *d_bhat, int d_mode, \
      int d_conj, int d_active, T d_scale
#define FLAGFFT_SRC_ARGS(S) S.in, S.base, S.es, S.nlog, S.chirp, S.mode, S.conj, S.active
#define FLAGFFT_DST_ARGS(D) D.out, D.base, D.es, D.nlog, D.chirp, D.bhat, D.mode, D.conj, D.active, D.scale

// One radix-RR Stockham stage over an LDS-resident sequence of length R.
// Prior product q; butterfly b: k = b mod q, inputs buf[b + j*R/RR] twiddled by
// w_R^{j k R/(q RR)}, outputs to buf[(b-k)RR + k + j q].
template <typename T, int RR, int E>
__device__ __noinline__ void lds_stage(LdsC<T> *buf, int R, int q, const FastDiv fdq, int twstep, int lt, int tpu,
                                       const GlbC<T> *__restrict__ tw) {
  constexpr int MAXB = (E + RR - 1) / RR;
  const int nb = R / RR;
  V2<T> v[MAXB][RR];
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
#pragma unroll
      for (int j = 0; j < RR; ++j) v[m][j] = ldc(buf + lpad(b + j * nb));
      if (q > 1) {
        const int k = b - static_cast<int>(fdiv(b, fdq)) * q;
        const int st = k * twstep;
#pragma unroll
        for (int j = 1; j < RR; ++j) v[m][j] = cmul(v[m][j], ldg(tw + j * st));
      }
      dft<RR>(v[m]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
      const int k = (q > 1) ? b - static_cast<int>(fdiv(b, fdq)) * q : 0;
      const int base = (b - k) * RR + k;
#pragma unroll
      for (int j = 0; j < RR; ++j) stc(buf + lpad(base + j * q), v[m][j]);
    }
  }
  __syncthreads();
}

// First stage (q = 1) reading its butterfly inputs straight from HBM.
template <typename T, int RR, int E>
__device__ __noinline__ void stage_first(LdsC<T> *buf, int R, FLAGFFT_SRC_PARAMS, int lt, int tpu) {
  constexpr int MAXB = (E + RR - 1) / RR;
  const int nb = R / RR;
  V2<T> v[MAXB][RR];
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
#pragma unroll
      for (int j = 0; j < RR; ++j)
        v[m][j] = s_active ? src_elem<T>(s_in, s_base, s_es, s_mode, s_nlog, s_chirp, s_conj, b + j * nb)
                           : V2<T>{0, 0};
    }
  }
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
      dft<RR>(v[m]);
#pragma unroll
      for (int j = 0; j < RR; ++j) stc(buf + lpad(b * RR + j), v[m][j]);
    }
  }
  __syncthreads();
}

// Last stage (q * RR == R, so k == b) writing its outputs straight to HBM.
template <typename T, int RR, int E>
__device__ __noinline__ void stage_last(LdsC<T> *buf, int R, int q, int twstep, const GlbC<T> *__restrict__ tw,
                                        FLAGFFT_DST_PARAMS, int lt, int tpu) {
  constexpr int MAXB = (E + RR - 1) / RR;
  const int nb = R / RR;
  V2<T> v[MAXB][RR];
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
#pragma unroll
      for (int j = 0; j < RR; ++j) v[m][j] = ldc(buf + lpad(b + j * nb));
      const int st = b * twstep;
#pragma unroll
      for (int j = 1; j < RR; ++j) v[m][j] = cmul(v[m][j], ldg(tw + j * st));
    }
  }
  if (!d_active) return;
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
      dft<RR>(v[m]);
#pragma unroll
      for (int j = 0; j < RR; ++j)
        dst_elem<T>(d_out, d_base, d_es, d_mode, d_nlog, d_chirp, d_bhat, d_conj, d_scale, b + j * q, v[m][j]);
    }
  }
}

// Single-stage transform (one radix: R <= 17 or R == 19): HBM -> registers -> HBM.
template <typename T, int RR, int E>
__device__ __noinline__ void stage_only(FLAGFFT_SRC_PARAMS, FLAGFFT_DST_PARAMS, int lt) {
  if (lt != 0 || !s_active) return;
  V2<T> v[RR];
#pragma unroll
  for (int j = 0; j < RR; ++j) v[j] = src_elem<T>(s_in, s_base, s_es, s_mode, s_nlog, s_chirp, s_conj, j);
  dft<RR>(v);
#pragma unroll
  for (int j = 0; j < RR; ++j) dst_elem<T>(d_out, d_base, d_es, d_mode, d_nlog, d_chirp, d_bhat, d_conj, d_scale, j, v[j]);
}

#define FLAGFFT_RADICES(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(19)

// Forward DFT of length R held in LDS (natural order in, natural order out).
// Every thread of the block must call this (it contains block barriers).
template <typename T, int E>
__device__ __forceinline__ void lds_fft(LdsC<T> *buf, int R, int nst, const int32_t *rad, const FastDiv *fdq,
                                        const int32_t *twstep, int lt, int tpu, const void *tw_small) {
  const GlbC<T> *tw = gc<T>(tw_small);
  int q = 1;
  for (int s = 0; s < nst; ++s) {
    const int r = rad[s];
    switch (r) {
#define FLAGFFT_MID(RR) \
  case RR: lds_stage<T, RR, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      FLAGFFT_RADICES(FLAGFFT_MID)
#undef FLAGFFT_MID
      default: break;  // the planner emits radices 2..17 and 19 only
    }
    q *= r;
  }
}

// Same transform with the first stage fed from HBM and the last stage writing HBM.
template <typename T, int E>
__device__ __forceinline__ void lds_fft_direct(LdsC<T> *buf, int R, int nst, const int32_t *rad, const FastDiv *fdq,
                                               const int32_t *twstep, int lt, int tpu, const void *tw_small,
                                               const SrcDesc &src, const DstDesc<T> &dst) {
  const GlbC<T> *tw = gc<T>(tw_small);
  if (nst == 0) {  // length 1
    if (lt == 0 && src.active)
      dst_elem<T>(dst.out, dst.base, dst.es, dst.mode, dst.nlog, dst.chirp, dst.bhat, dst.conj, dst.scale, 0,
                  src_elem<T>(src.in, src.base, src.es, src.mode, src.nlog, src.chirp, src.conj, 0));
    return;
  }
  if (nst == 1) {
    switch (rad[0]) {
#define FLAGFFT_ONLY(RR) \
  case RR: stage_only<T, RR, E>(FLAGFFT_SRC_ARGS(src), FLAGFFT_DST_ARGS(dst), lt); break;
      FLAGFFT_RADICES(FLAGFFT_ONLY)
#undef FLAGFFT_ONLY
      default: break;
    }
    return;
  }
  switch (rad[0]) {
#define FLAGFFT_FIRST(RR) \
  case RR: stage_first<T, RR, E>(buf, R, FLAGFFT_SRC_ARGS(src), lt, tpu); break;
    FLAGFFT_RADICES(FLAGFFT_FIRST)
#undef FLAGFFT_FIRST
    default: break;
  }
  int q = rad[0];
  for (int s = 1; s + 1 < nst; ++s) {
    const int r = rad[s];
    switch (r) {
#define FLAGFFT_MID(RR) \
  case RR: lds_stage<T, RR, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      FLAGFFT_RADICES(FLAGFFT_MID)
#undef FLAGFFT_MID
      default: break;
    }
    q *= r;
  }
  switch (rad[nst - 1]) {
#define FLAGFFT_LAST(RR) \
  case RR: stage_last<T, RR, E>(buf, R, q, twstep[nst - 1], tw, FLAGFFT_DST_ARGS(dst), lt, tpu); break;
    FLAGFFT_RADICES(FLAGFFT_LAST)
#undef FLAGFFT_LAST
    default: break;
  }
}

// Unit u -> (b, c, i) of a pass.
__device__ __forceinline__ void decompose_unit(const PassArgs &a, uint32_t uu, uint32_t &b, uint32_t &c, uint32_t &i) {
  if (a.col_fast) {
    const uint32_t q1 = fdiv(uu, a.fd_ncol);
    c = uu - q1 * a.ncol;
    b = fdiv(q1, a.fd_tcnt);
    i = q1 - b * a.tcnt;
  } else {
    const uint32_t q1 = fdiv(uu, a.fd_tcnt);
    i = uu - q1 * a.tcnt;
    b = fdiv(q1, a.fd_ncol);
    c = q1 - b * a.ncol;
  }
}

// ---------------------------------------------------------------------------
// Generic pass.  Direct form (contiguous single-pass sequences): first stage
// reads HBM, last stage writes HBM.  Staged form (strided sequences and
// multi-pass super-stages): batched coalesced copy into LDS with the source
// mode, conjugation and super-stage twiddle applied, LDS transform, batched
// coalesced copy out with the sink mode.
// ---------------------------------------------------------------------------
// Slot s and lane lt of thread tid for C slots of tpu threads; threads past
// C * tpu (a persistent workgroup running a smaller tile) get no butterflies
// but still reach every barrier.
__device__ __forceinline__ void slot_lane(int tid, int C, int tpu, int &s, int &lt) {
  s = tid / tpu;
  lt = tid - s * tpu;
  if (s >= C) {
    s = 0;
    lt = 1 << 24;
  }
}

template <typename T, int E>
__device__ __forceinline__ void pass_body(const PassArgs &a, int blk) {
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int64_t u0 = static_cast<int64_t>(blk) * a.C;
  const int LS = a.lds_stride;
  const T scale = static_cast<T>(a.scale);

  if (a.direct) {
    int s, lt;
    slot_lane(tid, a.C, a.tpu, s, lt);
    const int64_t u = u0 + s;
    const bool active = u < a.units && lt < a.tpu;
    uint32_t b = 0, c = 0, i = 0;
    if (active) decompose_unit(a, static_cast<uint32_t>(u), b, c, i);
    SrcDesc src{a.in, static_cast<int64_t>(b) * a.gi.bs + static_cast<int64_t>(c) * a.gi.cs, a.gi.es, a.nlog,
                a.chirp, a.src_mode, a.conj_in, active ? 1 : 0};
    DstDesc<T> dst{a.out, static_cast<int64_t>(b) * a.go.bs + static_cast<int64_t>(c) * a.go.cs, a.go.es, a.nlog,
                   a.chirp, a.bhat, a.dst_mode, a.conj_out, active ? 1 : 0, scale};
    lds_fft_direct<T, E>(lds + s * LS, a.R, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small, src, dst);
    return;
  }

  const int CR = a.C * a.R;
  const GlbC<T> *tw_lo = gc<T>(a.tw_lo);
  const GlbC<T> *tw_hi = gc<T>(a.tw_hi);
  const uint32_t tw_mask = (1u << a.tw_bits) - 1u;
  auto split = [&](int f, int slot_fast, int &s, int &e) __attribute__((always_inline)) {
    if (slot_fast) {
      e = static_cast<int>(fdiv(f, a.fd_C));
      s = f - e * a.C;
    } else {
      s = static_cast<int>(fdiv(f, a.fd_R));
      e = f - s * a.R;
    }
  };
  batched<T, kBatch>(
      CR, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int s, e;
        split(f, a.load_slot_fast, s, e);
        const int64_t u = u0 + s;
        V2<T> v{0, 0};
        if (u < a.units) {
          uint32_t b, c, i;
          decompose_unit(a, static_cast<uint32_t>(u), b, c, i);
          const int64_t base = static_cast<int64_t>(b) * a.gi.bs + static_cast<int64_t>(c) * a.gi.cs;
          const int64_t elog = static_cast<int64_t>(i) + static_cast<int64_t>(e) * a.tcnt;
          v = src_elem<T>(a.in, base, a.gi.es, a.src_mode, a.nlog, a.chirp, a.conj_in, elog);
          if (a.p > 1 && e > 0) {
            const uint32_t k = i - fdiv(i, a.fd_p) * a.p;
            const uint32_t mm = static_cast<uint32_t>(e) * k * static_cast<uint32_t>(a.twp_mult);
            v = cmul(v, cmul(ldg(tw_hi + (mm >> a.tw_bits)), ldg(tw_lo + (mm & tw_mask))));
          }
        }
        return v;
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int s, e;
        split(f, a.load_slot_fast, s, e);
        stc(lds + s * LS + lpad(e), v);
      });
  __syncthreads();

  {
    int s, lt;
    slot_lane(tid, a.C, a.tpu, s, lt);
    lds_fft<T, E>(lds + s * LS, a.R, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  }

  batched<T, kBatch>(
      CR, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int s, e;
        split(f, a.store_slot_fast, s, e);
        return ldc(lds + s * LS + lpad(e));
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int s, e;
        split(f, a.store_slot_fast, s, e);
        const int64_t u = u0 + s;
        if (u >= a.units) return;
        uint32_t b, c, i;
        decompose_unit(a, static_cast<uint32_t>(u), b, c, i);
        const uint32_t k = (a.p > 1) ? i - fdiv(i, a.fd_p) * a.p : 0u;
        const int64_t oidx = static_cast<int64_t>(i - k) * a.R + k + static_cast<int64_t>(e) * a.p;
        const int64_t base = static_cast<int64_t>(b) * a.go.bs + static_cast<int64_t>(c) * a.go.cs;
        dst_elem<T>(a.out, base, a.go.es, a.dst_mode, a.nlog, a.chirp, a.bhat, a.conj_out, scale, oidx, v);
      });
}

template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pass(const PassArgs a) {
  pass_body<T, E>(a, remap_block(blockIdx.x, a.nblocks, a.xcd_remap));
}

// Sequence q -> base offset of (c, b) for the given geometry.
__device__ __forceinline__ int64_t seq_base(uint32_t q, const FusedArgs &a, const Geometry &g) {
  const uint32_t b = fdiv(q, a.fd_ncol);
  const uint32_t c = q - b * a.ncol;
  return static_cast<int64_t>(b) * g.bs + static_cast<int64_t>(c) * g.cs;
}

__device__ __forceinline__ void slot_elem(int f, int slot_fast, const FusedArgs &a, int ext, const FastDiv &fd_ext,
                                          int &s, int &e) {
  if (slot_fast) {
    e = static_cast<int>(fdiv(f, a.fd_C));
    s = f - e * a.C;
  } else {
    s = static_cast<int>(fdiv(f, fd_ext));
    e = f - s * ext;
  }
}

// ---------------------------------------------------------------------------
// Rader, whole prime-length sequence in LDS (python/flagfft_codegen/jit_source.py:270-388):
//   a[q] = x[g^{-q}], A = FFT_m(a), X[0] = x0 + A[0], W = FFT_m(conj(A * bhat)),
//   X[g^q] = x0 + conj(W[q]).  bhat = FFT_m(w_p^{g^q}) / m.
// LDS per slot: [0, p) x / X, [p, p + m) a / A / W, one cell for A[0].
// ---------------------------------------------------------------------------
template <typename T, int E>
__device__ __forceinline__ void rader_body(const FusedArgs &a, int blk) {
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blk) * a.C;
  const int p = a.n, m = a.m, LS = a.lds_stride;
  const int xoff = 0, aoff = lpad(p) + 1, dcoff = aoff + lpad(m) + 1;
  const GlbR<int32_t> *idx = gr<int32_t>(a.aux);
  const GlbC<T> *bhat = gc<T>(a.bhat);

  batched<T, kBatch>(
      a.C * p, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int s, e;
        slot_elem(f, a.load_slot_fast, a, a.lext, a.fd_lext, s, e);
        const int64_t q = q0 + s;
        if (q >= a.nseqs) return V2<T>{0, 0};
        return src_elem<T>(a.in, seq_base(static_cast<uint32_t>(q), a, a.gi), a.gi.es, a.src_mode, p, nullptr,
                           a.conj_in, e);
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int s, e;
        slot_elem(f, a.load_slot_fast, a, a.lext, a.fd_lext, s, e);
        stc(lds + s * LS + xoff + lpad(e), v);
      });
  __syncthreads();
  int s, lt;
  slot_lane(tid, a.C, a.tpu, s, lt);
  LdsC<T> *X = lds + s * LS + xoff;
  LdsC<T> *A = lds + s * LS + aoff;
  for (int qq = lt; qq < m; qq += a.tpu) {
    const int src = idx[qq == 0 ? 0 : m - qq];
    stc(A + lpad(qq), ldc(X + lpad(src)));
  }
  __syncthreads();
  lds_fft<T, E>(A, m, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  for (int qq = lt; qq < m; qq += a.tpu) {
    const V2<T> av = ldc(A + lpad(qq));
    if (qq == 0) stc(lds + s * LS + dcoff, av);  // sum over x_j, j != 0
    stc(A + lpad(qq), cconj(cmul(av, ldg(bhat + qq))));
  }
  __syncthreads();
  lds_fft<T, E>(A, m, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  const V2<T> x0 = ldc(X);
  for (int qq = lt; qq < m; qq += a.tpu) stc(X + lpad(idx[qq]), cadd(x0, cconj(ldc(A + lpad(qq)))));
  __syncthreads();
  if (lt == 0) stc(X, cadd(x0, ldc(lds + s * LS + dcoff)));
  __syncthreads();
  const T scale = static_cast<T>(a.scale);
  batched<T, kBatch>(
      a.C * a.sext, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int ss, e;
        slot_elem(f, a.store_slot_fast, a, a.sext, a.fd_sext, ss, e);
        return ldc(lds + ss * LS + xoff + lpad(e));
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int ss, e;
        slot_elem(f, a.store_slot_fast, a, a.sext, a.fd_sext, ss, e);
        const int64_t q = q0 + ss;
        if (q >= a.nseqs) return;
        dst_elem<T>(a.out, seq_base(static_cast<uint32_t>(q), a, a.go), a.go.es, a.dst_mode, p, nullptr, nullptr,
                    a.conj_out, scale, e, v);
      });
}

template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_rader(const FusedArgs a) {
  rader_body<T, E>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Bluestein, whole convolution in LDS (src/exec/raw_nodes.cpp:153-204 order):
//   a = x * chirp (zero-padded to M), A = FFT_M(a), W = FFT_M(conj(A * bhat)),
//   y[k] = conj(W[k]) * chirp[k].  bhat = FFT_M(b) / M.
// ---------------------------------------------------------------------------
template <typename T, int E>
__device__ __forceinline__ void bluestein_body(const FusedArgs &a, int blk) {
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blk) * a.C;
  const int n = a.n, M = a.m, LS = a.lds_stride;
  const GlbC<T> *chirp = gc<T>(a.chirp);
  const GlbC<T> *bhat = gc<T>(a.bhat);
  int s, lt;
  slot_lane(tid, a.C, a.tpu, s, lt);
  LdsC<T> *A = lds + s * LS;
  for (int e = n + lt; e < M; e += a.tpu) stc(A + lpad(e), V2<T>{0, 0});
  batched<T, kBatch>(
      a.C * n, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int ss, e;
        slot_elem(f, a.load_slot_fast, a, a.lext, a.fd_lext, ss, e);
        const int64_t q = q0 + ss;
        if (q >= a.nseqs) return V2<T>{0, 0};
        const V2<T> v = src_elem<T>(a.in, seq_base(static_cast<uint32_t>(q), a, a.gi), a.gi.es, a.src_mode & 3, n,
                                    nullptr, a.conj_in, e);
        return cmul(v, ldg(chirp + e));
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int ss, e;
        slot_elem(f, a.load_slot_fast, a, a.lext, a.fd_lext, ss, e);
        stc(lds + ss * LS + lpad(e), v);
      });
  __syncthreads();
  lds_fft<T, E>(A, M, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  for (int qq = lt; qq < M; qq += a.tpu) stc(A + lpad(qq), cconj(cmul(ldc(A + lpad(qq)), ldg(bhat + qq))));
  __syncthreads();
  lds_fft<T, E>(A, M, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  const T scale = static_cast<T>(a.scale);
  batched<T, kBatch>(
      a.C * a.sext, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int ss, e;
        slot_elem(f, a.store_slot_fast, a, a.sext, a.fd_sext, ss, e);
        return cmul(cconj(ldc(lds + ss * LS + lpad(e))), ldg(chirp + e));
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int ss, e;
        slot_elem(f, a.store_slot_fast, a, a.sext, a.fd_sext, ss, e);
        const int64_t q = q0 + ss;
        if (q >= a.nseqs) return;
        dst_elem<T>(a.out, seq_base(static_cast<uint32_t>(q), a, a.go), a.go.es, a.dst_mode & 3, n, nullptr,
                    nullptr, a.conj_out, scale, e, v);
      });
}

template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_bluestein(const FusedArgs a) {
  bluestein_body<T, E>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Even-length R2C via an n/2-point complex transform, post-process fused:
//   z_j = x_2j + i x_2j+1, Z = FFT_h(z), X_k = E_k + w_n^k O_k with
//   E_k = (Z_k + conj Z_{h-k})/2, O_k = -i (Z_k - conj Z_{h-k})/2, k = 0..h (Z_h = Z_0).
// Input geometry counts reals (es must be 1 and bases even): z is read as V2.
// ---------------------------------------------------------------------------
template <typename T, int E>
__device__ __forceinline__ void r2c_even_body(const FusedArgs &a, int blk) {
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blk) * a.C;
  const int h = a.m, LS = a.lds_stride;
  const GlbC<T> *wn = gc<T>(a.aux);
  batched<T, kBatch>(
      a.C * h, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int s, e;
        slot_elem(f, a.load_slot_fast, a, h, a.fd_lext, s, e);
        const int64_t q = q0 + s;
        if (q >= a.nseqs) return V2<T>{0, 0};
        const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.gi);
        return ldg((const GlbC<T> *)(gr<T>(a.in) + base + 2 * static_cast<int64_t>(e)));
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int s, e;
        slot_elem(f, a.load_slot_fast, a, h, a.fd_lext, s, e);
        stc(lds + s * LS + lpad(e), v);
      });
  __syncthreads();
  int s, lt;
  slot_lane(tid, a.C, a.tpu, s, lt);
  lds_fft<T, E>(lds + s * LS, h, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  const T half = T(0.5);
  batched<T, kBatch>(
      a.C * (h + 1), tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int ss, k;
        slot_elem(f, a.store_slot_fast, a, h + 1, a.fd_sext, ss, k);
        const LdsC<T> *Z = lds + ss * LS;
        const V2<T> zk = ldc(Z + lpad(k == h ? 0 : k));
        const V2<T> zc = cconj(ldc(Z + lpad(k == 0 ? 0 : h - k)));
        const V2<T> ev{(zk.x + zc.x) * half, (zk.y + zc.y) * half};
        const V2<T> od = cmul_mi(V2<T>{(zk.x - zc.x) * half, (zk.y - zc.y) * half});
        V2<T> X = cadd(ev, cmul(ldg(wn + k), od));
        if (a.conj_out) X.y = -X.y;
        return X;
      },
      [&](int f, V2<T> X) __attribute__((always_inline)) {
        int ss, k;
        slot_elem(f, a.store_slot_fast, a, h + 1, a.fd_sext, ss, k);
        const int64_t q = q0 + ss;
        if (q >= a.nseqs) return;
        const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.go);
        stg(gc<T>(a.out) + base + static_cast<int64_t>(k) * a.go.es, X);
      });
}

template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_r2c_even(const FusedArgs a) {
  r2c_even_body<T, E>(a, blockIdx.x);
}

// ---------------------------------------------------------------------------
// Even-length C2R via an n/2-point transform, pre-process fused:
//   Z'_k = (X_k + conj X_{h-k}) + i w_n^{-k} (X_k - conj X_{h-k}), k < h,
//   Im X_0 and Im X_h dropped; z = IDFT_h(Z') (unnormalised) -> (x_2j, x_2j+1).
// Output geometry counts reals (es == 1, even bases).
// ---------------------------------------------------------------------------
template <typename T, int E>
__device__ __forceinline__ void c2r_even_body(const FusedArgs &a, int blk) {
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blk) * a.C;
  const int h = a.m, LS = a.lds_stride;
  const GlbC<T> *wn = gc<T>(a.aux);
  const GlbC<T> *X = gc<T>(a.in);
  batched<T, kBatch>(
      a.C * h, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int s, k;
        slot_elem(f, a.load_slot_fast, a, h, a.fd_lext, s, k);
        const int64_t q = q0 + s;
        if (q >= a.nseqs) return V2<T>{0, 0};
        const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.gi);
        V2<T> xk = ldg(X + base + static_cast<int64_t>(k) * a.gi.es);
        V2<T> xh = ldg(X + base + static_cast<int64_t>(h - k) * a.gi.es);
        if (k == 0) {
          xk.y = T(0);
          xh.y = T(0);
        }
        xh.y = -xh.y;  // conj X_{h-k}
        const V2<T> sum = cadd(xk, xh);
        const V2<T> dif = csub(xk, xh);
        // i * conj(w_n^k) * dif
        const V2<T> t = cmul_pi(cmul(dif, cconj(ldg(wn + k))));
        return cconj(cadd(sum, t));  // conj for the inverse-by-forward identity
      },
      [&](int f, V2<T> v) __attribute__((always_inline)) {
        int s, k;
        slot_elem(f, a.load_slot_fast, a, h, a.fd_lext, s, k);
        stc(lds + s * LS + lpad(k), v);
      });
  __syncthreads();
  int s, lt;
  slot_lane(tid, a.C, a.tpu, s, lt);
  lds_fft<T, E>(lds + s * LS, h, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, a.tw_small);
  const T scale = static_cast<T>(a.scale);
  batched<T, kBatch>(
      a.C * h, tid, nthr,
      [&](int f) __attribute__((always_inline)) {
        int ss, j;
        slot_elem(f, a.store_slot_fast, a, h, a.fd_sext, ss, j);
        V2<T> z = cconj(ldc(lds + ss * LS + lpad(j)));
        z.x *= scale;
        z.y *= scale;
        return z;
      },
      [&](int f, V2<T> z) __attribute__((always_inline)) {
        int ss, j;
        slot_elem(f, a.store_slot_fast, a, h, a.fd_sext, ss, j);
        const int64_t q = q0 + ss;
        if (q >= a.nseqs) return;
        const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.go);
        stg((GlbC<T> *)(gr<T>(a.out) + base + 2 * static_cast<int64_t>(j)), z);
      });
}

template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_c2r_even(const FusedArgs a) {
  c2r_even_body<T, E>(a, blockIdx.x);
}

template <typename T>
__device__ __forceinline__ void real_post_body(const RealPassArgs &a, int blk, int nblk);
template <typename T>
__device__ __forceinline__ void real_pre_body(const RealPassArgs &a, int blk, int nblk);

// ---------------------------------------------------------------------------
// Persistent mixed-length kernel.  A fixed grid of workgroups drains a queue of
// tiles from jobs of any length, type and precision (single-pass transforms,
// or one pass of multi-pass chains): each workgroup takes the next tile index
// from a global counter, finds its job by binary search over the jobs' first
// tiles and runs the same device body as the job's standalone kernel (so
// results are bit-identical to it).  Every workgroup exits once the counter
// passes the last tile.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kQueueThreads) void k_queue(const QueueArgs q) {
  __shared__ int item_s;
  for (;;) {
    __syncthreads();  // the previous tile is done with LDS and item_s
    if (threadIdx.x == 0) item_s = static_cast<int>(atomicAdd(q.counter, 1u));
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(item_s);  // uniform: the job's fields load as scalars
    if (item >= q.total) break;
    int lo = 0, hi = q.njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (q.jobs[mid].first <= item) lo = mid;
      else hi = mid - 1;
    }
    const QueueJob &j = q.jobs[lo];
    const int blk = item - j.first;
    switch (j.kind * 2 + j.dbl) {
      case kQueuePass * 2: pass_body<float, 16>(j.pa, blk); break;
      case kQueuePass * 2 + 1: pass_body<double, 8>(j.pa, blk); break;
      case kQueueRader * 2: rader_body<float, 16>(j.fa, blk); break;
      case kQueueRader * 2 + 1: rader_body<double, 8>(j.fa, blk); break;
      case kQueueBluestein * 2: bluestein_body<float, 16>(j.fa, blk); break;
      case kQueueBluestein * 2 + 1: bluestein_body<double, 8>(j.fa, blk); break;
      case kQueueR2CEven * 2: r2c_even_body<float, 16>(j.fa, blk); break;
      case kQueueR2CEven * 2 + 1: r2c_even_body<double, 8>(j.fa, blk); break;
      case kQueueC2REven * 2: c2r_even_body<float, 16>(j.fa, blk); break;
      case kQueueC2REven * 2 + 1: c2r_even_body<double, 8>(j.fa, blk); break;
      case kQueueRealPost * 2: real_post_body<float>(j.ra, blk, j.nblocks); break;
      case kQueueRealPost * 2 + 1: real_post_body<double>(j.ra, blk, j.nblocks); break;
      case kQueueRealPre * 2: real_pre_body<float>(j.ra, blk, j.nblocks); break;
      case kQueueRealPre * 2 + 1: real_pre_body<double>(j.ra, blk, j.nblocks); break;
      default: break;
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent kernels of the register-resident passes: one per instance (the
// queue variants of colpass.hip's catalogue), each draining the jobs of that
// instance with k_queue's tile loop, so that every instance keeps its own
// register count (a kernel holding several would need the registers of all).
// ---------------------------------------------------------------------------
template <typename T, int C, int RA, int RB, int NCH, int WPE>
__global__ __launch_bounds__(kQueueThreads) __attribute__((amdgpu_waves_per_eu(WPE))) void k_queue_col(const QueueArgs q) {
  __shared__ int item_s;
  for (;;) {
    __syncthreads();  // the previous tile is done with LDS and item_s
    if (threadIdx.x == 0) item_s = static_cast<int>(atomicAdd(q.counter, 1u));
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(item_s);  // uniform: the job's fields load as scalars
    if (item >= q.total) break;
    int lo = 0, hi = q.njobs - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (q.jobs[mid].first <= item) lo = mid;
      else hi = mid - 1;
    }
    // The job's arguments in scalar registers, as a standalone launch has its
    // kernel arguments: lane i loads dword i, readlane broadcasts each one
    // (a plain copy from the job array would sit in vector registers).
    constexpr int NW = static_cast<int>(sizeof(StageArgs) / 4);
    static_assert(NW <= 64 && sizeof(StageArgs) % 4 == 0, "StageArgs must fit one wavefront's lanes");
    union {
      StageArgs a;
      uint32_t w[NW];
    } u;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&q.jobs[lo].sa);
    const int lane = static_cast<int>(threadIdx.x & 63);
    const uint32_t mine = lane < NW ? src[lane] : 0u;
#pragma unroll
    for (int w = 0; w < NW; ++w) u.w[w] = __builtin_amdgcn_readlane(mine, w);
    const int first = __builtin_amdgcn_readfirstlane(q.jobs[lo].first);
    col_body<T, C, RA, RB, NCH, 1>(u.a, item - first);
  }
}

// ---------------------------------------------------------------------------
// Split-spectrum passes for even-length real transforms that run multi-pass:
//   post: Z (h complex, contiguous) -> X (h+1 bins);  pre: X -> conj(Z').
// ---------------------------------------------------------------------------
// Item g of a sequence set with `per` items per sequence: sequence q and
// item k (32-bit division when the count allows).
__device__ __forceinline__ void real_item(int64_t g, int64_t total, int per, int64_t &q, int &k) {
  if (total < (int64_t{1} << 31)) {
    const uint32_t gg = static_cast<uint32_t>(g), qq = gg / static_cast<uint32_t>(per);
    q = qq;
    k = static_cast<int>(gg - qq * static_cast<uint32_t>(per));
  } else {
    q = g / per;
    k = static_cast<int>(g - q * per);
  }
}

// Tile blk of nblk (256 threads each) strides over the items: item k <= h/2
// of a sequence reads Z_k and Z_{h-k} once and writes both X_k and X_{h-k}
// (X_k = E_k + w^k O_k, X_{h-k} = conj(E_k - w^k O_k), as the fused leaf).
template <typename T>
__device__ __forceinline__ void real_post_body(const RealPassArgs &a, int blk, int nblk) {
  const int h = a.n / 2;
  const int per = h / 2 + 1;
  const int64_t total = a.nseqs * per;
  const GlbC<T> *tw_lo = gc<T>(a.tw_lo);
  const GlbC<T> *tw_hi = gc<T>(a.tw_hi);
  const int mask = (1 << a.tw_bits) - 1;
  const T half = T(0.5);
  for (int64_t g = blk * static_cast<int64_t>(kQueueThreads) + threadIdx.x; g < total;
       g += static_cast<int64_t>(nblk) * kQueueThreads) {
    int64_t q;
    int k;
    real_item(g, total, per, q, k);
    const int64_t b = a.ncol == 1 ? q : q / a.ncol, c = q - b * a.ncol;
    const GlbC<T> *Z = gc<T>(a.in) + b * a.gi.bs + c * a.gi.cs;
    GlbC<T> *X = gc<T>(a.out) + b * a.go.bs + c * a.go.cs;
    const V2<T> zk = ldg(Z + k * a.gi.es);
    const V2<T> zc = cconj(ldg(Z + (k == 0 ? 0 : h - k) * a.gi.es));
    const V2<T> ev{(zk.x + zc.x) * half, (zk.y + zc.y) * half};
    const V2<T> od = cmul_mi(V2<T>{(zk.x - zc.x) * half, (zk.y - zc.y) * half});
    const V2<T> wo = cmul(cmul(ldg(tw_hi + (k >> a.tw_bits)), ldg(tw_lo + (k & mask))), od);
    stg(X + k * a.go.es, cadd(ev, wo));
    if (2 * k != h) stg(X + (h - k) * a.go.es, cconj(csub(ev, wo)));
  }
}

template <typename T>
__global__ __launch_bounds__(kQueueThreads) void k_real_post(const RealPassArgs a) {
  real_post_body<T>(a, blockIdx.x, gridDim.x);
}

// Item k <= h/2 reads X_k and X_{h-k} once and writes Z'_k and (k > 0)
// Z'_{h-k}: the roles of the two bins swap and conj(w^(h-k)) = -w^k.
template <typename T>
__device__ __forceinline__ void real_pre_body(const RealPassArgs &a, int blk, int nblk) {
  const int h = a.n / 2;
  const int per = h / 2 + 1;
  const int64_t total = a.nseqs * per;
  const GlbC<T> *tw_lo = gc<T>(a.tw_lo);
  const GlbC<T> *tw_hi = gc<T>(a.tw_hi);
  const int mask = (1 << a.tw_bits) - 1;
  for (int64_t g = blk * static_cast<int64_t>(kQueueThreads) + threadIdx.x; g < total;
       g += static_cast<int64_t>(nblk) * kQueueThreads) {
    int64_t q;
    int k;
    real_item(g, total, per, q, k);
    const int64_t b = a.ncol == 1 ? q : q / a.ncol, c = q - b * a.ncol;
    const GlbC<T> *Xs = gc<T>(a.in) + b * a.gi.bs + c * a.gi.cs;
    GlbC<T> *Zo = gc<T>(a.out) + b * a.go.bs + c * a.go.cs;
    V2<T> xk = ldg(Xs + k * a.gi.es);
    V2<T> xh = ldg(Xs + (h - k) * a.gi.es);
    if (k == 0) {
      xk.y = T(0);
      xh.y = T(0);
    }
    const V2<T> w = cmul(ldg(tw_hi + (k >> a.tw_bits)), ldg(tw_lo + (k & mask)));
    {
      const V2<T> cx = cconj(xh);
      const V2<T> t = cmul_pi(cmul(csub(xk, cx), cconj(w)));
      stg(Zo + k * a.go.es, cconj(cadd(cadd(xk, cx), t)));
    }
    if (k != 0 && 2 * k != h) {
      const V2<T> cx = cconj(xk);
      const V2<T> t = cmul_pi(cmul(csub(xh, cx), V2<T>{-w.x, -w.y}));
      stg(Zo + (h - k) * a.go.es, cconj(cadd(cadd(xh, cx), t)));
    }
  }
}

template <typename T>
__global__ __launch_bounds__(kQueueThreads) void k_real_pre(const RealPassArgs a) {
  real_pre_body<T>(a, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// FP64 table generation
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gen_roots(V2<T> *out, int64_t count, int64_t n, int64_t mult) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < count;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    // residues below n <= 2^32, so their product fits 64 unsigned bits
    const uint64_t un = static_cast<uint64_t>(n);
    const uint64_t r = (static_cast<uint64_t>(j) % un) * (static_cast<uint64_t>(mult) % un) % un;
    double s, c;
    sincospi(-2.0 * static_cast<double>(r) / static_cast<double>(n), &s, &c);
    out[j] = V2<T>{static_cast<T>(c), static_cast<T>(s)};
  }
}

template <typename T>
__global__ void k_gen_chirp(V2<T> *out, int64_t n) {
  for (int64_t k = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; k < n;
       k += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint64_t r = (static_cast<uint64_t>(k) * static_cast<uint64_t>(k)) % static_cast<uint64_t>(2 * n);
    double s, c;
    sincospi(-static_cast<double>(r) / static_cast<double>(n), &s, &c);
    out[k] = V2<T>{static_cast<T>(c), static_cast<T>(s)};
  }
}

// chirp[k] w_m^(h k) for h < p, k < n (FP64 phase, one rounding): the
// first-stage factors of a convolution split over p workgroups.
template <typename T>
__global__ void k_gen_chirp_split(V2<T> *out, int64_t n, int64_t m, int p) {
  const int64_t total = n * p;
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < total;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t h = j / n, k = j - h * n;
    const uint64_t r = (static_cast<uint64_t>(k) * static_cast<uint64_t>(k)) % static_cast<uint64_t>(2 * n);
    const int64_t t = (h * k) % m;
    // exp(-i pi r / n) exp(-2 i pi t / m), phase reduced to [-2, 0] half-turns
    double ph = -static_cast<double>(r) / static_cast<double>(n) - 2.0 * static_cast<double>(t) / static_cast<double>(m);
    if (ph < -2.0) ph += 2.0;
    double s, c;
    sincospi(ph, &s, &c);
    out[j] = V2<T>{static_cast<T>(c), static_cast<T>(s)};
  }
}

__global__ void k_gen_bluestein_b(V2<double> *out, int64_t n, int64_t m) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < m;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t k = -1;
    if (j < n) k = j;
    else if (m - j < n) k = m - j;
    if (k < 0) {
      out[j] = V2<double>{0.0, 0.0};
      continue;
    }
    const uint64_t r = (static_cast<uint64_t>(k) * static_cast<uint64_t>(k)) % static_cast<uint64_t>(2 * n);
    double s, c;
    sincospi(static_cast<double>(r) / static_cast<double>(n), &s, &c);
    out[j] = V2<double>{c, s};
  }
}

__global__ void k_gen_rader_b(V2<double> *out, const int32_t *idx, int64_t p) {
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < p - 1;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double s, c;
    sincospi(-2.0 * static_cast<double>(idx[q]) / static_cast<double>(p), &s, &c);
    out[q] = V2<double>{c, s};
  }
}

template <typename T>
__global__ void k_scale_cast(const V2<double> *in, V2<T> *out, int64_t count, double scale) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < count;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    out[j] = V2<T>{static_cast<T>(in[j].x * scale), static_cast<T>(in[j].y * scale)};
  }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
namespace {

dim3 grid_for(int64_t total, int threads) {
  int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  return dim3(static_cast<unsigned>(blocks));
}

}  // namespace

bool supported_ept(bool dbl, int ept) { return dbl ? (ept == 8 || ept == 16) : (ept == 8 || ept == 16 || ept == 32); }

hipError_t allow_lds(const void *f, size_t bytes) {
  if (bytes <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
}

#define FLAGFFT_KERNEL_EPT(KERNEL)                                                        \
  do {                                                                                  \
    if (dbl) {                                                                          \
      if (ept == 8) return reinterpret_cast<const void *>(dev::KERNEL<double, 8>);      \
      if (ept == 16) return reinterpret_cast<const void *>(dev::KERNEL<double, 16>);    \
    } else {                                                                            \
      if (ept == 8) return reinterpret_cast<const void *>(dev::KERNEL<float, 8>);       \
      if (ept == 16) return reinterpret_cast<const void *>(dev::KERNEL<float, 16>);     \
      if (ept == 32) return reinterpret_cast<const void *>(dev::KERNEL<float, 32>);     \
    }                                                                                   \
    return nullptr;                                                                     \
  } while (0)

const void *pass_kernel(bool dbl, int ept) { FLAGFFT_KERNEL_EPT(k_pass); }
const void *rader_kernel(bool dbl, int ept) { FLAGFFT_KERNEL_EPT(k_rader); }
const void *bluestein_kernel(bool dbl, int ept) { FLAGFFT_KERNEL_EPT(k_bluestein); }
const void *r2c_even_kernel(bool dbl, int ept) { FLAGFFT_KERNEL_EPT(k_r2c_even); }
const void *c2r_even_kernel(bool dbl, int ept) { FLAGFFT_KERNEL_EPT(k_c2r_even); }

const void *queue_kernel() { return reinterpret_cast<const void *>(dev::k_queue); }
// Indexed by QueueJob::variant (colpass.hip's queue variants); null where unused.
const void *queue_col_kernel(int variant) {
  static const void *const k[] = {
      reinterpret_cast<const void *>(dev::k_queue_col<float, 8, 32, 32, 2, 4>),
      reinterpret_cast<const void *>(dev::k_queue_col<float, 16, 32, 16, 1, 2>),
      reinterpret_cast<const void *>(dev::k_queue_col<float, 16, 16, 16, 1, 4>),
      reinterpret_cast<const void *>(dev::k_queue_col<float, 32, 16, 8, 1, 3>),
      reinterpret_cast<const void *>(dev::k_queue_col<float, 32, 8, 8, 1, 4>),
      reinterpret_cast<const void *>(dev::k_queue_col<float, 64, 8, 4, 1, 4>),
      reinterpret_cast<const void *>(dev::k_queue_col<float, 64, 4, 4, 1, 4>),
      reinterpret_cast<const void *>(dev::k_queue_col<double, 16, 16, 16, 1, 2>),
      nullptr,
      reinterpret_cast<const void *>(dev::k_queue_col<double, 32, 8, 8, 1, 4>),
  };
  return variant >= 0 && variant < static_cast<int>(sizeof(k) / sizeof(k[0])) ? k[variant] : nullptr;
}

const void *real_post_kernel(bool dbl) {
  return dbl ? reinterpret_cast<const void *>(dev::k_real_post<double>)
             : reinterpret_cast<const void *>(dev::k_real_post<float>);
}
const void *real_pre_kernel(bool dbl) {
  return dbl ? reinterpret_cast<const void *>(dev::k_real_pre<double>)
             : reinterpret_cast<const void *>(dev::k_real_pre<float>);
}
// (one item per pair of bins k, h - k: h / 2 + 1 per sequence, h = n / 2)
LaunchShape real_post_shape(const RealPassArgs &a) { return LaunchShape{grid_for(a.nseqs * (a.n / 4 + 1), 256), dim3(256), 0}; }
LaunchShape real_pre_shape(const RealPassArgs &a) { return LaunchShape{grid_for(a.nseqs * (a.n / 4 + 1), 256), dim3(256), 0}; }

hipError_t gen_roots(void *out, bool dbl, int64_t count, int64_t n, int64_t mult, hipStream_t st) {
  dim3 g = grid_for(count, 256);
  if (dbl) hipLaunchKernelGGL(dev::k_gen_roots<double>, g, dim3(256), 0, st, static_cast<dev::V2<double> *>(out), count, n, mult);
  else hipLaunchKernelGGL(dev::k_gen_roots<float>, g, dim3(256), 0, st, static_cast<dev::V2<float> *>(out), count, n, mult);
  return hipGetLastError();
}

hipError_t gen_chirp_split(void *out, bool dbl, int64_t n, int64_t m, int p, hipStream_t st) {
  const dim3 g = grid_for(n * p, 256);
  if (dbl)
    hipLaunchKernelGGL(dev::k_gen_chirp_split<double>, g, dim3(256), 0, st, static_cast<dev::V2<double> *>(out), n, m, p);
  else
    hipLaunchKernelGGL(dev::k_gen_chirp_split<float>, g, dim3(256), 0, st, static_cast<dev::V2<float> *>(out), n, m, p);
  return hipGetLastError();
}

hipError_t gen_chirp(void *out, bool dbl, int64_t n, hipStream_t st) {
  dim3 g = grid_for(n, 256);
  if (dbl) hipLaunchKernelGGL(dev::k_gen_chirp<double>, g, dim3(256), 0, st, static_cast<dev::V2<double> *>(out), n);
  else hipLaunchKernelGGL(dev::k_gen_chirp<float>, g, dim3(256), 0, st, static_cast<dev::V2<float> *>(out), n);
  return hipGetLastError();
}

hipError_t gen_bluestein_b(double2 *out, int64_t n, int64_t m, hipStream_t st) {
  hipLaunchKernelGGL(dev::k_gen_bluestein_b, grid_for(m, 256), dim3(256), 0, st, reinterpret_cast<dev::V2<double> *>(out), n, m);
  return hipGetLastError();
}

hipError_t gen_rader_b(double2 *out, const int32_t *idx, int64_t p, hipStream_t st) {
  hipLaunchKernelGGL(dev::k_gen_rader_b, grid_for(p - 1, 256), dim3(256), 0, st, reinterpret_cast<dev::V2<double> *>(out), idx, p);
  return hipGetLastError();
}

hipError_t scale_cast(const double2 *in, void *out, bool dbl, int64_t count, double scale, hipStream_t st) {
  dim3 g = grid_for(count, 256);
  if (dbl) hipLaunchKernelGGL(dev::k_scale_cast<double>, g, dim3(256), 0, st, reinterpret_cast<const dev::V2<double> *>(in), static_cast<dev::V2<double> *>(out), count, scale);
  else hipLaunchKernelGGL(dev::k_scale_cast<float>, g, dim3(256), 0, st, reinterpret_cast<const dev::V2<double> *>(in), static_cast<dev::V2<float> *>(out), count, scale);
  return hipGetLastError();
}

}  // namespace flagfft
