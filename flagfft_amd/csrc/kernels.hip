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
#include "fft_args.hpp"
#include "kernels.hpp"

namespace flagfft {
namespace dev {

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
  uint32_t t = __umulhi(n, f.mul);
  return (t + n) >> f.shift;
}

// LDS index padding: one spare slot every 32 complex elements.
__device__ __forceinline__ int lpad(int e) { return e + (e >> 5); }

__device__ __forceinline__ int remap_block(int b, int nblocks, int enable) {
  if (!enable || (nblocks & 7) != 0) return b;
  // blocks are dispatched round-robin over the 8 XCDs: keep consecutive tiles on one XCD
  return (b & 7) * (nblocks >> 3) + (b >> 3);
}

// Logical input element e: base source, then conjugation (inverse by the
// conj-forward-conj identity), then the Bluestein chirp.
template <typename T>
__device__ __forceinline__ V2<T> src_elem(const void *in, int64_t base, int64_t es, int mode, int64_t n,
                                          const V2<T> *chirp, int conj_in, int64_t e) {
  if ((mode & kSrcChirpFlag) && e >= n) return V2<T>{0, 0};
  const int bm = mode & 3;
  V2<T> v;
  if (bm == kSrcComplex) {
    v = static_cast<const V2<T> *>(in)[base + e * es];
  } else if (bm == kSrcReal) {
    v = V2<T>{static_cast<const T *>(in)[base + e * es], T(0)};
  } else {
    const V2<T> *X = static_cast<const V2<T> *>(in);
    const int64_t h = n >> 1;
    if (e <= h) {
      v = X[base + e * es];
      if (e == 0 || 2 * e == n) v.y = T(0);
    } else {
      v = X[base + (n - e) * es];
      v.y = -v.y;
    }
  }
  if (conj_in) v.y = -v.y;
  if (mode & kSrcChirpFlag) v = cmul(v, chirp[e]);
  return v;
}

template <typename T>
__device__ __forceinline__ void dst_elem(void *out, int64_t base, int64_t es, int mode, int64_t n,
                                         const V2<T> *chirp, const V2<T> *bhat, int conj_out, T scale, int64_t e,
                                         V2<T> v) {
  if (mode & kDstPointwiseFlag) {
    v = cconj(cmul(v, bhat[e]));
    static_cast<V2<T> *>(out)[base + e * es] = v;
    return;
  }
  if (mode & kDstChirpFlag) {
    if (e >= n) return;
    v = cmul(cconj(v), chirp[e]);
  }
  if (conj_out) v.y = -v.y;
  if (scale != T(1)) {
    v.x *= scale;
    v.y *= scale;
  }
  const int bm = mode & 3;
  if (bm == kDstComplex) {
    static_cast<V2<T> *>(out)[base + e * es] = v;
  } else if (bm == kDstHalf) {
    if (2 * e <= n) static_cast<V2<T> *>(out)[base + e * es] = v;
  } else {
    static_cast<T *>(out)[base + e * es] = v.x;
  }
}

// One radix-RR Stockham stage over an LDS-resident sequence of length R.
// Prior product q; butterfly b: k = b mod q, inputs buf[b + j*R/RR] twiddled by
// w_R^{j k R/(q RR)}, outputs to buf[(b-k)RR + k + j q].
template <typename T, int RR, int E>
__device__ __noinline__ void lds_stage(V2<T> *buf, int R, int q, const FastDiv fdq, int twstep, int lt, int tpu,
                                       const V2<T> *__restrict__ tw) {
  constexpr int MAXB = (E + RR - 1) / RR;
  const int nb = R / RR;
  V2<T> v[MAXB][RR];
#pragma unroll
  for (int m = 0; m < MAXB; ++m) {
    const int b = lt + m * tpu;
    if (b < nb) {
#pragma unroll
      for (int j = 0; j < RR; ++j) v[m][j] = buf[lpad(b + j * nb)];
      if (q > 1) {
        const int k = b - static_cast<int>(fdiv(b, fdq)) * q;
        const int st = k * twstep;
#pragma unroll
        for (int j = 1; j < RR; ++j) v[m][j] = cmul(v[m][j], tw[j * st]);
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
      for (int j = 0; j < RR; ++j) buf[lpad(base + j * q)] = v[m][j];
    }
  }
  __syncthreads();
}

// Forward DFT of length R held in LDS (natural order in, natural order out).
// Every thread of the block must call this (it contains block barriers).
template <typename T, int E>
__device__ __forceinline__ void lds_fft(V2<T> *buf, int R, int nst, const int32_t *rad, const FastDiv *fdq, const int32_t *twstep,
                        int lt, int tpu, const V2<T> *__restrict__ tw) {
  int q = 1;
  for (int s = 0; s < nst; ++s) {
    const int r = rad[s];
    switch (r) {
      case 2: lds_stage<T, 2, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 3: lds_stage<T, 3, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 4: lds_stage<T, 4, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 5: lds_stage<T, 5, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 6: lds_stage<T, 6, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 7: lds_stage<T, 7, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 8: lds_stage<T, 8, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 9: lds_stage<T, 9, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 10: lds_stage<T, 10, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 11: lds_stage<T, 11, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 12: lds_stage<T, 12, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 13: lds_stage<T, 13, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 14: lds_stage<T, 14, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 15: lds_stage<T, 15, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      case 16: lds_stage<T, 16, E>(buf, R, q, fdq[s], twstep[s], lt, tpu, tw); break;
      default: break;  // the planner never emits other radices
    }
    q *= r;
  }
}

// ---------------------------------------------------------------------------
// Generic pass: staged load (+source mode, conj, super-stage twiddle) ->
// LDS transform -> staged store (+sink mode).
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_pass(const PassArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V2<T> *lds = reinterpret_cast<V2<T> *>(smem_raw);
  const int tid = threadIdx.x;
  const int nthr = blockDim.x;
  const int blk = remap_block(blockIdx.x, a.nblocks, a.xcd_remap);
  const int64_t u0 = static_cast<int64_t>(blk) * a.C;
  const int CR = a.C * a.R;
  const int LS = a.lds_stride;
  const V2<T> *tw_lo = static_cast<const V2<T> *>(a.tw_lo);
  const V2<T> *tw_hi = static_cast<const V2<T> *>(a.tw_hi);
  const V2<T> *chirp = static_cast<const V2<T> *>(a.chirp);
  const uint32_t tw_mask = (1u << a.tw_bits) - 1u;

  for (int f = tid; f < CR; f += nthr) {
    int s, e;
    if (a.load_slot_fast) {
      e = static_cast<int>(fdiv(f, a.fd_C));
      s = f - e * a.C;
    } else {
      s = static_cast<int>(fdiv(f, a.fd_R));
      e = f - s * a.R;
    }
    const int64_t u = u0 + s;
    V2<T> v{0, 0};
    if (u < a.units) {
      const uint32_t uu = static_cast<uint32_t>(u);
      uint32_t i, c, b;
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
      const int64_t base = static_cast<int64_t>(b) * a.gi.bs + static_cast<int64_t>(c) * a.gi.cs;
      const int64_t elog = static_cast<int64_t>(i) + static_cast<int64_t>(e) * a.tcnt;
      v = src_elem<T>(a.in, base, a.gi.es, a.src_mode, a.nlog, chirp, a.conj_in, elog);
      if (a.p > 1 && e > 0) {
        const uint32_t k = i - fdiv(i, a.fd_p) * a.p;
        const uint32_t mm = static_cast<uint32_t>(e) * k * static_cast<uint32_t>(a.twp_mult);
        v = cmul(v, cmul(tw_hi[mm >> a.tw_bits], tw_lo[mm & tw_mask]));
      }
    }
    lds[s * LS + lpad(e)] = v;
  }
  __syncthreads();

  {
    const int s = tid / a.tpu;
    const int lt = tid - s * a.tpu;
    lds_fft<T, E>(lds + s * LS, a.R, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu,
                  static_cast<const V2<T> *>(a.tw_small));
  }

  const V2<T> *bhat = static_cast<const V2<T> *>(a.bhat);
  const T scale = static_cast<T>(a.scale);
  for (int f = tid; f < CR; f += nthr) {
    int s, e;
    if (a.store_slot_fast) {
      e = static_cast<int>(fdiv(f, a.fd_C));
      s = f - e * a.C;
    } else {
      s = static_cast<int>(fdiv(f, a.fd_R));
      e = f - s * a.R;
    }
    const int64_t u = u0 + s;
    if (u >= a.units) continue;
    const uint32_t uu = static_cast<uint32_t>(u);
    uint32_t i, c, b;
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
    const uint32_t k = (a.p > 1) ? i - fdiv(i, a.fd_p) * a.p : 0u;
    const int64_t oidx = static_cast<int64_t>(i - k) * a.R + k + static_cast<int64_t>(e) * a.p;
    const int64_t base = static_cast<int64_t>(b) * a.go.bs + static_cast<int64_t>(c) * a.go.cs;
    dst_elem<T>(a.out, base, a.go.es, a.dst_mode, a.nlog, chirp, bhat, a.conj_out, scale, oidx,
                lds[s * LS + lpad(e)]);
  }
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
__global__ __launch_bounds__(kMaxBlockThreads) void k_rader(const FusedArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V2<T> *lds = reinterpret_cast<V2<T> *>(smem_raw);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * a.C;
  const int p = a.n, m = a.m, LS = a.lds_stride;
  const int xoff = 0, aoff = lpad(p) + 1, dcoff = aoff + lpad(m) + 1;
  const int32_t *idx = static_cast<const int32_t *>(a.aux);
  const V2<T> *bhat = static_cast<const V2<T> *>(a.bhat);

  for (int f = tid; f < a.C * p; f += nthr) {
    int s, e;
    slot_elem(f, a.load_slot_fast, a, a.lext, a.fd_lext, s, e);
    const int64_t q = q0 + s;
    V2<T> v{0, 0};
    if (q < a.nseqs) {
      v = src_elem<T>(a.in, seq_base(static_cast<uint32_t>(q), a, a.gi), a.gi.es, a.src_mode, p, nullptr, a.conj_in,
                      e);
    }
    lds[s * LS + xoff + lpad(e)] = v;
  }
  __syncthreads();
  const int s = tid / a.tpu;
  const int lt = tid - s * a.tpu;
  V2<T> *X = lds + s * LS + xoff;
  V2<T> *A = lds + s * LS + aoff;
  for (int qq = lt; qq < m; qq += a.tpu) {
    const int src = idx[qq == 0 ? 0 : m - qq];
    A[lpad(qq)] = X[lpad(src)];
  }
  __syncthreads();
  lds_fft<T, E>(A, m, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, static_cast<const V2<T> *>(a.tw_small));
  for (int qq = lt; qq < m; qq += a.tpu) {
    const V2<T> av = A[lpad(qq)];
    if (qq == 0) lds[s * LS + dcoff] = av;  // sum over x_j, j != 0
    A[lpad(qq)] = cconj(cmul(av, bhat[qq]));
  }
  __syncthreads();
  lds_fft<T, E>(A, m, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, static_cast<const V2<T> *>(a.tw_small));
  const V2<T> x0 = X[0];
  for (int qq = lt; qq < m; qq += a.tpu) X[lpad(idx[qq])] = cadd(x0, cconj(A[lpad(qq)]));
  __syncthreads();
  if (lt == 0) X[0] = cadd(x0, lds[s * LS + dcoff]);
  __syncthreads();
  const T scale = static_cast<T>(a.scale);
  for (int f = tid; f < a.C * a.sext; f += nthr) {
    int ss, e;
    slot_elem(f, a.store_slot_fast, a, a.sext, a.fd_sext, ss, e);
    const int64_t q = q0 + ss;
    if (q >= a.nseqs) continue;
    dst_elem<T>(a.out, seq_base(static_cast<uint32_t>(q), a, a.go), a.go.es, a.dst_mode, p, nullptr, nullptr,
                a.conj_out, scale, e, lds[ss * LS + xoff + lpad(e)]);
  }
}

// ---------------------------------------------------------------------------
// Bluestein, whole convolution in LDS (src/exec/raw_nodes.cpp:153-204 order):
//   a = x * chirp (zero-padded to M), A = FFT_M(a), W = FFT_M(conj(A * bhat)),
//   y[k] = conj(W[k]) * chirp[k].  bhat = FFT_M(b) / M.
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_bluestein(const FusedArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V2<T> *lds = reinterpret_cast<V2<T> *>(smem_raw);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * a.C;
  const int n = a.n, M = a.m, LS = a.lds_stride;
  const V2<T> *chirp = static_cast<const V2<T> *>(a.chirp);
  const V2<T> *bhat = static_cast<const V2<T> *>(a.bhat);
  const int s = tid / a.tpu;
  const int lt = tid - s * a.tpu;
  V2<T> *A = lds + s * LS;
  for (int e = n + lt; e < M; e += a.tpu) A[lpad(e)] = V2<T>{0, 0};
  for (int f = tid; f < a.C * n; f += nthr) {
    int ss, e;
    slot_elem(f, a.load_slot_fast, a, a.lext, a.fd_lext, ss, e);
    const int64_t q = q0 + ss;
    V2<T> v{0, 0};
    if (q < a.nseqs) {
      v = src_elem<T>(a.in, seq_base(static_cast<uint32_t>(q), a, a.gi), a.gi.es, a.src_mode & 3, n, nullptr,
                      a.conj_in, e);
      v = cmul(v, chirp[e]);
    }
    lds[ss * LS + lpad(e)] = v;
  }
  __syncthreads();
  const V2<T> *tw = static_cast<const V2<T> *>(a.tw_small);
  lds_fft<T, E>(A, M, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, tw);
  for (int qq = lt; qq < M; qq += a.tpu) A[lpad(qq)] = cconj(cmul(A[lpad(qq)], bhat[qq]));
  __syncthreads();
  lds_fft<T, E>(A, M, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu, tw);
  const T scale = static_cast<T>(a.scale);
  for (int f = tid; f < a.C * a.sext; f += nthr) {
    int ss, e;
    slot_elem(f, a.store_slot_fast, a, a.sext, a.fd_sext, ss, e);
    const int64_t q = q0 + ss;
    if (q >= a.nseqs) continue;
    V2<T> v = cmul(cconj(lds[ss * LS + lpad(e)]), chirp[e]);
    dst_elem<T>(a.out, seq_base(static_cast<uint32_t>(q), a, a.go), a.go.es, a.dst_mode & 3, n, nullptr, nullptr,
                a.conj_out, scale, e, v);
  }
}

// ---------------------------------------------------------------------------
// Even-length R2C via an n/2-point complex transform, post-process fused:
//   z_j = x_2j + i x_2j+1, Z = FFT_h(z), X_k = E_k + w_n^k O_k with
//   E_k = (Z_k + conj Z_{h-k})/2, O_k = -i (Z_k - conj Z_{h-k})/2, k = 0..h (Z_h = Z_0).
// Input geometry counts reals (es must be 1 and bases even): z is read as V2.
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_r2c_even(const FusedArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V2<T> *lds = reinterpret_cast<V2<T> *>(smem_raw);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * a.C;
  const int h = a.m, LS = a.lds_stride;
  const V2<T> *wn = static_cast<const V2<T> *>(a.aux);
  for (int f = tid; f < a.C * h; f += nthr) {
    int s, e;
    slot_elem(f, a.load_slot_fast, a, h, a.fd_lext, s, e);
    const int64_t q = q0 + s;
    V2<T> v{0, 0};
    if (q < a.nseqs) {
      const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.gi);
      v = *reinterpret_cast<const V2<T> *>(static_cast<const T *>(a.in) + base + 2 * static_cast<int64_t>(e));
    }
    lds[s * LS + lpad(e)] = v;
  }
  __syncthreads();
  const int s = tid / a.tpu;
  const int lt = tid - s * a.tpu;
  lds_fft<T, E>(lds + s * LS, h, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu,
                static_cast<const V2<T> *>(a.tw_small));
  const T half = T(0.5);
  for (int f = tid; f < a.C * (h + 1); f += nthr) {
    int ss, k;
    slot_elem(f, a.store_slot_fast, a, h + 1, a.fd_sext, ss, k);
    const int64_t q = q0 + ss;
    if (q >= a.nseqs) continue;
    const V2<T> *Z = lds + ss * LS;
    const V2<T> zk = Z[lpad(k == h ? 0 : k)];
    const V2<T> zc = cconj(Z[lpad(k == 0 ? 0 : h - k)]);
    const V2<T> ev{(zk.x + zc.x) * half, (zk.y + zc.y) * half};
    const V2<T> od = cmul_mi(V2<T>{(zk.x - zc.x) * half, (zk.y - zc.y) * half});
    V2<T> X = cadd(ev, cmul(wn[k], od));
    if (a.conj_out) X.y = -X.y;
    const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.go);
    static_cast<V2<T> *>(a.out)[base + static_cast<int64_t>(k) * a.go.es] = X;
  }
}

// ---------------------------------------------------------------------------
// Even-length C2R via an n/2-point transform, pre-process fused:
//   Z'_k = (X_k + conj X_{h-k}) + i w_n^{-k} (X_k - conj X_{h-k}), k < h,
//   Im X_0 and Im X_h dropped; z = IDFT_h(Z') (unnormalised) -> (x_2j, x_2j+1).
// Output geometry counts reals (es == 1, even bases).
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ __launch_bounds__(kMaxBlockThreads) void k_c2r_even(const FusedArgs a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  V2<T> *lds = reinterpret_cast<V2<T> *>(smem_raw);
  const int tid = threadIdx.x, nthr = blockDim.x;
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * a.C;
  const int h = a.m, LS = a.lds_stride;
  const V2<T> *wn = static_cast<const V2<T> *>(a.aux);
  const V2<T> *X = static_cast<const V2<T> *>(a.in);
  for (int f = tid; f < a.C * h; f += nthr) {
    int s, k;
    slot_elem(f, a.load_slot_fast, a, h, a.fd_lext, s, k);
    const int64_t q = q0 + s;
    V2<T> v{0, 0};
    if (q < a.nseqs) {
      const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.gi);
      V2<T> xk = X[base + static_cast<int64_t>(k) * a.gi.es];
      V2<T> xh = X[base + static_cast<int64_t>(h - k) * a.gi.es];
      if (k == 0) {
        xk.y = T(0);
        xh.y = T(0);
      }
      xh.y = -xh.y;  // conj X_{h-k}
      const V2<T> sum = cadd(xk, xh);
      const V2<T> dif = csub(xk, xh);
      // i * conj(w_n^k) * dif
      const V2<T> t = cmul_pi(cmul(dif, cconj(wn[k])));
      v = cconj(cadd(sum, t));  // conj for the inverse-by-forward trick
    }
    lds[s * LS + lpad(k)] = v;
  }
  __syncthreads();
  const int s = tid / a.tpu;
  const int lt = tid - s * a.tpu;
  lds_fft<T, E>(lds + s * LS, h, a.nst, a.rad, a.fd_q, a.twq_step, lt, a.tpu,
                static_cast<const V2<T> *>(a.tw_small));
  const T scale = static_cast<T>(a.scale);
  for (int f = tid; f < a.C * h; f += nthr) {
    int ss, j;
    slot_elem(f, a.store_slot_fast, a, h, a.fd_sext, ss, j);
    const int64_t q = q0 + ss;
    if (q >= a.nseqs) continue;
    V2<T> z = cconj(lds[ss * LS + lpad(j)]);
    z.x *= scale;
    z.y *= scale;
    const int64_t base = seq_base(static_cast<uint32_t>(q), a, a.go);
    *reinterpret_cast<V2<T> *>(static_cast<T *>(a.out) + base + 2 * static_cast<int64_t>(j)) = z;
  }
}

// ---------------------------------------------------------------------------
// Split-spectrum passes for even-length real transforms that run multi-pass:
//   post: Z (h complex, contiguous) -> X (h+1 bins);  pre: X -> conj(Z').
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_real_post(const RealPassArgs a) {
  const int h = a.n / 2;
  const int64_t total = a.nseqs * (h + 1);
  const V2<T> *Zall = static_cast<const V2<T> *>(a.in);
  const V2<T> *wn = static_cast<const V2<T> *>(a.tw_half);
  const T half = T(0.5);
  for (int64_t g = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t q = g / (h + 1);
    const int k = static_cast<int>(g - q * (h + 1));
    const int64_t b = q / a.ncol, c = q - b * a.ncol;
    const V2<T> *Z = Zall + b * a.gi.bs + c * a.gi.cs;
    const V2<T> zk = Z[(k == h ? 0 : k) * a.gi.es];
    const V2<T> zc = cconj(Z[(k == 0 ? 0 : h - k) * a.gi.es]);
    const V2<T> ev{(zk.x + zc.x) * half, (zk.y + zc.y) * half};
    const V2<T> od = cmul_mi(V2<T>{(zk.x - zc.x) * half, (zk.y - zc.y) * half});
    V2<T> X = cadd(ev, cmul(wn[k], od));
    static_cast<V2<T> *>(a.out)[b * a.go.bs + c * a.go.cs + k * a.go.es] = X;
  }
}

template <typename T>
__global__ void k_real_pre(const RealPassArgs a) {
  const int h = a.n / 2;
  const int64_t total = a.nseqs * h;
  const V2<T> *wn = static_cast<const V2<T> *>(a.tw_half);
  for (int64_t g = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; g < total;
       g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t q = g / h;
    const int k = static_cast<int>(g - q * h);
    const int64_t b = q / a.ncol, c = q - b * a.ncol;
    const V2<T> *Xs = static_cast<const V2<T> *>(a.in) + b * a.gi.bs + c * a.gi.cs;
    V2<T> xk = Xs[k * a.gi.es];
    V2<T> xh = Xs[(h - k) * a.gi.es];
    if (k == 0) {
      xk.y = T(0);
      xh.y = T(0);
    }
    xh.y = -xh.y;
    const V2<T> t = cmul_pi(cmul(csub(xk, xh), cconj(wn[k])));
    static_cast<V2<T> *>(a.out)[b * a.go.bs + c * a.go.cs + k * a.go.es] = cconj(cadd(cadd(xk, xh), t));
  }
}

// ---------------------------------------------------------------------------
// FP64 table generation
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_gen_roots(V2<T> *out, int64_t count, int64_t n, int64_t mult) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < count;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const int64_t r = (j % n) * (mult % n) % n;
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

__global__ void k_gen_bluestein_b(double2 *out, int64_t n, int64_t m) {
  for (int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; j < m;
       j += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    int64_t k = -1;
    if (j < n) k = j;
    else if (m - j < n) k = m - j;
    if (k < 0) {
      out[j] = double2{0.0, 0.0};
      continue;
    }
    const uint64_t r = (static_cast<uint64_t>(k) * static_cast<uint64_t>(k)) % static_cast<uint64_t>(2 * n);
    double s, c;
    sincospi(static_cast<double>(r) / static_cast<double>(n), &s, &c);
    out[j] = double2{c, s};
  }
}

__global__ void k_gen_rader_b(double2 *out, const int32_t *idx, int64_t p) {
  for (int64_t q = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; q < p - 1;
       q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    double s, c;
    sincospi(-2.0 * static_cast<double>(idx[q]) / static_cast<double>(p), &s, &c);
    out[q] = double2{c, s};
  }
}

template <typename T>
__global__ void k_scale_cast(const double2 *in, V2<T> *out, int64_t count, double scale) {
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

template <typename K>
hipError_t prep_lds(K kernel, size_t lds) {
  if (lds > 64 * 1024) {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                               static_cast<int>(lds));
  }
  return hipSuccess;
}

template <typename Args, typename K>
hipError_t launch_one(K kernel, const Args &a, const LaunchShape &s, hipStream_t st) {
  hipError_t err = prep_lds(kernel, s.lds_bytes);
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(kernel, s.grid, s.block, s.lds_bytes, st, a);
  return hipGetLastError();
}

dim3 grid_for(int64_t total, int threads) {
  int64_t blocks = (total + threads - 1) / threads;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  return dim3(static_cast<unsigned>(blocks));
}

}  // namespace

bool supported_ept(bool dbl, int ept) { return dbl ? (ept == 8 || ept == 16) : (ept == 16 || ept == 32); }

#define FLAGFFT_DISPATCH_EPT(KERNEL, ARGS)                                                  \
  do {                                                                                      \
    if (dbl) {                                                                              \
      if (ept == 8) return launch_one(dev::KERNEL<double, 8>, ARGS, s, st);                 \
      if (ept == 16) return launch_one(dev::KERNEL<double, 16>, ARGS, s, st);               \
    } else {                                                                                \
      if (ept == 16) return launch_one(dev::KERNEL<float, 16>, ARGS, s, st);                \
      if (ept == 32) return launch_one(dev::KERNEL<float, 32>, ARGS, s, st);                \
    }                                                                                       \
    return hipErrorInvalidValue;                                                            \
  } while (0)

hipError_t launch_pass(const PassArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st) {
  FLAGFFT_DISPATCH_EPT(k_pass, a);
}
hipError_t launch_rader(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st) {
  FLAGFFT_DISPATCH_EPT(k_rader, a);
}
hipError_t launch_bluestein(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st) {
  FLAGFFT_DISPATCH_EPT(k_bluestein, a);
}
hipError_t launch_r2c_even(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st) {
  FLAGFFT_DISPATCH_EPT(k_r2c_even, a);
}
hipError_t launch_c2r_even(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st) {
  FLAGFFT_DISPATCH_EPT(k_c2r_even, a);
}

hipError_t launch_real_post(const RealPassArgs &a, bool dbl, hipStream_t st) {
  const int64_t total = a.nseqs * (a.n / 2 + 1);
  LaunchShape s{grid_for(total, 256), dim3(256), 0};
  if (dbl) return launch_one(dev::k_real_post<double>, a, s, st);
  return launch_one(dev::k_real_post<float>, a, s, st);
}

hipError_t launch_real_pre(const RealPassArgs &a, bool dbl, hipStream_t st) {
  const int64_t total = a.nseqs * (a.n / 2);
  LaunchShape s{grid_for(total, 256), dim3(256), 0};
  if (dbl) return launch_one(dev::k_real_pre<double>, a, s, st);
  return launch_one(dev::k_real_pre<float>, a, s, st);
}

hipError_t gen_roots(void *out, bool dbl, int64_t count, int64_t n, int64_t mult, hipStream_t st) {
  dim3 g = grid_for(count, 256);
  if (dbl) hipLaunchKernelGGL(dev::k_gen_roots<double>, g, dim3(256), 0, st, static_cast<double2 *>(out), count, n, mult);
  else hipLaunchKernelGGL(dev::k_gen_roots<float>, g, dim3(256), 0, st, static_cast<float2 *>(out), count, n, mult);
  return hipGetLastError();
}

hipError_t gen_chirp(void *out, bool dbl, int64_t n, hipStream_t st) {
  dim3 g = grid_for(n, 256);
  if (dbl) hipLaunchKernelGGL(dev::k_gen_chirp<double>, g, dim3(256), 0, st, static_cast<double2 *>(out), n);
  else hipLaunchKernelGGL(dev::k_gen_chirp<float>, g, dim3(256), 0, st, static_cast<float2 *>(out), n);
  return hipGetLastError();
}

hipError_t gen_bluestein_b(double2 *out, int64_t n, int64_t m, hipStream_t st) {
  hipLaunchKernelGGL(dev::k_gen_bluestein_b, grid_for(m, 256), dim3(256), 0, st, out, n, m);
  return hipGetLastError();
}

hipError_t gen_rader_b(double2 *out, const int32_t *idx, int64_t p, hipStream_t st) {
  hipLaunchKernelGGL(dev::k_gen_rader_b, grid_for(p - 1, 256), dim3(256), 0, st, out, idx, p);
  return hipGetLastError();
}

hipError_t scale_cast(const double2 *in, void *out, bool dbl, int64_t count, double scale, hipStream_t st) {
  dim3 g = grid_for(count, 256);
  if (dbl) hipLaunchKernelGGL(dev::k_scale_cast<double>, g, dim3(256), 0, st, in, static_cast<double2 *>(out), count, scale);
  else hipLaunchKernelGGL(dev::k_scale_cast<float>, g, dim3(256), 0, st, in, static_cast<float2 *>(out), count, scale);
  return hipGetLastError();
}

}  // namespace flagfft
