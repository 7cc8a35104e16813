// colpass_body.hpp — register-resident two-stage super-stages (gfx950 device code).
//
// One launch is one pass of a multi-pass chain (the reference's four-step
// column / row passes, python/flagfft_codegen/kernels.py:710-742) for a
// super-radix R = RA x RB whose two stages both run in registers:
//   * C adjacent units per workgroup, lane -> (unit slot = tid % C, t = tid / C),
//     so every HBM access of a wavefront is C-element row segments (C x 8 bytes
//     in FP32: 128-512 byte runs, the width a strided HBM stream needs);
//   * stage A: thread t loads the RA elements t + RB j (j < RA), applies the
//     inter-pass twiddles, and transforms them in registers (radix RA);
//   * inner twiddles w_R^(t k1), then the RA x RB transpose within each unit
//     goes through LDS in NCH chunks of k1, so a workgroup needs 1 / NCH of
//     its tile in LDS and two workgroups share a CU;
//   * stage B: thread t' transforms the RB values of k1 = t' + RB m (radix RB)
//     and holds outputs k1 + RA k2;
//   * the first pass of contiguous rows stores each unit's R outputs as one
//     contiguous block (through LDS, NCH chunks of k2); later passes store
//     C-element runs directly.
// Same index map, twiddles and results as the LDS Stockham super-stage
// (stage.hip, mode kStagePass): unit i of sequence b reads x[i + e tcnt]
// (e < R), multiplies element e by w_n^(e kt twp_mult), transforms, and
// writes y[(i - k) R + k + e' p], k = i mod p.
//
// Included by colpass.hip (the standalone kernels) and kernels.hip (the
// persistent queue kernel runs the same bodies).
#pragma once

#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "dev_common.hpp"
#include "fft_args.hpp"

namespace flagfft {
namespace dev {

// LDS slot of exchange entry (kk, t, slot) of a chunk: slots fastest, so a
// wavefront's stores of one kk are 64 consecutive entries; one pad entry per 32.
template <int C, int RB>
__device__ __forceinline__ int xslot(int kk, int t, int sl) {
  const int idx = (kk * RB + t) * C + sl;
  return idx + (idx >> 5);
}

// Store to element `idx` of an array with a uniform base and a 32-bit lane
// offset (global_store with saddr), as ldt loads.
template <typename T>
__device__ __forceinline__ void stt(GlbC<T> *base, uint32_t idx, V2<T> v) {
  FLAGFFT_GLB char *b = reinterpret_cast<FLAGFFT_GLB char *>(base);
  stg(reinterpret_cast<GlbC<T> *>(b + idx * static_cast<uint32_t>(sizeof(V2<T>))), v);
}

// p[j] = w^j for j < 8 (p[0] = 1), each at most three products deep.
template <typename T>
__device__ __forceinline__ void pow8(V2<T> w, V2<T> *p) {
  p[0] = V2<T>{1, 0};
  p[1] = w;
  p[2] = cmul(w, w);
  p[3] = cmul(p[2], w);
  p[4] = cmul(p[2], p[2]);
  p[5] = cmul(p[4], w);
  p[6] = cmul(p[3], p[3]);
  p[7] = cmul(p[4], p[3]);
}

// Loads of tile `blk`: element e = t + RB j of the thread's unit (zero
// outside the units and, for a Bluestein chirp source, past the user length).
template <typename T, int C, int RA, int RB, bool PAIR = false>
__device__ __forceinline__ void col_load(const StageArgs &a, int blk, V2<T> (&x)[RA]) {
  const int tid = threadIdx.x;
  const int sl = tid % C;
  const int t = tid / C;
  const int64_t u0 = static_cast<int64_t>(blk) * C;
  const bool active = u0 + sl < a.units;
  const uint32_t bq = fdiv(static_cast<uint32_t>(u0), a.fd_tcnt);
  const uint32_t i = static_cast<uint32_t>(u0) - bq * static_cast<uint32_t>(a.tcnt) + sl;
  const uint32_t tcnt = static_cast<uint32_t>(a.tcnt);
  const GlbC<T> *src = gc<T>(a.in) + static_cast<int64_t>(bq) * a.bs_in;
  const bool chirp_in = (a.src_mode & kSrcChirpFlag) != 0;
  const uint32_t nlog = static_cast<uint32_t>(a.nlog);
  if (PAIR && a.dc_nyq_pair == 2 && blk == 0 && sl == 0) {
    // unit 0 of an inverse 2D real column pass: Herm(A) + i Herm(B) for the
    // DC column A (column 0) and the Nyquist column B (column `units`)
    constexpr int R = RA * RB;
    const uint32_t cb = static_cast<uint32_t>(a.units);
    const T half = T(0.5);
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      const uint32_t e = static_cast<uint32_t>(t + RB * j);
      const uint32_t em = e == 0 ? 0u : static_cast<uint32_t>(R) - e;
      const V2<T> va = ldt(src, e * tcnt), vam = ldt(src, em * tcnt);
      const V2<T> vb = ldt(src, cb + e * tcnt), vbm = ldt(src, cb + em * tcnt);
      const V2<T> ha{(va.x + vam.x) * half, (va.y - vam.y) * half};
      const V2<T> hb{(vb.x + vbm.x) * half, (vb.y - vbm.y) * half};
      x[j] = V2<T>{ha.x - hb.y, ha.y + hb.x};
    }
    return;
  }
#pragma unroll
  for (int j = 0; j < RA; ++j) {
    const uint32_t e = i + static_cast<uint32_t>(t + RB * j) * tcnt;
    x[j] = (active && (!chirp_in || e < nlog)) ? ldt(src, e) : V2<T>{0, 0};
  }
}

// Everything after the loads of tile `blk` (x: what col_load gave).
// TWM: twiddles from table loads per power (0) or from two or three loads and
// products (1: w^j = w^(j mod 8) (w^8)^(j / 8), fewer memory instructions).
// PAIR: the instance also runs the DC / Nyquist pair modes (StageArgs::dc_nyq_pair).
template <typename T, int C, int RA, int RB, int NCH, int TWM, bool PAIR = false>
__device__ __forceinline__ void col_rest(const StageArgs &a, int blk, V2<T> (&x)[RA]) {
  constexpr int R = RA * RB;
  constexpr int NM = RA / RB;   // stage-B transforms per thread
  constexpr int KC = RA / NCH;  // k1 values per exchange chunk
  constexpr int KD = RB / NCH;  // k2 values per block-store chunk
  constexpr int BL = R / NCH;   // block-store chunk length per unit
  constexpr int NT = C * RB;
  static_assert(RA % RB == 0 && RA % NCH == 0 && RB % NCH == 0, "shape");
  static_assert((C & (C - 1)) == 0, "C must be a power of two");
  LdsC<T> *lds = lds_base<T>();
  const int tid = threadIdx.x;
  const int sl = tid % C;
  const int t = tid / C;
  // The host launches this kernel only when C divides tcnt (every tile lies
  // in one sequence) and a sequence spans under 4 GiB: the sequence base is
  // uniform and every element offset fits 32 bits.
  const int64_t u0 = static_cast<int64_t>(blk) * C;
  const bool active = u0 + sl < a.units;
  const uint32_t bq = fdiv(static_cast<uint32_t>(u0), a.fd_tcnt);
  const uint32_t i = static_cast<uint32_t>(u0) - bq * static_cast<uint32_t>(a.tcnt) + sl;
  const bool first = a.p == a.ncol;
  const uint32_t k = i - fdiv(i, a.fd_p) * static_cast<uint32_t>(a.p);
  const uint32_t kt = first ? 0u : fdiv(k, a.fd_ncol);
  const uint32_t tcnt = static_cast<uint32_t>(a.tcnt);
  const GlbC<T> *tws = gc<T>(a.tw_small);
  const GlbC<T> *tw_lo = gc<T>(a.tw_lo);
  const GlbC<T> *tw_hi = gc<T>(a.tw_hi);
  const uint32_t tw_mask = (1u << a.tw_bits) - 1u;
  const uint32_t kmul = kt * static_cast<uint32_t>(a.twp_mult);
  const T scale = static_cast<T>(a.scale);

  // ---- input modes of the loaded elements e = t + RB j of the unit -----------
  // Bluestein chains: the first pass reads chirp[e] x[e] for e < n (zero past
  // the user length n = nlog), the last passes write conj(X bhat) (forward
  // chain) or conj(X) chirp for e < n (inverse chain)
  const bool chirp_in = (a.src_mode & kSrcChirpFlag) != 0;
  const uint32_t nlog = static_cast<uint32_t>(a.nlog);
  const GlbC<T> *chirp = gc<T>(a.chirp);
  constexpr int L8 = RA < 8 ? RA : 8;
  constexpr int H8 = (RA + 7) / 8;
  if (a.conj_in) {
#pragma unroll
    for (int j = 0; j < RA; ++j) x[j].y = -x[j].y;
  }
  if (chirp_in) {
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      const uint32_t e = i + static_cast<uint32_t>(t + RB * j) * tcnt;
      if (active && e < nlog) x[j] = cmul(x[j], ldt(chirp, e));
    }
  }
  // inter-pass twiddles: element e = t + RB j times w^(e kmul) =
  // w^((t + RB j0) kmul) w^(8 RB jh kmul), j = j0 + 8 jh, from the two-level table
  if (!first) {
    auto tw2 = [&](uint32_t mm) __attribute__((always_inline)) {
      return cmul(ldg(tw_hi + (mm >> a.tw_bits)), ldg(tw_lo + (mm & tw_mask)));
    };
    V2<T> blo[8], shi[H8];
    if constexpr (TWM == 0) {
#pragma unroll
      for (int j = 0; j < L8; ++j) blo[j] = tw2(static_cast<uint32_t>(t + RB * j) * kmul);
#pragma unroll
      for (int j = 1; j < H8; ++j) shi[j] = tw2(static_cast<uint32_t>(8 * RB * j) * kmul);
    } else {
      const V2<T> base = tw2(static_cast<uint32_t>(t) * kmul);
      V2<T> sp[8];
      pow8(tw2(static_cast<uint32_t>(RB) * kmul), sp);
      blo[0] = base;
#pragma unroll
      for (int j = 1; j < L8; ++j) blo[j] = cmul(base, sp[j]);
      if constexpr (H8 > 1) {
        shi[1] = tw2(static_cast<uint32_t>(8 * RB) * kmul);
#pragma unroll
        for (int j = 2; j < H8; ++j) shi[j] = cmul(shi[j - 1], shi[1]);
      }
    }
#pragma unroll
    for (int j = 0; j < RA; ++j) {
      const V2<T> w = (j >= 8) ? cmul(blo[j & 7], shi[j >> 3]) : blo[j];
      x[j] = cmul(x[j], w);
    }
  }

  // ---- stage A (radix RA over j) and the inner twiddles ------------------------
  dft<RA>(x);
  // w_R^(t k1) = lo[k1 mod 8] hi[k1 / 8], straight from the length-R table
  // (t k1 < R): two roundings per twiddle
  V2<T> wlo[8], whi[H8];
  if constexpr (TWM == 0) {
#pragma unroll
    for (int j = 0; j < L8; ++j) wlo[j] = ldt(tws, static_cast<uint32_t>(t * j));
#pragma unroll
    for (int j = 0; j < H8; ++j) whi[j] = ldt(tws, static_cast<uint32_t>(t * 8 * j));
  } else {
    pow8(ldt(tws, static_cast<uint32_t>(t)), wlo);
    if constexpr (H8 > 1) {
      whi[1] = ldt(tws, static_cast<uint32_t>(t * 8));
#pragma unroll
      for (int j = 2; j < H8; ++j) whi[j] = cmul(whi[j - 1], whi[1]);
    }
  }
#pragma unroll
  for (int j = 1; j < RA; ++j) {
    const V2<T> w = (j >= 8) ? ((j & 7) ? cmul(wlo[j & 7], whi[j >> 3]) : whi[j >> 3]) : wlo[j];
    x[j] = cmul(x[j], w);
  }

  // ---- RA x RB transpose within each unit and stage B (radix RB over t) -------
  V2<T> y[NM][RB];
  if constexpr (NCH == 2 && RA == RB) {
    // Two chunks, each holding half of every unit: entry (k1, tt) travels in
    // chunk c = (tt mod 2) xor (k1 >= H).  Every thread writes H values and
    // reads H values per chunk (the values of its k1 = t of one tt parity),
    // so no thread holds more than RA values at a time; stage B is then
    // radix 2 over the even / odd tt transforms of length H.
    //
    // Stage B runs on (unit slot, k1) = (tid % C, tid / C) for passes that
    // store C-element runs, or (tid / RA, tid % RA) for the first pass of
    // contiguous rows: there 32 consecutive lanes hold consecutive outputs
    // k1 + RA k2 of one unit and store RA-element runs directly.  The row
    // pitch of the chunk buffer keeps either reader at two accesses per bank.
    constexpr int H = RA / 2;
    V2<T> ye[H], yo[H];
    const bool bs = a.block_store != 0;
    const int pitch = bs ? H * C + 1 : H * C + 8;
    const int rsl = bs ? tid / RA : sl;  // unit slot of this thread's stage-B transform
    const int rk1 = bs ? tid % RA : t;   // and its k1
    const int hw = rk1 / H;
    // odd t write their upper k1 half in chunk 0: swap the halves in
    // registers first (constant indices: a runtime index into x would put the
    // array in scratch memory)
    if (t & 1) {
#pragma unroll
      for (int kk = 0; kk < H; ++kk) {
        const V2<T> tmp = x[kk];
        x[kk] = x[H + kk];
        x[H + kk] = tmp;
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (c > 0) __syncthreads();  // the previous chunk's reads are done
      const int h = (t & 1) ^ c;    // the k1 half this thread writes in chunk c
#pragma unroll
      for (int kk = 0; kk < H; ++kk) stc(lds + (h * H + kk) * pitch + (t >> 1) * C + sl, x[c * H + kk]);
      __syncthreads();
      const int par = hw ^ c;  // tt parity this thread reads in chunk c
#pragma unroll
      for (int q = 0; q < H; ++q) {
        const V2<T> v = ldc(lds + rk1 * pitch + q * C + rsl);
        if (par) yo[q] = v;
        else ye[q] = v;
      }
    }
    dft<H>(ye);
    dft<H>(yo);
    sfor<H>([&](auto K) __attribute__((always_inline)) {
      constexpr int k2 = decltype(K)::value;
      const V2<T> o = twiddle_ct<k2, RA>(yo[k2]);
      y[0][k2] = cadd(ye[k2], o);
      y[0][k2 + H] = csub(ye[k2], o);
    });
    if (bs) {
      // unit u0 + rsl owns outputs (u0 + rsl) R + e', e' = rk1 + RA k2
      if (u0 + rsl < a.units) {
        GlbC<T> *dst = gc<T>(a.out) + u0 * R;
        const uint32_t o0 = static_cast<uint32_t>(rsl * R + rk1);
#pragma unroll
        for (int q = 0; q < RB; ++q) {
          V2<T> v = y[0][q];
          if (a.conj_out) v.y = -v.y;
          if (scale != T(1)) {
            v.x *= scale;
            v.y *= scale;
          }
          stt(dst, o0 + static_cast<uint32_t>(RA * q), v);
        }
      }
      return;
    }
  } else {
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
      if (c > 0) __syncthreads();  // the previous chunk's reads are done
#pragma unroll
      for (int kk = 0; kk < KC; ++kk) stc(lds + xslot<C, RB>(kk, t, sl), x[c * KC + kk]);
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const int k1 = t + RB * m;
        if (k1 / KC == c) {
          const int kk = k1 - c * KC;
#pragma unroll
          for (int tt = 0; tt < RB; ++tt) y[m][tt] = ldc(lds + xslot<C, RB>(kk, tt, sl));
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NM; ++m) dft<RB>(y[m]);
  }

  auto finish = [&](V2<T> v) __attribute__((always_inline)) {
    if (a.conj_out) v.y = -v.y;
    if (scale != T(1)) {
      v.x *= scale;
      v.y *= scale;
    }
    return v;
  };

  if (a.block_store) {
    // the C units' outputs are one contiguous block of C R elements; each
    // chunk d carries outputs e' = k1 + RA k2 with k2 in [d KD, (d + 1) KD),
    // a contiguous range of BL elements per unit
    // row pitch = 4 (mod 32) entries: the 16 slots of a wavefront's store land
    // on distinct bank octets
    constexpr int LSB = (BL + BL / 32 + 28) / 32 * 32 + 4;
    GlbC<T> *dst = gc<T>(a.out) + u0 * R;
    const int64_t total = a.units * static_cast<int64_t>(R);
#pragma unroll
    for (int d = 0; d < NCH; ++d) {
      __syncthreads();  // the exchange (or previous chunk) reads are done
#pragma unroll
      for (int m = 0; m < NM; ++m) {
        const int k1 = t + RB * m;
#pragma unroll
        for (int q = 0; q < KD; ++q) {
          const int e2 = k1 + RA * q;  // e' - d BL
          stc(lds + sl * LSB + e2 + (e2 >> 5), y[m][d * KD + q]);
        }
      }
      __syncthreads();
      constexpr int PER = (C * BL + NT - 1) / NT;
#pragma unroll
      for (int r = 0; r < PER; ++r) {
        const int f = tid + r * NT;
        if (f < C * BL) {
          const int s2 = f / BL, e2 = f - s2 * BL;
          const int64_t g = static_cast<int64_t>(s2) * R + d * BL + e2;
          if (u0 * R + g < total) stg(dst + g, finish(ldc(lds + s2 * LSB + e2 + (e2 >> 5))));
        }
      }
    }
  } else {
    // forward 2D real column pass with the DC / Nyquist pair in unit 0: its
    // outputs go through LDS so that each thread sees Z_e and Z_-e
    // (workgroup 0 only, uniform over the workgroup)
    const bool unpack = PAIR && a.dc_nyq_pair == 1 && blk == 0;
    if (unpack) {
      __syncthreads();  // the exchange's reads are done
      if (sl == 0) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
#pragma unroll
          for (int q = 0; q < RB; ++q) stc(lds + (t + RB * m + RA * q), y[m][q]);
        }
      }
      __syncthreads();
    }
    if (active) {
      GlbC<T> *dst = gc<T>(a.out) + static_cast<int64_t>(bq) * a.bs_out;
      const uint32_t o0 = (i - k) * static_cast<uint32_t>(R) + k;
      const uint32_t pp = static_cast<uint32_t>(a.p);
      const GlbC<T> *bhat = gc<T>(a.bhat);
      if (unpack && sl == 0) {
        // unit 0 is column 0 of the single pass (o0 = 0); B goes to column `units`
        const uint32_t cb = static_cast<uint32_t>(a.units);
        const T half = T(0.5);
#pragma unroll
        for (int m = 0; m < NM; ++m) {
#pragma unroll
          for (int q = 0; q < RB; ++q) {
            const int e = t + RB * m + RA * q;
            const V2<T> z = y[m][q];
            const V2<T> zm = ldc(lds + (e == 0 ? 0 : R - e));  // Z_-e (conjugated below)
            const V2<T> va{(z.x + zm.x) * half, (z.y - zm.y) * half};
            const V2<T> vb{(z.y + zm.y) * half, (zm.x - z.x) * half};
            stt(dst, static_cast<uint32_t>(e) * pp, finish(va));
            stt(dst, cb + static_cast<uint32_t>(e) * pp, finish(vb));
          }
        }
      } else {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
          const int k1 = t + RB * m;
#pragma unroll
          for (int q = 0; q < RB; ++q) {
            const uint32_t o = o0 + static_cast<uint32_t>(k1 + RA * q) * pp;
            if (a.dst_mode & kDstPointwiseFlag) {
              stt(dst, o, cconj(cmul(y[m][q], ldt(bhat, o))));
            } else if (a.dst_mode & kDstChirpFlag) {
              if (o < nlog) stt(dst, o, finish(cmul(cconj(y[m][q]), ldt(chirp, o))));
            } else {
              stt(dst, o, finish(y[m][q]));
            }
          }
        }
      }
    }
  }
}

// One tile (workgroup `blk` of the launch) of a register-resident pass; every
// one of the C * RB threads of the workgroup must call it (it synchronises).
template <typename T, int C, int RA, int RB, int NCH, int TWM, bool PAIR = false>
__device__ __forceinline__ void col_body(const StageArgs &a, int blk) {
  V2<T> x[RA];
  col_load<T, C, RA, RB, PAIR>(a, blk, x);
  col_rest<T, C, RA, RB, NCH, TWM, PAIR>(a, blk, x);
}

}  // namespace dev
}  // namespace flagfft
