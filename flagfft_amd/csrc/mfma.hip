// mfma.hip — split-precision matrix-core transform of length 1024 = 32 x 32 (gfx950).
//
// x[32 n1 + n2] is viewed as a 32x32 matrix.  The two radix-32 DFT stages are
// dense complex matrix products on v_mfma_f32_32x32x16_{bf16,f16}:
//
//   step 1   B^T[n2][k1] = sum_n1 x[32 n1 + n2] F[n1][k1]      (data = A operand)
//   twiddle  B^T[n2][k1] *= w_1024^(n2 k1)                    (FP32, in the accumulators)
//   step 2   C^T[k2][k1] = sum_n2 F[k2][n2] B^T[n2][k1]         (data = B operand)
//   out      X[k1 + 32 k2] = C^T[k2][k1]
//
// with F[a][b] = w_32^(ab).  Complex products are real products on stacked
// real/imaginary parts.  Every FP32 operand v is split into v_hi = half(v) and
// v_lo = half(v - v_hi); each real product keeps hi*hi + hi*lo + lo*hi (three
// MFMAs, FP32 accumulation), dropping only lo*lo.
//
// Lane layout of the 32x32x16 instructions (wave64, lane l, c = l % 32,
// h = l / 32): A operand A[c][8h + v], B operand B[8h + v][c] (v < 8),
// accumulator register r holds D[8(r/4) + 4h + r%4][c].  The accumulator of
// step 1 is reused as step 2's B operand without any data movement: chunk q
// of the K loop pairs lane group h, element v with n2 = 16q + 8(v/4) + 4h + v%4,
// which is accumulator register 8q + v, and step 2's F operand is permuted the
// same way.  Step 1 sums over n1 = 16q + 2v + h and step 2's output rows are
// permuted (k2 = pi(row), see out_row) so that the two half waves load
// (x[32 n1 + c]) and store (X[c + 32 k2]) adjacent rows: one 512-byte run per
// instruction.
//
// FP16 has a narrow exponent range, so each transform is scaled by a power of
// two (exact) that brings max |x| below 1 before the split, and scaled back on
// store.
#include <hip/hip_runtime.h>

#include "dev_common.hpp"
#include "fft_args.hpp"
#include "kernels.hpp"

#include <string>

namespace flagfft {
namespace dev {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

template <bool F16>
struct HalfT;
template <>
struct HalfT<false> {
  using E = __bf16;
  using V8 = bf16x8;
};
template <>
struct HalfT<true> {
  using E = _Float16;
  using V8 = f16x8;
};

template <bool F16>
struct Split8 {
  typename HalfT<F16>::V8 hi, lo;
};

template <bool F16>
__device__ __forceinline__ f32x16 mma(typename HalfT<F16>::V8 a, typename HalfT<F16>::V8 b, f32x16 c) {
  if constexpr (F16)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// acc += a * b with both operands split: hi*hi + hi*lo + lo*hi.
template <bool F16>
__device__ __forceinline__ f32x16 mma3(const Split8<F16> &a, const Split8<F16> &b, f32x16 acc) {
  acc = mma<F16>(a.lo, b.hi, acc);
  acc = mma<F16>(a.hi, b.lo, acc);
  return mma<F16>(a.hi, b.hi, acc);
}

// element v of s := split of x
template <bool F16>
__device__ __forceinline__ void split_into(Split8<F16> &s, int v, float x) {
  using E = typename HalfT<F16>::E;
  const E hi = static_cast<E>(x);
  s.hi[v] = hi;
  s.lo[v] = static_cast<E>(x - static_cast<float>(hi));
}

template <bool F16>
__device__ __forceinline__ Split8<F16> neg(const Split8<F16> &s) {
  return Split8<F16>{-s.hi, -s.lo};
}

// Output row permutation of step 2: accumulator row i = 8a + 4h + b holds
// k2 = 8a + 2b + h, so the two half waves store adjacent 256-byte rows.
__device__ __forceinline__ int out_row(int i) { return 8 * (i >> 3) + 2 * (i & 3) + ((i >> 2) & 1); }

// w_1024^e from the FP32 table (e reduced mod 1024).
__device__ __forceinline__ V2<float> w1024(const GlbC<float> *tw, int e) { return ldg(tw + (e & 1023)); }

// Maximum of a non-negative float over the wavefront: DPP exchanges within
// each row of 16 lanes (quad swaps, half-row and row mirrors), then the four
// row maxima through scalar registers.  Non-negative floats order as their
// bit patterns.
__device__ __forceinline__ float wave_max_nonneg(float m) {
  int b = __float_as_int(m);
  b = max(b, __builtin_amdgcn_mov_dpp(b, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
  b = max(b, __builtin_amdgcn_mov_dpp(b, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
  b = max(b, __builtin_amdgcn_mov_dpp(b, 0x141, 0xF, 0xF, false));  // row_half_mirror
  b = max(b, __builtin_amdgcn_mov_dpp(b, 0x140, 0xF, 0xF, false));  // row_mirror
  const int r = max(max(__builtin_amdgcn_readlane(b, 0), __builtin_amdgcn_readlane(b, 16)),
                    max(__builtin_amdgcn_readlane(b, 32), __builtin_amdgcn_readlane(b, 48)));
  return __int_as_float(r);
}

// PF: transforms whose loads are in flight per wave (a ring of PF register
// sets); WPE: waves per SIMD the register allocation leaves room for.  The
// launch puts one workgroup on each CU, all resident at once (no tail of
// workgroups that start when the first ones finish).
template <bool F16, int PF, int WPE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_mfma1024(const MfmaArgs a) {
  const int lane = threadIdx.x & 63;
  const int c = lane & 31;
  const int h = lane >> 5;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int64_t nwaves = static_cast<int64_t>(gridDim.x) * (blockDim.x >> 6);
  const GlbC<float> *tw = gc<float>(a.tw);
  using V8 = typename HalfT<F16>::V8;

  // per-lane constant operands (identical for every wave) in LDS, 24 x 16 B
  // per lane, read back at each use: registers for a third wave per SIMD
  __shared__ uint4 kc[24][64];
  if (threadIdx.x < 64) {
    const int pc = out_row(c);
    Split8<F16> fb_r[2], fb_i[2], fa_r[2], fa_i[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const V2<float> f1 = w1024(tw, 32 * (16 * q + 2 * v + h) * c);
        const V2<float> f2 = w1024(tw, 32 * pc * (16 * q + 8 * (v >> 2) + 4 * h + (v & 3)));
        split_into<F16>(fb_r[q], v, f1.x);
        split_into<F16>(fb_i[q], v, f1.y);
        split_into<F16>(fa_r[q], v, f2.x);
        split_into<F16>(fa_i[q], v, f2.y);
      }
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      kc[0 + 2 * q][lane] = __builtin_bit_cast(uint4, fb_r[q].hi);
      kc[1 + 2 * q][lane] = __builtin_bit_cast(uint4, fb_r[q].lo);
      kc[4 + 2 * q][lane] = __builtin_bit_cast(uint4, fb_i[q].hi);
      kc[5 + 2 * q][lane] = __builtin_bit_cast(uint4, fb_i[q].lo);
      kc[8 + 2 * q][lane] = __builtin_bit_cast(uint4, fa_r[q].hi);
      kc[9 + 2 * q][lane] = __builtin_bit_cast(uint4, fa_r[q].lo);
      kc[12 + 2 * q][lane] = __builtin_bit_cast(uint4, fa_i[q].hi);
      kc[13 + 2 * q][lane] = __builtin_bit_cast(uint4, fa_i[q].lo);
    }
#pragma unroll
    for (int r = 0; r < 16; r += 2) {
      const V2<float> t0 = w1024(tw, (8 * (r >> 2) + 4 * h + (r & 3)) * c);
      const V2<float> t1 = w1024(tw, (8 * ((r + 1) >> 2) + 4 * h + ((r + 1) & 3)) * c);
      kc[16 + r / 2][lane] = uint4{__float_as_uint(t0.x), __float_as_uint(t0.y), __float_as_uint(t1.x),
                                   __float_as_uint(t1.y)};
    }
  }
  __syncthreads();
  auto ks = [&](int i) __attribute__((always_inline)) {
    return Split8<F16>{__builtin_bit_cast(V8, kc[i][lane]), __builtin_bit_cast(V8, kc[i + 1][lane])};
  };

  const float cin = a.conj_in ? -1.f : 1.f;
  const float cout = a.conj_out ? -1.f : 1.f;
  const float user_scale = static_cast<float>(a.scale);

  // software pipeline: the loads of the next PF transforms are in flight
  // while this one runs through the matrix cores
  auto load = [&](int64_t sq, V2<float> (&x)[2][8]) __attribute__((always_inline)) {
    const GlbC<float> *src = gc<float>(a.in) + sq * a.idist + c;
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 8; ++v) x[q][v] = ldg(src + 32 * (16 * q + 2 * v + h));
  };
  V2<float> ring[PF][2][8];
#pragma unroll
  for (int s = 0; s < PF; ++s)
    if (wave + s * nwaves < a.nseq) load(wave + s * nwaves, ring[s]);
  for (int64_t base = wave; base < a.nseq; base += PF * nwaves) {
#pragma unroll
  for (int s = 0; s < PF; ++s) {
    const int64_t seq = base + s * nwaves;
    if (seq >= a.nseq) break;
    GlbC<float> *dst = gc<float>(a.out) + seq * a.odist + c;
    V2<float> x[2][8];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 8; ++v) x[q][v] = ring[s][q][v];
    if (seq + PF * nwaves < a.nseq) load(seq + PF * nwaves, ring[s]);
    asm volatile("" ::: "memory");  // the constants are read from LDS in every iteration

    // exact power-of-two range scaling (FP16 only)
    float pre = 1.f, post = 1.f;
    if constexpr (F16) {
      float m = 0.f;
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int v = 0; v < 8; ++v) m = fmaxf(m, fmaxf(fabsf(x[q][v].x), fabsf(x[q][v].y)));
      m = wave_max_nonneg(m);
      if (m > 0.f && m == m && m <= 3.0e38f) {
        int e = 0;
        (void)frexpf(m, &e);  // m = f 2^e, f in [0.5, 1)
        pre = ldexpf(1.f, -e);
        post = ldexpf(1.f, e);
      }
    }

    Split8<F16> ar[2], ai[2];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        split_into<F16>(ar[q], v, x[q][v].x * pre);
        split_into<F16>(ai[q], v, cin * x[q][v].y * pre);
      }

    // step 1: B^T = X^T F  (re: Ar Fr - Ai Fi, im: Ar Fi + Ai Fr)
    f32x16 br = {}, bi = {};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const Split8<F16> fbr = ks(0 + 2 * q), fbi = ks(4 + 2 * q);
      br = mma3<F16>(ar[q], fbr, br);
      br = mma3<F16>(neg<F16>(ai[q]), fbi, br);
      bi = mma3<F16>(ar[q], fbi, bi);
      bi = mma3<F16>(ai[q], fbr, bi);
    }

    // twiddle in FP32, then split into step 2's B operand (register 8q + v)
    Split8<F16> sr[2], si[2];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const uint4 tt = kc[16 + r / 2][lane];
      const V2<float> t = (r & 1) ? V2<float>{__uint_as_float(tt.z), __uint_as_float(tt.w)}
                                  : V2<float>{__uint_as_float(tt.x), __uint_as_float(tt.y)};
      const float re = br[r] * t.x - bi[r] * t.y;
      const float im = br[r] * t.y + bi[r] * t.x;
      split_into<F16>(sr[r >> 3], r & 7, re);
      split_into<F16>(si[r >> 3], r & 7, im);
    }

    // step 2: C^T = F B^T  (re: Fr Br - Fi Bi, im: Fr Bi + Fi Br)
    f32x16 cr = {}, ci = {};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const Split8<F16> far = ks(8 + 2 * q), fai = ks(12 + 2 * q);
      cr = mma3<F16>(far, sr[q], cr);
      cr = mma3<F16>(fai, neg<F16>(si[q]), cr);
      ci = mma3<F16>(far, si[q], ci);
      ci = mma3<F16>(fai, sr[q], ci);
    }

    const float so = post * user_scale;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int k2 = 8 * (r >> 2) + 2 * (r & 3) + h;  // pi(8(r/4) + 4h + r%4)
      stg(dst + 32 * k2, V2<float>{cr[r] * so, cout * ci[r] * so});
    }
  }
  }
}

}  // namespace dev

namespace {

struct MfmaEntry {
  int pf, wpe;
  const void *f16, *bf16;
};

#define FLAGFFT_MFMA(PF, WPE) \
  {PF, WPE, (const void *)dev::k_mfma1024<true, PF, WPE>, (const void *)dev::k_mfma1024<false, PF, WPE>}

// Variant 0 is the default.  Measured on MI355X at 1024 x 65536 (one
// workgroup of four waves per CU, the fastest launch for every variant):
// BF16 191-192 us with PF = 2, FP16 197-200 us with PF = 2 or 3; two or three
// workgroups per CU ran 5-10 % slower.
const MfmaEntry kMfmaTable[] = {
    FLAGFFT_MFMA(2, 2), FLAGFFT_MFMA(3, 2), FLAGFFT_MFMA(1, 3), FLAGFFT_MFMA(2, 1), FLAGFFT_MFMA(4, 1),
};

}  // namespace

int mfma_variants() { return static_cast<int>(sizeof(kMfmaTable) / sizeof(kMfmaTable[0])); }

const void *mfma_kernel(int mode, int variant) {
  if (variant < 0 || variant >= mfma_variants()) return nullptr;
  if (mode == kComputeFp16x3) return kMfmaTable[variant].f16;
  if (mode == kComputeBf16x3) return kMfmaTable[variant].bf16;
  return nullptr;
}

std::string mfma_kernel_targs(int variant) {
  return std::to_string(kMfmaTable[variant].pf) + "," + std::to_string(kMfmaTable[variant].wpe);
}

}  // namespace flagfft
