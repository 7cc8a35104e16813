// dev_common.hpp — device helpers shared by the gfx950 kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "codelets.hpp"
#include "fft_args.hpp"

namespace flagfft {
namespace dev {

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
  uint32_t t = __umulhi(n, f.mul);
  return (t + n) >> f.shift;
}

// Address-space-qualified views.  Non-inlined stage functions receive these
// so their LDS traffic is ds_read/ds_write and their table reads are global
// loads (a generic pointer would turn both into FLAT accesses).
#define FLAGFFT_LDS __attribute__((address_space(3)))
#define FLAGFFT_GLB __attribute__((address_space(1)))
template <typename T>
using LdsC = FLAGFFT_LDS V2<T>;
template <typename T>
using GlbC = FLAGFFT_GLB V2<T>;
template <typename T>
using GlbR = FLAGFFT_GLB T;

template <typename T>
__device__ __forceinline__ V2<T> ldc(const LdsC<T> *p) {
  return V2<T>{p->x, p->y};
}
template <typename T>
__device__ __forceinline__ void stc(LdsC<T> *p, V2<T> v) {
  p->x = v.x;
  p->y = v.y;
}
template <typename T>
__device__ __forceinline__ V2<T> ldg(const GlbC<T> *p) {
  return V2<T>{p->x, p->y};
}
template <typename T>
__device__ __forceinline__ void stg(GlbC<T> *p, V2<T> v) {
  p->x = v.x;
  p->y = v.y;
}
// Element `idx` of a table with a 32-bit index: the address stays a uniform
// base plus one 32-bit lane offset (global_load with saddr), which keeps many
// outstanding table loads cheap in registers.
template <typename T>
__device__ __forceinline__ V2<T> ldt(const GlbC<T> *base, uint32_t idx) {
  const FLAGFFT_GLB char *b = reinterpret_cast<const FLAGFFT_GLB char *>(base);
  return ldg(reinterpret_cast<const GlbC<T> *>(b + idx * static_cast<uint32_t>(sizeof(V2<T>))));
}

template <typename T>
__device__ __forceinline__ const GlbC<T> *gc(const void *p) {
  return (const GlbC<T> *)p;
}
template <typename T>
__device__ __forceinline__ GlbC<T> *gc(void *p) {
  return (GlbC<T> *)p;
}
template <typename T>
__device__ __forceinline__ const GlbR<T> *gr(const void *p) {
  return (const GlbR<T> *)p;
}
template <typename T>
__device__ __forceinline__ GlbR<T> *gr(void *p) {
  return (GlbR<T> *)p;
}
template <typename T>
__device__ __forceinline__ LdsC<T> *lds_base() {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  return (LdsC<T> *)smem_raw;
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
                                          const void *chirp, int conj_in, int64_t e) {
  if ((mode & kSrcChirpFlag) && e >= n) return V2<T>{0, 0};
  const int bm = mode & 3;
  V2<T> v;
  if (bm == kSrcComplex) {
    v = ldg(gc<T>(in) + base + e * es);
  } else if (bm == kSrcReal) {
    v = V2<T>{gr<T>(in)[base + e * es], T(0)};
  } else {
    const int64_t h = n >> 1;
    if (e <= h) {
      v = ldg(gc<T>(in) + base + e * es);
      if (e == 0 || 2 * e == n) v.y = T(0);
    } else {
      v = ldg(gc<T>(in) + base + (n - e) * es);
      v.y = -v.y;
    }
  }
  if (conj_in) v.y = -v.y;
  if (mode & kSrcChirpFlag) v = cmul(v, ldg(gc<T>(chirp) + e));
  return v;
}

template <typename T>
__device__ __forceinline__ void dst_elem(void *out, int64_t base, int64_t es, int mode, int64_t n, const void *chirp,
                                         const void *bhat, int conj_out, T scale, int64_t e, V2<T> v) {
  if (mode & kDstPointwiseFlag) {
    stg(gc<T>(out) + base + e * es, cconj(cmul(v, ldg(gc<T>(bhat) + e))));
    return;
  }
  if (mode & kDstChirpFlag) {
    if (e >= n) return;
    v = cmul(cconj(v), ldg(gc<T>(chirp) + e));
  }
  if (conj_out) v.y = -v.y;
  if (scale != T(1)) {
    v.x *= scale;
    v.y *= scale;
  }
  const int bm = mode & 3;
  if (bm == kDstComplex) {
    stg(gc<T>(out) + base + e * es, v);
  } else if (bm == kDstHalf) {
    if (2 * e <= n) stg(gc<T>(out) + base + e * es, v);
  } else {
    gr<T>(out)[base + e * es] = v.x;
  }
}

// Batched staged copy: every thread issues U independent loads before it
// commits any of them, so U global (or LDS) round trips overlap instead of
// serialising one per loop iteration.
template <typename T, int U, typename Load, typename Commit>
__device__ __forceinline__ void batched(int total, int tid, int nthr, Load load, Commit commit) {
  for (int f0 = tid; f0 < total; f0 += U * nthr) {
    V2<T> v[U];
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      const int f = f0 + uu * nthr;
      if (f < total) v[uu] = load(f);
    }
#pragma unroll
    for (int uu = 0; uu < U; ++uu) {
      const int f = f0 + uu * nthr;
      if (f < total) commit(f, v[uu]);
    }
  }
}

}  // namespace dev
}  // namespace flagfft
