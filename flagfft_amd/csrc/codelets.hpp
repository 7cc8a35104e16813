// codelets.hpp — forward radix-R DFT butterflies on register arrays (gfx950 device code).
//
// Arithmetic contract (the reference's codelets python/flagfft_codegen/codelet/radix*.py
// compute the same forward DFT_R with exp(-2 pi i jk / R)): v[k] <- sum_j v[j] w_R^{jk}.
// Inverse transforms are formed by the callers as conj(DFT(conj(x))).
//
// Structure: radix 2 and 4 by hand; prime radices by the symmetric-pair
// formula X_k = x0 + sum_j (x_j + x_{R-j}) cos - i (x_j - x_{R-j}) sin;
// composite radices by one Cooley-Tukey split R = A*B with compile-time
// twiddles, trivial ones (1, -1, +-i, (+-1 +- i)/sqrt2) specialised.
#pragma once

#include <hip/hip_runtime.h>

#include <utility>

namespace flagfft {
namespace dev {

// Interleaved complex value.  A plain aggregate (not HIP's vector type) so it
// can be loaded member-wise through address-space-qualified pointers.
template <typename T>
struct alignas(2 * sizeof(T)) Cx {
  T x, y;
};
template <typename T>
using V2 = Cx<T>;

template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  return C{a.x + b.x, a.y + b.y};
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  return C{a.x - b.x, a.y - b.y};
}
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C>
__device__ __forceinline__ C cconj(C a) {
  return C{a.x, -a.y};
}
// -i * a
template <typename C>
__device__ __forceinline__ C cmul_mi(C a) {
  return C{a.y, -a.x};
}
// +i * a
template <typename C>
__device__ __forceinline__ C cmul_pi(C a) {
  return C{-a.y, a.x};
}

// ---- compile-time trigonometry -------------------------------------------
struct CosSin {
  double c, s;
};

constexpr double kPiD = 3.141592653589793238462643383279502884;

// Taylor series on |x| <= pi/4 (terms up to x^31: far below double rounding).
constexpr double taylor_sin(double x) {
  double term = x, sum = x, x2 = x * x;
  for (int i = 1; i < 16; ++i) {
    term *= -x2 / static_cast<double>((2 * i) * (2 * i + 1));
    sum += term;
  }
  return sum;
}
constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0, x2 = x * x;
  for (int i = 1; i < 16; ++i) {
    term *= -x2 / static_cast<double>((2 * i - 1) * (2 * i));
    sum += term;
  }
  return sum;
}

// cos and sin of 2 pi m / n, octant-reduced so special angles are exact.
constexpr CosSin unit_root(long long m, long long n) {
  m %= n;
  if (m < 0) m += n;
  long long a = 8 * m;
  long long oct = a / n;
  long long rem = a - oct * n;
  long long r = (oct & 1) ? n - rem : rem;
  double t = kPiD * static_cast<double>(r) / (4.0 * static_cast<double>(n));
  double cr = taylor_cos(t), sr = taylor_sin(t);
  switch (oct) {
    case 0: return CosSin{cr, sr};
    case 1: return CosSin{sr, cr};
    case 2: return CosSin{-sr, cr};
    case 3: return CosSin{-cr, sr};
    case 4: return CosSin{-cr, -sr};
    case 5: return CosSin{-sr, -cr};
    case 6: return CosSin{sr, -cr};
    default: return CosSin{cr, -sr};
  }
}

constexpr bool is_prime_ct(int n) {
  if (n < 2) return false;
  for (int d = 2; d * d <= n; ++d)
    if (n % d == 0) return false;
  return true;
}

constexpr int first_split(int R) {
  if (R % 4 == 0 && R > 4) return 4;
  for (int d = 2; d < R; ++d)
    if (R % d == 0) return d;
  return R;
}

// ---- compile-time loop ----------------------------------------------------
template <typename F, int... Is>
__device__ __forceinline__ void sfor_impl(F &&f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void sfor(F &&f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// v * w_R^M (forward twiddle exp(-2 pi i M / R)) with trivial angles specialised.
template <int M, int R, typename C>
__device__ __forceinline__ C twiddle_ct(C v) {
  using T = decltype(v.x);
  constexpr int mm = ((M % R) + R) % R;
  constexpr T h = static_cast<T>(0.707106781186547524400844362104849039);
  if constexpr (mm == 0) {
    return v;
  } else if constexpr (4 * mm == R) {
    return cmul_mi(v);
  } else if constexpr (2 * mm == R) {
    return C{-v.x, -v.y};
  } else if constexpr (4 * mm == 3 * R) {
    return cmul_pi(v);
  } else if constexpr (8 * mm == R) {  // (1 - i)/sqrt2
    return C{(v.x + v.y) * h, (v.y - v.x) * h};
  } else if constexpr (8 * mm == 3 * R) {  // (-1 - i)/sqrt2
    return C{(v.y - v.x) * h, (-v.x - v.y) * h};
  } else if constexpr (8 * mm == 5 * R) {  // (-1 + i)/sqrt2
    return C{(-v.x - v.y) * h, (v.x - v.y) * h};
  } else if constexpr (8 * mm == 7 * R) {  // (1 + i)/sqrt2
    return C{(v.x - v.y) * h, (v.x + v.y) * h};
  } else {
    constexpr CosSin w = unit_root(mm, R);
    constexpr T c = static_cast<T>(w.c);
    constexpr T s = static_cast<T>(-w.s);
    return C{v.x * c - v.y * s, v.x * s + v.y * c};
  }
}

template <int R, typename C>
__device__ __forceinline__ void dft(C *v);

template <typename C>
__device__ __forceinline__ void dft2(C *v) {
  C a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <typename C>
__device__ __forceinline__ void dft4(C *v) {
  C s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
  C s1 = cadd(v[1], v[3]), d1 = csub(v[1], v[3]);
  v[0] = cadd(s0, s1);
  v[2] = csub(s0, s1);
  v[1] = cadd(d0, cmul_mi(d1));
  v[3] = csub(d0, cmul_mi(d1));
}

template <int R, typename C>
__device__ __forceinline__ void dft_prime(C *v) {
  using T = decltype(v[0].x);
  constexpr int H = (R - 1) / 2;
  C a[H + 1], b[H + 1];
  const C x0 = v[0];
  C sum = x0;
  sfor<H>([&](auto J) __attribute__((always_inline)) {
    constexpr int j = decltype(J)::value + 1;
    a[j] = cadd(v[j], v[R - j]);
    b[j] = csub(v[j], v[R - j]);
    sum = cadd(sum, a[j]);
  });
  sfor<H>([&](auto K) __attribute__((always_inline)) {
    constexpr int k = decltype(K)::value + 1;
    C A = x0;
    C B = C{0, 0};
    sfor<H>([&](auto J) __attribute__((always_inline)) {
      constexpr int j = decltype(J)::value + 1;
      constexpr CosSin w = unit_root(static_cast<long long>(j) * k, R);
      constexpr T c = static_cast<T>(w.c);
      constexpr T s = static_cast<T>(w.s);
      A.x += a[j].x * c;
      A.y += a[j].y * c;
      B.x += b[j].x * s;
      B.y += b[j].y * s;
    });
    // X_k = A - iB, X_{R-k} = A + iB
    v[k] = C{A.x + B.y, A.y - B.x};
    v[R - k] = C{A.x - B.y, A.y + B.x};
  });
  v[0] = sum;
}

// Cooley-Tukey split R = A*B: n = B*n1 + n2, k = k1 + A*k2.
template <int R, typename C>
__device__ __forceinline__ void dft_split(C *v) {
  constexpr int A = first_split(R);
  constexpr int B = R / A;
  C y[R];
  sfor<B>([&](auto N2) __attribute__((always_inline)) {
    constexpr int n2 = decltype(N2)::value;
    C t[A];
    sfor<A>([&](auto N1) __attribute__((always_inline)) {
      constexpr int n1 = decltype(N1)::value;
      t[n1] = v[B * n1 + n2];
    });
    dft<A>(t);
    sfor<A>([&](auto K1) __attribute__((always_inline)) {
      constexpr int k1 = decltype(K1)::value;
      y[n2 * A + k1] = twiddle_ct<n2 * k1, R>(t[k1]);
    });
  });
  sfor<A>([&](auto K1) __attribute__((always_inline)) {
    constexpr int k1 = decltype(K1)::value;
    C t[B];
    sfor<B>([&](auto N2) __attribute__((always_inline)) {
      constexpr int n2 = decltype(N2)::value;
      t[n2] = y[n2 * A + k1];
    });
    dft<B>(t);
    sfor<B>([&](auto K2) __attribute__((always_inline)) {
      constexpr int k2 = decltype(K2)::value;
      v[k1 + A * k2] = t[k2];
    });
  });
}

template <int R, typename C>
__device__ __forceinline__ void dft(C *v) {
  if constexpr (R == 1) {
    (void)v;
  } else if constexpr (R == 2) {
    dft2(v);
  } else if constexpr (R == 4) {
    dft4(v);
  } else if constexpr (is_prime_ct(R)) {
    dft_prime<R>(v);
  } else {
    dft_split<R>(v);
  }
}

}  // namespace dev
}  // namespace flagfft
