/*
 * oracle.c — CPU restatement of FlagFFT's transform path (TEST INFRASTRUCTURE).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, as the checker and as the timed CPU baseline.  It is never
 * linked into, loaded by or used as a fallback of the HIP product path.
 *
 * How each function is pinned (tests/test_oracle.py):
 *   ora_rng_*            bit-exact SplitMix64 outputs for fixed seeds (published
 *                        SplitMix64 test vector for seed 0) and moment checks.
 *   ora_dft_ld           closed forms: DFT(delta_j) = w^{jk}, DFT(1) = N delta_0,
 *                        DFT(w^{-mj}) = N delta_m; numpy pocketfft fixtures
 *                        (tests/golden) agree to 1e-15.
 *   ora_fft_c2c_{d,f}    vs ora_dft_ld for every N in 1..300 and the reference's
 *                        test sizes; golden fixtures; Parseval; roundtrip = N x.
 *   ora_fft_r2c/c2r      first N/2+1 bins of the complex DFT of the promoted
 *                        input; C2R(X) = Re IDFT(Hermitian(X)) with DC/Nyquist
 *                        imaginary parts dropped; R2C->C2R = N x; zero -> zero.
 *   ora_fft2d_*          separability against 1D DFTs of rows and columns;
 *                        numpy fft2/rfft2/irfft2 fixtures.
 *   ora_primitive_root / ora_rader_index_table
 *                        brute-force order check; permutation of 1..p-1.
 *   ora_bluestein_chirp  exact k^2 mod 2N in integers vs direct angle (k < 2^26).
 *   ora_rader_c2c_d / ora_bluestein_c2c_d
 *                        vs ora_dft_ld at the reference's prime sizes.
 *   ora_rel_l2           hand-computed small cases.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

/* ------------------------------------------------------------------------- */
/* threading helper                                                          */
/* ------------------------------------------------------------------------- */
typedef void (*range_fn)(void *ctx, int64_t begin, int64_t end);
typedef struct {
  range_fn fn;
  void *ctx;
  int64_t begin, end;
} range_job;

static void *range_trampoline(void *arg) {
  range_job *job = (range_job *)arg;
  if (job->begin < job->end) job->fn(job->ctx, job->begin, job->end);
  return NULL;
}

static int resolve_threads(int nthreads) {
  if (nthreads > 0) return nthreads;
  long cores = sysconf(_SC_NPROCESSORS_ONLN);
  return cores > 0 ? (int)cores : 1;
}

static void parallel_for(int64_t count, int nthreads, range_fn fn, void *ctx) {
  nthreads = resolve_threads(nthreads);
  if (nthreads > count) nthreads = (int)(count > 0 ? count : 1);
  if (nthreads <= 1) {
    fn(ctx, 0, count);
    return;
  }
  pthread_t *tids = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  range_job *jobs = (range_job *)malloc(sizeof(range_job) * (size_t)nthreads);
  int64_t chunk = (count + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    jobs[t].fn = fn;
    jobs[t].ctx = ctx;
    jobs[t].begin = t * chunk;
    jobs[t].end = (t + 1) * chunk < count ? (t + 1) * chunk : count;
    pthread_create(&tids[t], NULL, range_trampoline, &jobs[t]);
  }
  for (int t = 0; t < nthreads; ++t) pthread_join(tids[t], NULL);
  free(tids);
  free(jobs);
}

/* ------------------------------------------------------------------------- */
/* deterministic inputs                                                      */
/* ------------------------------------------------------------------------- */
static uint64_t splitmix_next(uint64_t *state) {
  *state += 0x9e3779b97f4a7c15ULL;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

void ora_rng_signed_unit(uint64_t seed, int64_t count, double *out) {
  uint64_t state = seed;
  const double scale = 1.0 / 9007199254740992.0; /* 2^-53 */
  for (int64_t i = 0; i < count; ++i) {
    uint64_t bits = splitmix_next(&state) >> 11;
    out[i] = ((double)bits * scale) * 2.0 - 1.0;
  }
}

void ora_rng_normal(uint64_t seed, int64_t count, double *out) {
  uint64_t state = seed;
  const double scale = 1.0 / 9007199254740992.0;
  const double two_pi = 6.283185307179586476925286766559;
  for (int64_t i = 0; i < count; i += 2) {
    double u1 = ((double)(splitmix_next(&state) >> 11) + 0.5) * scale;
    double u2 = (double)(splitmix_next(&state) >> 11) * scale;
    double r = sqrt(-2.0 * log(u1));
    out[i] = r * cos(two_pi * u2);
    if (i + 1 < count) out[i + 1] = r * sin(two_pi * u2);
  }
}

uint64_t ora_accuracy_seed(int type, int n, int batch, uint64_t variant) {
  return 0x4654464654455354ULL ^ ((uint64_t)(uint32_t)type << 48) ^ ((uint64_t)(uint32_t)n << 16) ^
         (uint64_t)(uint32_t)batch ^ (variant * 0x9e3779b97f4a7c15ULL);
}

/* ------------------------------------------------------------------------- */
/* naive long-double DFT                                                     */
/* ------------------------------------------------------------------------- */
static const long double kPiL = 3.141592653589793238462643383279502884L;

/* cos/sin(2*pi*m/n) for 0 <= m < n with octant symmetry so the argument
 * handed to cosl/sinl never exceeds pi/4 (exact special values). */
static void unit_root_ld(int64_t m, int64_t n, long double *c, long double *s) {
  /* angle = 2 pi m / n ; work in units of 1/(8n) turns: a = 8m in [0, 8n) */
  m %= n;
  if (m < 0) m += n;
  int64_t a = 8 * m; /* octant index = a / n */
  int64_t oct = a / n;
  int64_t rem = a - oct * n; /* angle within octant: rem/(8n) turn */
  long double cr, sr;
  long double t;
  int64_t r = rem;
  if (oct & 1) r = n - rem; /* mirror within odd octants */
  t = 2.0L * kPiL * (long double)r / (8.0L * (long double)n);
  cr = cosl(t);
  sr = sinl(t);
  long double x, y;
  switch (oct) {
    case 0: x = cr; y = sr; break;
    case 1: x = sr; y = cr; break;
    case 2: x = -sr; y = cr; break;
    case 3: x = -cr; y = sr; break;
    case 4: x = -cr; y = -sr; break;
    case 5: x = -sr; y = -cr; break;
    case 6: x = sr; y = -cr; break;
    default: x = cr; y = -sr; break;
  }
  *c = x;
  *s = y;
}

typedef struct {
  const double *in;
  double *out;
  int64_t n, batch;
  int sign;
  const long double *cs; /* interleaved cos, sin of 2 pi m / n */
} dft_ctx;

static void dft_rows(void *vctx, int64_t begin, int64_t end) {
  dft_ctx *c = (dft_ctx *)vctx;
  const int64_t n = c->n;
  for (int64_t work = begin; work < end; ++work) {
    int64_t b = work / n, k = work % n;
    const double *x = c->in + 2 * b * n;
    long double sr = 0.0L, si = 0.0L;
    int64_t idx = 0; /* j*k mod n, incrementally */
    for (int64_t j = 0; j < n; ++j) {
      long double wr = c->cs[2 * idx], wi = (long double)c->sign * c->cs[2 * idx + 1];
      long double xr = x[2 * j], xi = x[2 * j + 1];
      sr += xr * wr - xi * wi;
      si += xr * wi + xi * wr;
      idx += k;
      if (idx >= n) idx -= n;
    }
    c->out[2 * (b * n + k)] = (double)sr;
    c->out[2 * (b * n + k) + 1] = (double)si;
  }
}

void ora_dft_ld(const double *in, double *out, int64_t n, int64_t batch, int sign, int nthreads) {
  long double *cs = (long double *)malloc(sizeof(long double) * 2 * (size_t)n);
  for (int64_t m = 0; m < n; ++m) unit_root_ld(m, n, &cs[2 * m], &cs[2 * m + 1]);
  dft_ctx ctx = {in, out, n, batch, sign > 0 ? 1 : -1, cs};
  parallel_for(n * batch, nthreads, dft_rows, &ctx);
  free(cs);
}

/* ------------------------------------------------------------------------- */
/* fast FFT, instantiated for double and float                               */
/* ------------------------------------------------------------------------- */
#define ORA_REAL double
#define ORA_SFX(name) name##_d
#include "oracle_fft_impl.h"
#undef ORA_REAL
#undef ORA_SFX

#define ORA_REAL float
#define ORA_SFX(name) name##_f
#include "oracle_fft_impl.h"
#undef ORA_REAL
#undef ORA_SFX

/* ------------------------------------------------------------------------- */
/* 2D (double): rows (last axis) then columns                                */
/* ------------------------------------------------------------------------- */
typedef struct {
  double *data;
  int64_t n0, n1, ncols, batch;
  int sign;
} col_ctx;

static void col_pass(void *vctx, int64_t begin, int64_t end) {
  col_ctx *c = (col_ctx *)vctx;
  double *buf = (double *)malloc(sizeof(double) * 4 * (size_t)c->n0);
  for (int64_t work = begin; work < end; ++work) {
    int64_t b = work / c->ncols, col = work % c->ncols;
    double *base = c->data + 2 * b * c->n0 * c->ncols;
    for (int64_t r = 0; r < c->n0; ++r) {
      buf[2 * r] = base[2 * (r * c->ncols + col)];
      buf[2 * r + 1] = base[2 * (r * c->ncols + col) + 1];
    }
    ora_fft_c2c_d(buf, buf + 2 * c->n0, c->n0, 1, c->sign, 1);
    for (int64_t r = 0; r < c->n0; ++r) {
      base[2 * (r * c->ncols + col)] = buf[2 * c->n0 + 2 * r];
      base[2 * (r * c->ncols + col) + 1] = buf[2 * c->n0 + 2 * r + 1];
    }
  }
  free(buf);
}

int ora_fft2d_c2c_d(const double *in, double *out, int64_t n0, int64_t n1, int64_t batch, int sign, int nthreads) {
  int rc = ora_fft_c2c_d(in, out, n1, n0 * batch, sign, nthreads);
  if (rc) return rc;
  col_ctx ctx = {out, n0, n1, n1, batch, sign};
  parallel_for(batch * n1, nthreads, col_pass, &ctx);
  return 0;
}

int ora_fft2d_r2c_d(const double *in, double *out, int64_t n0, int64_t n1, int64_t batch, int nthreads) {
  const int64_t h = n1 / 2 + 1;
  int rc = ora_fft_r2c_d(in, out, n1, n0 * batch, n1, h, nthreads);
  if (rc) return rc;
  col_ctx ctx = {out, n0, n1, h, batch, -1};
  parallel_for(batch * h, nthreads, col_pass, &ctx);
  return 0;
}

int ora_fft2d_c2r_d(const double *in, double *out, int64_t n0, int64_t n1, int64_t batch, int nthreads) {
  const int64_t h = n1 / 2 + 1;
  double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)(n0 * h * batch));
  memcpy(tmp, in, sizeof(double) * 2 * (size_t)(n0 * h * batch));
  col_ctx ctx = {tmp, n0, n1, h, batch, 1};
  parallel_for(batch * h, nthreads, col_pass, &ctx);
  /* row C2R drops Im of each row's DC/Nyquist bin, as the reference's
   * row expand kernel does after the column pass (src/exec/raw_nodes.cpp:770-852). */
  int rc = ora_fft_c2r_d(tmp, out, n1, n0 * batch, h, n1, nthreads);
  free(tmp);
  return rc;
}

/* ------------------------------------------------------------------------- */
/* Rader / Bluestein restatements                                            */
/* ------------------------------------------------------------------------- */
static int64_t modpow_i64(int64_t base, int64_t e, int64_t mod) {
  __int128 result = 1, b = base % mod;
  while (e > 0) {
    if (e & 1) result = (result * b) % mod;
    b = (b * b) % mod;
    e >>= 1;
  }
  return (int64_t)result;
}

static int is_prime_i64(int64_t n) {
  if (n < 2) return 0;
  if (n % 2 == 0) return n == 2;
  for (int64_t d = 3; d * d <= n; d += 2)
    if (n % d == 0) return 0;
  return 1;
}

/* Smallest primitive root (src/plan/rader_utils.cpp:55-77). Returns -1 if p is not prime. */
int64_t ora_primitive_root(int64_t p) {
  if (!is_prime_i64(p)) return -1;
  if (p == 2) return 1;
  int64_t phi = p - 1, rest = phi;
  int64_t factors[64];
  int nf = 0;
  for (int64_t f = 2; f * f <= rest; ++f) {
    if (rest % f) continue;
    factors[nf++] = f;
    while (rest % f == 0) rest /= f;
  }
  if (rest > 1) factors[nf++] = rest;
  for (int64_t g = 2; g < p; ++g) {
    int ok = 1;
    for (int i = 0; i < nf && ok; ++i)
      if (modpow_i64(g, phi / factors[i], p) == 1) ok = 0;
    if (ok) return g;
  }
  return -1;
}

/* idx[i] = g^i mod p, i < p-1 (src/plan/rader_utils.cpp:79-90). */
void ora_rader_index_table(int64_t p, int64_t g, int64_t *idx) {
  int64_t v = 1;
  for (int64_t i = 0; i < p - 1; ++i) {
    idx[i] = v;
    v = (int64_t)(((__int128)v * g) % p);
  }
}

/* chirp[k] = exp(sign*i*pi*(k^2 mod 2n)/n) (src/codegen/tables.cpp:142-168),
 * with k^2 mod 2n reduced exactly in 128-bit integers. */
void ora_bluestein_chirp(int64_t n, int sign, double *out) {
  for (int64_t k = 0; k < n; ++k) {
    int64_t r = (int64_t)(((__int128)k * k) % (2 * n));
    long double c, s;
    unit_root_ld(r, 2 * n, &c, &s); /* angle 2 pi r / (2n) = pi r / n */
    out[2 * k] = (double)c;
    out[2 * k + 1] = (double)(sign > 0 ? s : -s);
  }
}

static void cmul_ld(long double ar, long double ai, long double br, long double bi, long double *cr, long double *ci) {
  *cr = ar * br - ai * bi;
  *ci = ar * bi + ai * br;
}

/* Rader following src/exec/raw_nodes.cpp:253-305 and the kernels of
 * python/flagfft_codegen/jit_source.py:270-388:
 *   b[q] = w_p^{sign*idx[q]},  B = FFT_m(b)            (tables.cpp:224-258)
 *   a[q] = x[idx[(m-q) mod m]], A = FFT_m(a)           (prepare)
 *   t = conj(A * B), W = FFT_m(t)                      (pointwise + forward FFT)
 *   X[idx[q]] = x0 + conj(W[q]) / m,  X[0] = sum x     (finalize)
 * FFT_m is the double fast FFT (forward, as the reference's children are). */
int ora_rader_c2c_d(const double *in, double *out, int64_t p, int64_t batch, int sign) {
  if (!is_prime_i64(p) || p < 3) return 1;
  const int64_t m = p - 1;
  int64_t g = ora_primitive_root(p);
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)m);
  ora_rader_index_table(p, g, idx);
  double *b = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  double *bh = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  double *a = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  double *ah = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  for (int64_t q = 0; q < m; ++q) {
    long double c, s;
    unit_root_ld(idx[q], p, &c, &s);
    b[2 * q] = (double)c;
    b[2 * q + 1] = (double)(sign > 0 ? s : -s);
  }
  ora_fft_c2c_d(b, bh, m, 1, -1, 1);
  for (int64_t bt = 0; bt < batch; ++bt) {
    const double *x = in + 2 * bt * p;
    double *y = out + 2 * bt * p;
    for (int64_t q = 0; q < m; ++q) {
      int64_t src = idx[q == 0 ? 0 : m - q];
      a[2 * q] = x[2 * src];
      a[2 * q + 1] = x[2 * src + 1];
    }
    ora_fft_c2c_d(a, ah, m, 1, -1, 1);
    for (int64_t q = 0; q < m; ++q) {
      long double pr, pi;
      cmul_ld(ah[2 * q], ah[2 * q + 1], bh[2 * q], bh[2 * q + 1], &pr, &pi);
      a[2 * q] = (double)pr;
      a[2 * q + 1] = (double)(-pi);
    }
    ora_fft_c2c_d(a, ah, m, 1, -1, 1);
    long double sr = 0.0L, si = 0.0L;
    for (int64_t j = 0; j < p; ++j) {
      sr += x[2 * j];
      si += x[2 * j + 1];
    }
    double x0r = x[0], x0i = x[1];
    for (int64_t q = 0; q < m; ++q) {
      y[2 * idx[q]] = x0r + ah[2 * q] / (double)m;
      y[2 * idx[q] + 1] = x0i - ah[2 * q + 1] / (double)m;
    }
    y[0] = (double)sr;
    y[1] = (double)si;
  }
  free(idx);
  free(b);
  free(bh);
  free(a);
  free(ah);
  return 0;
}

/* Bluestein following src/exec/raw_nodes.cpp:153-204 with the tables of
 * src/codegen/tables.cpp:142-215 and kernels of jit_source.py:125-228:
 *   chirp[k] = exp(-+ i pi k^2/n)  (minus for forward), b[k] = conj-sign chirp, b mirrored to m-k
 *   a = x*chirp zero-padded to m; A = FFT_m(a); t = conj(A*B); W = FFT_m(t)
 *   y[k] = conj(W[k])/m * chirp[k]                                          */
int ora_bluestein_c2c_d(const double *in, double *out, int64_t n, int64_t m, int64_t batch, int sign) {
  if (m < 2 * n - 1) return 1;
  double *chirp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  double *bchirp = (double *)malloc(sizeof(double) * 2 * (size_t)n);
  ora_bluestein_chirp(n, sign > 0 ? 1 : -1, chirp);
  ora_bluestein_chirp(n, sign > 0 ? -1 : 1, bchirp);
  double *b = (double *)calloc(2 * (size_t)m, sizeof(double));
  double *bh = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  double *a = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  double *ah = (double *)malloc(sizeof(double) * 2 * (size_t)m);
  for (int64_t k = 0; k < n; ++k) {
    b[2 * k] = bchirp[2 * k];
    b[2 * k + 1] = bchirp[2 * k + 1];
    if (k) {
      b[2 * (m - k)] = bchirp[2 * k];
      b[2 * (m - k) + 1] = bchirp[2 * k + 1];
    }
  }
  ora_fft_c2c_d(b, bh, m, 1, -1, 1);
  for (int64_t bt = 0; bt < batch; ++bt) {
    const double *x = in + 2 * bt * n;
    double *y = out + 2 * bt * n;
    for (int64_t j = 0; j < m; ++j) {
      if (j < n) {
        long double pr, pi;
        cmul_ld(x[2 * j], x[2 * j + 1], chirp[2 * j], chirp[2 * j + 1], &pr, &pi);
        a[2 * j] = (double)pr;
        a[2 * j + 1] = (double)pi;
      } else {
        a[2 * j] = 0.0;
        a[2 * j + 1] = 0.0;
      }
    }
    ora_fft_c2c_d(a, ah, m, 1, -1, 1);
    for (int64_t j = 0; j < m; ++j) {
      long double pr, pi;
      cmul_ld(ah[2 * j], ah[2 * j + 1], bh[2 * j], bh[2 * j + 1], &pr, &pi);
      a[2 * j] = (double)pr;
      a[2 * j + 1] = (double)(-pi);
    }
    ora_fft_c2c_d(a, ah, m, 1, -1, 1);
    for (int64_t k = 0; k < n; ++k) {
      long double wr = ah[2 * k] / (long double)m, wi = -ah[2 * k + 1] / (long double)m;
      long double pr, pi;
      cmul_ld(wr, wi, chirp[2 * k], chirp[2 * k + 1], &pr, &pi);
      y[2 * k] = (double)pr;
      y[2 * k + 1] = (double)pi;
    }
  }
  free(chirp);
  free(bchirp);
  free(b);
  free(bh);
  free(a);
  free(ah);
  return 0;
}

/* ------------------------------------------------------------------------- */
/* error metrics (ctest/flagfft_test.h:222-282)                              */
/* ------------------------------------------------------------------------- */
#define ORA_METRIC_BODY(VALUE_T)                                                        \
  long double worst_l2 = 0.0L, worst_linf = 0.0L;                                        \
  for (int64_t b = 0; b < batch; ++b) {                                                  \
    long double err_sq = 0.0L, ref_sq = 0.0L, err_max = 0.0L, ref_max = 0.0L;            \
    for (int64_t i = 0; i < elems; ++i) {                                                \
      int64_t at = b * elems + i;                                                        \
      long double d, r;                                                                  \
      if (cplx) {                                                                        \
        long double dr = (long double)value[2 * at] - (long double)ref[2 * at];          \
        long double di = (long double)value[2 * at + 1] - (long double)ref[2 * at + 1];  \
        d = sqrtl(dr * dr + di * di);                                                    \
        r = sqrtl((long double)ref[2 * at] * ref[2 * at] +                               \
                  (long double)ref[2 * at + 1] * ref[2 * at + 1]);                       \
      } else {                                                                           \
        d = fabsl((long double)value[at] - (long double)ref[at]);                        \
        r = fabsl((long double)ref[at]);                                                 \
      }                                                                                  \
      if (!isfinite(d) || !isfinite(r)) return INFINITY;                                 \
      err_sq += d * d;                                                                   \
      ref_sq += r * r;                                                                   \
      if (d > err_max) err_max = d;                                                      \
      if (r > ref_max) ref_max = r;                                                      \
    }                                                                                    \
    long double l2 = ref_sq == 0.0L ? (err_sq == 0.0L ? 0.0L : INFINITY) : sqrtl(err_sq / ref_sq); \
    long double linf = ref_max == 0.0L ? (err_max == 0.0L ? 0.0L : INFINITY) : err_max / ref_max;  \
    if (l2 > worst_l2) worst_l2 = l2;                                                    \
    if (linf > worst_linf) worst_linf = linf;                                            \
  }                                                                                      \
  if (rel_linf) *rel_linf = (double)worst_linf;                                          \
  return (double)worst_l2;

double ora_rel_l2(const double *value, const double *ref, int64_t elems, int64_t batch, int cplx, double *rel_linf) {
  ORA_METRIC_BODY(double)
}

double ora_rel_l2_f(const float *value, const double *ref, int64_t elems, int64_t batch, int cplx, double *rel_linf) {
  ORA_METRIC_BODY(float)
}

double ora_work_factor(int64_t n) {
  if (n <= 64) return (double)n;
  int64_t v = 2 * n - 1, covered = 1;
  int p = 0;
  while (covered < v) {
    covered <<= 1;
    ++p;
  }
  return 3.0 * p + 3.0;
}
