/*
 * oracle.h — CPU restatement of FlagFFT's transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (flagfft_amd/libflagfft.so).  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product never links, loads or
 * falls back to it.
 *
 * Conventions pinned by the reference (citations relative to the reference
 * repository root):
 *   - forward sign exp(-2*pi*i*j*k/N), FLAGFFT_FORWARD = -1   include/flagfft.h:37
 *   - inverse is unnormalised (roundtrip = N*x)               ctest/test_exec_c2c_inv_bs_s.cpp:73-91
 *   - R2C emits N/2+1 bins                                     python/flagfft_codegen/kernels.py:1127-1156
 *   - C2R ignores Im of the DC and (even N) Nyquist input bins  python/flagfft_codegen/kernels.py:1159-1192
 *   - 2D row-major: row (last axis) transform then columns     src/exec/raw_nodes.cpp:549-610
 *   - Rader primitive root = smallest, idx[i] = g^i mod p       src/plan/rader_utils.cpp:55-90
 *   - Bluestein chirp exp(sign*i*pi*(k^2 mod 2N)/N), b mirrored src/codegen/tables.cpp:142-215
 *
 * Complex data is interleaved (re, im) like flagfftComplex / flagfftDoubleComplex.
 * Every transform takes `sign` = -1 (forward) or +1 (inverse, unnormalised).
 *
 * Parity status of each function (how it is pinned) is listed in the header
 * of oracle.c and checked by tests/test_oracle.py.
 */
#ifndef FLAGFFT_ORACLE_H
#define FLAGFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- deterministic inputs ------------------------------------------------ */
/* SplitMix64 signed_unit in [-1, 1): ctest/flagfft_test.h:390-411. */
void ora_rng_signed_unit(uint64_t seed, int64_t count, double *out);
/* Same stream, Box-Muller N(0,1): u1 = (a + 0.5) * 2^-53, u2 = b * 2^-53
 * (a, b = next() >> 11), z0 = r cos(2 pi u2), z1 = r sin(2 pi u2). */
void ora_rng_normal(uint64_t seed, int64_t count, double *out);
/* accuracy_seed(type, n, batch, variant): ctest/flagfft_test.h:413-417. */
uint64_t ora_accuracy_seed(int type, int n, int batch, uint64_t variant);

/* ---- definitions (long double, O(N^2)) ----------------------------------- */
/* Naive DFT of `batch` contiguous length-n complex rows, long double accumulation
 * with exact (j*k mod n) angle reduction.  Threads over rows/outputs. */
void ora_dft_ld(const double *in, double *out, int64_t n, int64_t batch, int sign, int nthreads);

/* ---- fast FFT (mixed-radix Stockham + Bluestein), the CPU baseline ------- */
/* Contiguous batched C2C.  nthreads <= 0 means one per online core. */
int ora_fft_c2c_d(const double *in, double *out, int64_t n, int64_t batch, int sign, int nthreads);
int ora_fft_c2c_f(const float *in, float *out, int64_t n, int64_t batch, int sign, int nthreads);

/* R2C: real rows of length n (distance idist reals) -> n/2+1 bins (distance odist complex). */
int ora_fft_r2c_d(const double *in, double *out, int64_t n, int64_t batch, int64_t idist, int64_t odist, int nthreads);
int ora_fft_r2c_f(const float *in, float *out, int64_t n, int64_t batch, int64_t idist, int64_t odist, int nthreads);
/* C2R: n/2+1 bins (distance idist complex) -> n reals (distance odist reals); Im of
 * bin 0 and (even n) bin n/2 are ignored; unnormalised. */
int ora_fft_c2r_d(const double *in, double *out, int64_t n, int64_t batch, int64_t idist, int64_t odist, int nthreads);
int ora_fft_c2r_f(const float *in, float *out, int64_t n, int64_t batch, int64_t idist, int64_t odist, int nthreads);

/* 2D row-major n0 x n1 (n1 contiguous), batch contiguous. */
int ora_fft2d_c2c_d(const double *in, double *out, int64_t n0, int64_t n1, int64_t batch, int sign, int nthreads);
int ora_fft2d_r2c_d(const double *in, double *out, int64_t n0, int64_t n1, int64_t batch, int nthreads);
int ora_fft2d_c2r_d(const double *in, double *out, int64_t n0, int64_t n1, int64_t batch, int nthreads);

/* ---- restatements of the reference's prime-length algorithms ------------- */
int64_t ora_primitive_root(int64_t p);                       /* src/plan/rader_utils.cpp:55-77 */
void ora_rader_index_table(int64_t p, int64_t g, int64_t *idx); /* src/plan/rader_utils.cpp:79-90 */
/* Bluestein chirp c[k] = exp(sign*i*pi*(k^2 mod 2n)/n), k < n: src/codegen/tables.cpp:142-168 */
void ora_bluestein_chirp(int64_t n, int sign, double *out);
/* Rader (p prime) and Bluestein (conv length m >= 2n-1) C2C following the
 * reference's step order (src/exec/raw_nodes.cpp:153-305). */
int ora_rader_c2c_d(const double *in, double *out, int64_t p, int64_t batch, int sign);
int ora_bluestein_c2c_d(const double *in, double *out, int64_t n, int64_t m, int64_t batch, int sign);

/* ---- error metrics: ctest/flagfft_test.h:222-320 ------------------------- */
/* Worst-row normwise relative L2 / Linf over `batch` rows of `elems` values
 * (`cplx` = 1 for interleaved complex).  Returns rel_l2; *rel_linf optional. */
double ora_rel_l2(const double *value, const double *ref, int64_t elems, int64_t batch, int cplx, double *rel_linf);
double ora_rel_l2_f(const float *value, const double *ref, int64_t elems, int64_t batch, int cplx, double *rel_linf);
/* work(N): N for N <= 64, else 3*ceil_log2(2N-1)+3 (ctest/flagfft_test.h:274-282). */
double ora_work_factor(int64_t n);

#ifdef __cplusplus
}
#endif
#endif
