/*
 * A C program that uses the library the way the reference's own C++ tests
 * and flagfft-cli do: through include/flagfft.h only, with device buffers
 * from the HIP runtime (ctest/test_plan.cpp:26-61 for the contract checks,
 * ctest/test_exec_c2c_inv_bs_s.cpp:73-91 and test_exec_r2c_c2r_ct_b.cpp:54-70
 * for the roundtrips).  Built by __graft_entry__.build() with gcc.
 *
 *   abi_caller --no-device   argument checks, then plan creation must stop at
 *                            FLAGFFT_INVALID_DEVICE (no GPU)
 *   abi_caller               argument checks, then C2C 4096 x 4 and R2C/C2R
 *                            6000 x 2 on the GPU: delta -> ones, roundtrip = N x
 * Exit status 0 and a last line "abi caller ok" on success.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <hip/hip_runtime_api.h>

#include "flagfft.h"

static int failures = 0;
#define EXPECT(cond, what)                                           \
  do {                                                               \
    if (!(cond)) {                                                   \
      fprintf(stderr, "FAILED: %s (%s:%d)\n", what, __FILE__, __LINE__); \
      ++failures;                                                    \
    }                                                                \
  } while (0)

static double signed_unit(unsigned long long *s) {
  /* SplitMix64, mapped to [-1, 1) */
  unsigned long long z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return (double)(z >> 11) * (1.0 / 4503599627370496.0) - 1.0;
}

static void contract_checks(void) {
  flagfftHandle plan = (flagfftHandle)0x1;
  EXPECT(flagfftPlan1d(&plan, 0, FLAGFFT_C2C, 1) == FLAGFFT_INVALID_SIZE && plan == NULL, "Plan1d n=0");
  EXPECT(flagfftPlan1d(NULL, 64, FLAGFFT_C2C, 1) == FLAGFFT_INVALID_VALUE, "Plan1d null out");
  EXPECT(flagfftPlan1d(&plan, 64, FLAGFFT_C2C, 0) == FLAGFFT_INVALID_VALUE, "Plan1d batch 0");
  EXPECT(flagfftPlan2d(&plan, 0, 32, FLAGFFT_C2C) == FLAGFFT_INVALID_SIZE && plan == NULL, "Plan2d n=0");
  plan = (flagfftHandle)0x1;
  EXPECT(flagfftPlan3d(&plan, 32, 16, 8, FLAGFFT_C2C) == FLAGFFT_NOT_SUPPORTED && plan == NULL, "Plan3d");
  EXPECT(flagfftDestroy(NULL) == FLAGFFT_INVALID_PLAN, "Destroy null");
  EXPECT(flagfftExecC2C(NULL, NULL, NULL, FLAGFFT_FORWARD) == FLAGFFT_INVALID_PLAN, "Exec null plan");
  EXPECT(flagfftGetPlanDescription(NULL) == NULL, "description of null");
  EXPECT(sizeof(flagfftComplex) == 8 && sizeof(flagfftDoubleComplex) == 16, "struct sizes");
}

static int c2c_roundtrip(hipStream_t st) {
  enum { N = 4096, B = 4 };
  flagfftHandle plan = NULL;
  if (flagfftPlan1d(&plan, N, FLAGFFT_C2C, B) != FLAGFFT_SUCCESS) {
    EXPECT(0, "Plan1d C2C");
    return 1;
  }
  const char *desc = flagfftGetPlanDescription(plan);
  EXPECT(desc != NULL && strlen(desc) > 0, "plan description");
  EXPECT(flagfftSetStream(plan, (flagfftStream_t)st) == FLAGFFT_SUCCESS, "SetStream");
  size_t bytes = sizeof(flagfftComplex) * N * B;
  flagfftComplex *x = malloc(bytes), *y = malloc(bytes), *z = malloc(bytes);
  unsigned long long seed = 12345;
  for (int i = 0; i < N * B; ++i) {
    x[i].x = (float)signed_unit(&seed);
    x[i].y = (float)signed_unit(&seed);
  }
  for (int i = 0; i < N; ++i) x[i].x = x[i].y = 0.0f; /* row 0: delta -> all ones */
  x[0].x = 1.0f;
  flagfftComplex *dx = NULL, *dy = NULL, *dz = NULL;
  EXPECT(hipMalloc((void **)&dx, bytes) == hipSuccess && hipMalloc((void **)&dy, bytes) == hipSuccess &&
             hipMalloc((void **)&dz, bytes) == hipSuccess, "hipMalloc");
  EXPECT(hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) == hipSuccess, "copy in");
  EXPECT(flagfftExecC2C(plan, dx, dy, FLAGFFT_FORWARD) == FLAGFFT_SUCCESS, "ExecC2C forward");
  EXPECT(flagfftExecC2C(plan, dy, dz, FLAGFFT_INVERSE) == FLAGFFT_SUCCESS, "ExecC2C inverse");
  EXPECT(flagfftExecC2C(plan, dy, dz, 7) == FLAGFFT_INVALID_VALUE, "bad direction");
  EXPECT(flagfftExecR2C(plan, (flagfftReal *)dx, dy) == FLAGFFT_INVALID_TYPE, "wrong exec for the type");
  EXPECT(hipStreamSynchronize(st) == hipSuccess, "sync");
  EXPECT(hipMemcpy(y, dy, bytes, hipMemcpyDeviceToHost) == hipSuccess, "copy out");
  EXPECT(hipMemcpy(z, dz, bytes, hipMemcpyDeviceToHost) == hipSuccess, "copy out");
  double dmax = 0, num = 0, den = 0;
  for (int k = 0; k < N; ++k) dmax = fmax(dmax, fabs(y[k].x - 1.0) + fabs(y[k].y));
  for (int i = 0; i < N * B; ++i) {
    double ex = (double)N * x[i].x, ey = (double)N * x[i].y;
    num += (z[i].x - ex) * (z[i].x - ex) + (z[i].y - ey) * (z[i].y - ey);
    den += ex * ex + ey * ey;
  }
  printf("C2C %d x %d: delta max err %.3e, roundtrip rel L2 %.3e\n", N, B, dmax, sqrt(num / den));
  EXPECT(dmax < 1e-5, "delta -> ones");
  EXPECT(sqrt(num / den) < 2e-6 * 12, "roundtrip = N x");
  EXPECT(flagfftDestroy(plan) == FLAGFFT_SUCCESS, "Destroy");
  EXPECT(flagfftExecC2C(plan, dx, dy, FLAGFFT_FORWARD) == FLAGFFT_INVALID_PLAN, "destroyed handle");
  hipFree(dx);
  hipFree(dy);
  hipFree(dz);
  free(x);
  free(y);
  free(z);
  return 0;
}

static int real_roundtrip(hipStream_t st) {
  enum { N = 6000, B = 2, H = N / 2 + 1 };
  flagfftHandle fwd = NULL, inv = NULL;
  EXPECT(flagfftPlan1d(&fwd, N, FLAGFFT_R2C, B) == FLAGFFT_SUCCESS, "Plan1d R2C");
  EXPECT(flagfftPlan1d(&inv, N, FLAGFFT_C2R, B) == FLAGFFT_SUCCESS, "Plan1d C2R");
  if (!fwd || !inv) return 1;
  flagfftSetStream(fwd, (flagfftStream_t)st);
  flagfftSetStream(inv, (flagfftStream_t)st);
  float *x = malloc(sizeof(float) * N * B), *z = malloc(sizeof(float) * N * B);
  unsigned long long seed = 777;
  for (int i = 0; i < N * B; ++i) x[i] = (float)signed_unit(&seed);
  float *dx = NULL, *dz = NULL;
  flagfftComplex *dX = NULL;
  EXPECT(hipMalloc((void **)&dx, sizeof(float) * N * B) == hipSuccess &&
             hipMalloc((void **)&dz, sizeof(float) * N * B) == hipSuccess &&
             hipMalloc((void **)&dX, sizeof(flagfftComplex) * H * B) == hipSuccess, "hipMalloc");
  hipMemcpy(dx, x, sizeof(float) * N * B, hipMemcpyHostToDevice);
  EXPECT(flagfftExecR2C(fwd, dx, dX) == FLAGFFT_SUCCESS, "ExecR2C");
  EXPECT(flagfftExecC2R(inv, dX, dz) == FLAGFFT_SUCCESS, "ExecC2R");
  hipStreamSynchronize(st);
  hipMemcpy(z, dz, sizeof(float) * N * B, hipMemcpyDeviceToHost);
  double num = 0, den = 0;
  for (int i = 0; i < N * B; ++i) {
    double e = (double)N * x[i];
    num += (z[i] - e) * (z[i] - e);
    den += e * e;
  }
  printf("R2C/C2R %d x %d: roundtrip rel L2 %.3e\n", N, B, sqrt(num / den));
  EXPECT(sqrt(num / den) < 2e-6 * 13 * 2, "R2C -> C2R = N x");
  flagfftDestroy(fwd);
  flagfftDestroy(inv);
  hipFree(dx);
  hipFree(dz);
  hipFree(dX);
  free(x);
  free(z);
  return 0;
}

int main(int argc, char **argv) {
  const int no_device = argc > 1 && strcmp(argv[1], "--no-device") == 0;
  contract_checks();
  if (no_device) {
    flagfftHandle plan = (flagfftHandle)0x1;
    EXPECT(flagfftPlan1d(&plan, 4096, FLAGFFT_C2C, 4) == FLAGFFT_INVALID_DEVICE && plan == NULL, "no device");
  } else {
    hipStream_t st;
    EXPECT(hipStreamCreate(&st) == hipSuccess, "hipStreamCreate");
    c2c_roundtrip(st);
    real_roundtrip(st);
    hipStreamDestroy(st);
  }
  if (failures) {
    fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  printf("abi caller ok\n");
  return 0;
}
