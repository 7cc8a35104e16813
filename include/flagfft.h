/*
 * flagfft.h — public C API of the MI355X-native FlagFFT library.
 *
 * Drop-in for the reference's include/flagfft.h:1-86: the same 13 extern "C"
 * symbols, enum values (flagfftType equal to cuFFT's), struct layouts and
 * error semantics.  flagfftStream_t carries a hipStream_t.
 *
 * Semantics (reference anchors):
 *   - forward = exp(-2 pi i jk/N) (FLAGFFT_FORWARD = -1); inverse unnormalised
 *   - R2C writes N/2+1 bins; C2R ignores Im of the DC / Nyquist input bins
 *   - Exec* are asynchronous on the stream set by flagfftSetStream (default: null stream)
 *   - Plan* null the handle first; sizes <= 0 -> FLAGFFT_INVALID_SIZE,
 *     null handle pointer / rank <= 0 / batch <= 0 -> FLAGFFT_INVALID_VALUE,
 *     unsupported layouts and rank 3 -> FLAGFFT_NOT_SUPPORTED
 *     (src/exec/c_api_plan.cpp:274-336, src/exec/c_api_internal.cpp:60-120)
 */
#pragma once

#ifdef __cplusplus
extern "C" {
#endif

/* Opaque stream handle: pass a hipStream_t (or NULL for the null stream). */
typedef void *flagfftStream_t;

typedef enum flagfftResult_t {
  FLAGFFT_SUCCESS = 0,
  FLAGFFT_INVALID_PLAN = 1,
  FLAGFFT_ALLOC_FAILED = 2,
  FLAGFFT_INVALID_TYPE = 3,
  FLAGFFT_INVALID_VALUE = 4,
  FLAGFFT_INTERNAL_ERROR = 5,
  FLAGFFT_EXEC_FAILED = 6,
  FLAGFFT_SETUP_FAILED = 7,
  FLAGFFT_INVALID_SIZE = 8,
  FLAGFFT_UNALIGNED_DATA = 9,
  FLAGFFT_INCOMPLETE_PARAMETER_LIST = 10,
  FLAGFFT_INVALID_DEVICE = 11,
  FLAGFFT_PARSE_ERROR = 12,
  FLAGFFT_NO_WORKSPACE = 13,
  FLAGFFT_NOT_SUPPORTED = 14,
} flagfftResult;

typedef enum flagfftType_t {
  FLAGFFT_R2C = 0x2a,
  FLAGFFT_C2R = 0x2c,
  FLAGFFT_C2C = 0x29,
  FLAGFFT_D2Z = 0x6a,
  FLAGFFT_Z2D = 0x6c,
  FLAGFFT_Z2Z = 0x69,
} flagfftType;

#define FLAGFFT_FORWARD (-1)
#define FLAGFFT_INVERSE 1

typedef struct flagfftComplex_t {
  float x;
  float y;
} flagfftComplex;

typedef struct flagfftDoubleComplex_t {
  double x;
  double y;
} flagfftDoubleComplex;

typedef float flagfftReal;
typedef double flagfftDoubleReal;

typedef struct flagfftPlan_t *flagfftHandle;

/* reference: include/flagfft.h:55 / src/exec/c_api_plan.cpp:249-257 */
flagfftResult flagfftPlan1d(flagfftHandle *plan, int nx, flagfftType type, int batch);
/* reference: include/flagfft.h:56 / src/exec/c_api_plan.cpp:259-267 (batch forced to 1) */
flagfftResult flagfftPlan2d(flagfftHandle *plan, int nx, int ny, flagfftType type);
/* reference: include/flagfft.h:57 / src/exec/c_api_plan.cpp:269-272 (always NOT_SUPPORTED) */
flagfftResult flagfftPlan3d(flagfftHandle *plan, int nx, int ny, int nz, flagfftType type);
/* reference: include/flagfft.h:58-68 / src/exec/c_api_plan.cpp:274-336 */
flagfftResult flagfftPlanMany(flagfftHandle *plan,
                              int rank,
                              int *n,
                              int *inembed,
                              int istride,
                              int idist,
                              int *onembed,
                              int ostride,
                              int odist,
                              flagfftType type,
                              int batch);

/* reference: include/flagfft.h:70-78 / src/exec/c_api_exec.cpp:5-167 */
flagfftResult flagfftExecC2C(flagfftHandle plan, flagfftComplex *idata, flagfftComplex *odata, int direction);
flagfftResult flagfftExecZ2Z(flagfftHandle plan,
                             flagfftDoubleComplex *idata,
                             flagfftDoubleComplex *odata,
                             int direction);
flagfftResult flagfftExecR2C(flagfftHandle plan, flagfftReal *idata, flagfftComplex *odata);
flagfftResult flagfftExecD2Z(flagfftHandle plan, flagfftDoubleReal *idata, flagfftDoubleComplex *odata);
flagfftResult flagfftExecC2R(flagfftHandle plan, flagfftComplex *idata, flagfftReal *odata);
flagfftResult flagfftExecZ2D(flagfftHandle plan, flagfftDoubleComplex *idata, flagfftDoubleReal *odata);

/* reference: include/flagfft.h:80-82 / src/exec/c_api_exec.cpp:169-236 */
flagfftResult flagfftSetStream(flagfftHandle plan, flagfftStream_t stream);
flagfftResult flagfftDestroy(flagfftHandle plan);
const char *flagfftGetPlanDescription(flagfftHandle plan);

#ifdef __cplusplus
}
#endif
