// kernels.hpp — host-side launchers of the gfx950 kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fft_args.hpp"

namespace flagfft {

struct LaunchShape {
  dim3 grid;
  dim3 block;
  size_t lds_bytes;
};

// Elements-per-thread variants compiled for every kernel family.
bool supported_ept(bool dbl, int ept);  // fp32: 8, 16, 32; fp64: 8, 16
constexpr int kMaxBlockThreads = 512;

hipError_t launch_pass(const PassArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st);
hipError_t launch_rader(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st);
hipError_t launch_bluestein(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st);
hipError_t launch_r2c_even(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st);
hipError_t launch_c2r_even(const FusedArgs &a, bool dbl, int ept, const LaunchShape &s, hipStream_t st);
// Specialised single-pass kernels: id of the compiled radix sequence (-1 if none).
int leaf_kernel_id(const std::vector<int> &radices);
int leaf_kernel_ept(int id);
hipError_t launch_leaf(const LeafArgs &a, bool dbl, int id, const LaunchShape &s, hipStream_t st);

hipError_t launch_real_post(const RealPassArgs &a, bool dbl, hipStream_t st);
hipError_t launch_real_pre(const RealPassArgs &a, bool dbl, hipStream_t st);

// ---- table generation on the GPU, FP64 arithmetic ------------------------
// out[j] = w_n^{(j*mult) mod n} = exp(-2 pi i ((j*mult) mod n) / n), j < count.
hipError_t gen_roots(void *out, bool dbl, int64_t count, int64_t n, int64_t mult, hipStream_t st);
// out[k] = exp(-i pi (k^2 mod 2n) / n), k < n.
hipError_t gen_chirp(void *out, bool dbl, int64_t n, hipStream_t st);
// FP64 Bluestein kernel b[k] = exp(+i pi (k^2 mod 2n)/n) for k < n, mirrored to M-k, zero elsewhere.
hipError_t gen_bluestein_b(double2 *out, int64_t n, int64_t m, hipStream_t st);
// FP64 Rader kernel b[q] = w_p^{idx[q]}, q < p-1.
hipError_t gen_rader_b(double2 *out, const int32_t *idx, int64_t p, hipStream_t st);
// out[j] = in[j] * scale, cast to the plan precision.
hipError_t scale_cast(const double2 *in, void *out, bool dbl, int64_t count, double scale, hipStream_t st);

}  // namespace flagfft
