// kernels.hpp — host-side launchers of the gfx950 kernels in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
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

// Host handles of the kernels every plan step launches (nullptr when no such
// instance is compiled).  A step is launched with hipLaunchKernel(handle,
// grid, block, {&args}, lds, stream), its only argument being its args struct.
const void *pass_kernel(bool dbl, int ept);
const void *rader_kernel(bool dbl, int ept);
const void *bluestein_kernel(bool dbl, int ept);
const void *r2c_even_kernel(bool dbl, int ept);
const void *c2r_even_kernel(bool dbl, int ept);
const void *real_post_kernel(bool dbl);
const void *real_pre_kernel(bool dbl);
LaunchShape real_post_shape(const RealPassArgs &a);
LaunchShape real_pre_shape(const RealPassArgs &a);

// Specialised single-pass kernels: id of the compiled radix sequence (-1 if none).
int leaf_kernel_id(const std::vector<int> &radices);
int leaf_kernel_ept(int id, bool dbl);
const void *leaf_kernel(bool dbl, int id, int mode);  // mode: 0 complex, 1 R2C, 2 C2R

// Specialised super-stage kernels: id of the compiled radix sequence (-1 if none),
// its units per workgroup and elements per thread.
int stage_kernel_id(const std::vector<int> &radices);
// The instance with `units` per workgroup (0: the default one), -1 if none.
int stage_kernel_id(const std::vector<int> &radices, bool dbl, int units);
void stage_kernel_shape(int id, bool dbl, int &units_per_group, int &ept);
const void *stage_kernel(bool dbl, int id, int mode);  // mode: 0 chain pass, 1 Bluestein convolution
// False for the instances that beat only the run-time-radix pass (5- and
// 10-smooth radices), which a Bluestein length search weighs like it.
bool stage_kernel_fast(int id);

// Register-resident two-stage super-stages (colpass.hip): the instance for
// super-radix r with `units` per workgroup (0: the default one), -1 if none;
// its units, threads, dynamic LDS bytes and element bytes; the kernel.
int col_kernel_id(int64_t r, bool dbl, int units);
int col_kernel_id(int64_t r, bool dbl, int units, bool first);
// Row-segment bytes (units x element bytes) of the default instance, 0 if none.
int col_run_bytes(int64_t r, bool dbl, bool first);
// An instance of radix r whose units per workgroup divide tcnt (a default
// instance first), or -1.
int col_kernel_id_dividing(int64_t r, bool dbl, int64_t tcnt);
void col_kernel_shape(int id, int &units, int &threads, size_t &lds_bytes, int &elem_bytes);
const void *col_kernel(int id);
// The same instance with the DC / Nyquist pair modes of a 2D real column pass
// (StageArgs::dc_nyq_pair), null if it has none; the first-pass default
// instance of r when it has one, else -1.
const void *col_pair_kernel(int id);
int col_pair_kernel_id(int64_t r, bool dbl);
// The instance of 256 threads the persistent queue kernel also runs (-1 if
// none for r), and its queue variant (QueueJob::variant; -1 for others).
int col_queue_kernel_id(int64_t r, bool dbl);
int col_queue_variant(int id);
// Template arguments after the element type, as the kernel symbol lists them.
std::string col_kernel_targs(int id, bool pair = false);

// Persistent mixed-length kernel (kQueueThreads threads per workgroup, one
// QueueArgs argument).
const void *queue_kernel();
// Persistent kernel of the register-resident pass jobs (kQueueCol) of one
// queue variant (QueueJob::variant), null if none.
const void *queue_col_kernel(int variant);

// Bluestein kernels with the convolution in LDS (FP32): id for the
// convolution's radix sequence (-1 if none), elements per thread, handle.
// LDS-resident Bluestein instances: the radices of the part in LDS, and
// whether the kernel runs the convolution as two halves (M = 2 x that part).
struct BluesteinLdsShape {
  std::vector<int> radices;
  bool halves = false;
  int64_t m = 0;  // convolution length
};
int bluestein_lds_kernel_id(const std::vector<int> &radices, bool dbl, bool halves);
int bluestein_lds_ept(int id);
bool bluestein_lds_has_stockham(int id);
// variant 0: Stockham stages; 1: Stockham, base twiddles prefetched; 2 / 3:
// in-place stages with table / binary-powered twiddles (bhat in
// digit-reversed order); instances without Stockham kernels ignore it
const void *bluestein_lds_kernel(int id, int variant);
int bluestein_lds_threads(int id, int variant);
std::vector<BluesteinLdsShape> bluestein_lds_candidates(bool dbl);
// Split form of a length-m convolution over `split` workgroups per sequence
// (-1 when no instance), the radices of its LDS block, and the combine kernel.
int bluestein_split_kernel_id(int64_t m, bool dbl, int split);
std::vector<int> bluestein_lds_kernel_radices(int id);
const void *bluestein_combine_kernel(bool dbl, int split);
// Row kinds of the single-pass kernels (the kernels' LeafMode values).
constexpr int kRowComplex = 0, kRowR2C = 1, kRowC2R = 2;

// Single-pass transforms spread over P workgroups per sequence (leaf_split.hip):
// instance for length n (-1 when none), its split P, block length L = n / P,
// elements per thread and block radices, and the kernel.
int leaf_split_kernel_id(int64_t n);
void leaf_split_shape(int id, bool dbl, int &p, int &l, int &ept, std::vector<int> &radices);
const void *leaf_split_kernel(int id, bool dbl, int mode);  // mode: LeafMode (real modes: P = 2 instances)

// Split-precision matrix-core transform of length 1024 (kComputeBf16x3 / kComputeFp16x3).
// Split-precision 1024-point kernels: `variant` < mfma_variants() picks the
// prefetch depth / occupancy instance (0: the default).
int mfma_variants();
const void *mfma_kernel(int mode, int variant);
std::string mfma_kernel_targs(int variant);

// Raise a kernel's dynamic-LDS limit when a launch needs more than 64 KiB.
hipError_t allow_lds(const void *f, size_t bytes);

// ---- table generation on the GPU, FP64 arithmetic ------------------------
// out[j] = w_n^{(j*mult) mod n} = exp(-2 pi i ((j*mult) mod n) / n), j < count.
hipError_t gen_roots(void *out, bool dbl, int64_t count, int64_t n, int64_t mult, hipStream_t st);
// out[k] = exp(-i pi (k^2 mod 2n) / n), k < n.
hipError_t gen_chirp(void *out, bool dbl, int64_t n, hipStream_t st);
// chirp[k] w_m^(h k), h < p, k < n (p blocks of n)
hipError_t gen_chirp_split(void *out, bool dbl, int64_t n, int64_t m, int p, hipStream_t st);
// FP64 Bluestein kernel b[k] = exp(+i pi (k^2 mod 2n)/n) for k < n, mirrored to M-k, zero elsewhere.
hipError_t gen_bluestein_b(double2 *out, int64_t n, int64_t m, hipStream_t st);
// FP64 Rader kernel b[q] = w_p^{idx[q]}, q < p-1.
hipError_t gen_rader_b(double2 *out, const int32_t *idx, int64_t p, hipStream_t st);
// out[j] = in[j] * scale, cast to the plan precision.
hipError_t scale_cast(const double2 *in, void *out, bool dbl, int64_t count, double scale, hipStream_t st);

}  // namespace flagfft
