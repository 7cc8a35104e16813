// planner.hpp — decomposition choice for the MI355X FFT path (host only, no HIP).
//
// Mirrors the role of the reference's src/plan (PlanBuilder::build,
// build_auto_candidates, select_leaf_factors: src/plan/builder_context.cpp:42-67,
// src/plan/auto_candidates.cpp:56-106, src/plan/factorization.cpp:70-193) for a
// different machine: the decision is between a single LDS-resident pass, a
// multi-pass Stockham super-stage chain, fused Rader and Bluestein, with
// real transforms split into an n/2-point complex core where possible.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace flagfft {

enum class Prec { F32, F64 };

inline int elem_bytes(Prec p) { return p == Prec::F32 ? 8 : 16; }

// Tunables; defaults are the untuned heuristics, the tuned-plan DB may override.
struct PlanOptions {
  int64_t leaf_max_f32 = 8192;    // longest single-pass LDS transform
  int64_t leaf_max_f64 = 4096;
  int64_t pass_max_f32 = 4096;    // longest super-radix of a multi-pass chain
  int64_t pass_max_f64 = 2048;
  int lds_pref_bytes = 64 * 1024;   // preferred LDS per workgroup
  int lds_cap_bytes = 160 * 1024;   // hard cap (gfx950 LDS per CU)
  int target_threads = 256;         // workgroup size aimed for by single-pass rows
  int chunk_bytes = 128;            // contiguous bytes per row segment for strided passes
  int ept = 0;                      // 0 = automatic (16, or 32 when needed)
  int xcd_remap = 1;
  std::vector<int64_t> forced_passes;  // explicit super-radix split (tuned plans)
  std::vector<int> forced_radices;     // explicit single-pass radix order (tuned plans)
  int forced_units = 0;                // explicit units per workgroup (tuned plans)
  int kernel_variant = 0;              // 1-based instance of a kernel family with several launch shapes (0: default)
  int stage_form = 0;                  // multi-pass chains: 0 automatic, 1 register-resident passes, 2 LDS super-stages
  int max_threads = 512;               // workgroup size limit of the launches
  bool generic_kernels = false;        // only the run-time-radix kernels (persistent-queue jobs)
  bool queue_form = false;             // with generic_kernels: register-resident passes the queue kernel runs
  bool in_place = false;               // input and output are one buffer: no sequence split over workgroups
};

enum class CoreKind { Leaf, MultiPass, Rader, Bluestein, MatrixCore };

// A length-n complex transform of a batch of sequences.
struct Core {
  CoreKind kind = CoreKind::Leaf;
  int64_t n = 0;
  std::vector<int> radices;                  // Leaf; Rader/Bluestein fused inner transform
  std::vector<int64_t> passes;               // MultiPass super-radices
  std::vector<std::vector<int>> pass_radices;
  int64_t m = 0;                             // Rader: p-1; Bluestein: M
  int64_t root = 0;                          // Rader primitive root
  std::shared_ptr<Core> conv;                // Bluestein: multi-pass core of M (null when fused)
  int compute = 0;                           // MatrixCore: 1 = BF16 hi/lo, 2 = FP16 hi/lo operands
  bool lds_conv = false;                     // Bluestein: convolution in LDS by a specialised kernel
  bool lds_halves = false;                   // lds_conv as two halves of M / 2 (radices cover M / 2)
  std::string describe(int indent = 0) const;
};

enum class RealForm { None, EvenFused, EvenMulti, Full };

struct Plan1D {
  int64_t n = 0;
  bool real_fwd = false, real_inv = false;
  RealForm real_form = RealForm::None;
  Core core;  // length n, or n/2 for EvenFused / EvenMulti
  std::string describe(int indent = 0) const;
};

struct PlanDesc {
  int rank = 1;
  int64_t n0 = 0, n1 = 0, n2 = 0;  // rank 1: n0 = n; row-major n0 x n1 (x n2), last axis contiguous
  int type = 0;                // flagfftType value
  int64_t batch = 1;
  Prec prec = Prec::F32;
  bool real_fwd = false, real_inv = false;
  int compute = 0;             // 0 = vector FP32/FP64; 1/2 = split BF16/FP16 matrix cores (C2C n=1024)
};

struct PlanTree {
  PlanDesc desc;
  Plan1D rows;   // rank 1: the transform; rank 2 / 3: along the last axis (contiguous)
  Plan1D mid;    // rank 3: C2C along n1 (strided)
  Plan1D cols;   // rank 2 / 3: C2C along n0 (strided)
  // rank 2 real, one transform: the DC and Nyquist half-spectrum columns
  // share one unit of a single register-resident column pass (n1 / 2 units
  // of work instead of n1 / 2 + 1, which would leave one tile alone)
  bool dc_nyq_pair = false;
  std::string describe() const;
};

// Helpers exposed for tests and the executor.
bool is_prime(int64_t n);
int64_t primitive_root(int64_t p);                  // smallest, -1 if p not prime
bool radix_factor(int64_t n, std::vector<int> &out);  // small radices 2..16 (product n)
bool smooth13(int64_t n);
int64_t leaf_max(const PlanOptions &o, Prec p);
int64_t pass_max(const PlanOptions &o, Prec p);
inline int lds_pad(int e) { return e + (e >> 5); }   // must match the kernels' lpad()

// Catalogue of the compile-time specialised super-stage kernels (leaf.hip):
// id of the kernel for this radix sequence, -1 when none is compiled.
int stage_kernel_id(const std::vector<int> &radices);
// Same for the specialised single-pass kernels.
int leaf_kernel_id(const std::vector<int> &radices);
// Register-resident two-stage pass instance for super-radix r (colpass.hip), -1 if none.
int col_kernel_id(int64_t r, bool dbl, int units);
int col_queue_kernel_id(int64_t r, bool dbl);
int col_run_bytes(int64_t r, bool dbl, bool first);
int col_pair_kernel_id(int64_t r, bool dbl);
int col_kernel_id_dividing(int64_t r, bool dbl, int64_t tcnt);
// True when every super-radix of the split has a specialised kernel (fast
// only: one that runs near the power-of-two passes' rate).
bool passes_specialised(const std::vector<int64_t> &passes, bool fast_only = false);
bool stage_kernel_fast(int id);

Core plan_core(int64_t n, Prec p, const PlanOptions &o);
PlanTree plan_tree(const PlanDesc &d, const PlanOptions &o);

}  // namespace flagfft
