// exec.hpp — executable plans: ordered kernel launches over user buffers,
// plan-owned workspaces and FP64-generated device tables.
//
// Plays the role of the reference's CompiledRaw* node tree
// (src/exec/raw_nodes.cpp:41-852) with ahead-of-time compiled HIP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "fft_args.hpp"
#include "kernels.hpp"
#include "planner.hpp"

namespace flagfft {

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string &what) : std::runtime_error(what), code(c) {}
};

void hip_check(hipError_t e, const char *what);

// Device allocation owned by a plan.
struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
  explicit DevBuf(size_t b);
  ~DevBuf();
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
};

enum BufId : int { kBufIn = 0, kBufOut = 1, kBufWs0 = 2 };

struct Step {
  enum Kind {
    Pass,
    Leaf,
    Stage,
    Rader,
    Bluestein,
    BluesteinLds,
    BluesteinCombine,
    LeafSplit,
    R2CEven,
    C2REven,
    RealPost,
    RealPre,
    Mfma,
    Col
  } kind = Pass;
  bool dbl = false;
  int ept = 16;
  int leaf_id = -1;  // Leaf / Stage: entry of the specialised kernel table
  int leaf_mode = 0;  // Leaf: 0 complex rows, 1 R2C, 2 C2R
  int src = kBufIn, dst = kBufOut;
  int aux = -1;  // BluesteinLds halves / split: workspace id of the park buffer
  int split = 0;  // BluesteinLds / BluesteinCombine: workgroups per sequence of a split convolution
  PassArgs pa{};
  LeafArgs la{};
  StageArgs sa{};
  FusedArgs fa{};
  RealPassArgs ra{};
  MfmaArgs ma{};
  BluesteinLdsArgs ba{};
  BluesteinCombineArgs ca{};
  LeafSplitArgs lsa{};
  int compute = 0;  // Mfma: kComputeBf16x3 / kComputeFp16x3
  int64_t groups_max = 0;  // Mfma: workgroups of a launch (resident at once)
  LaunchShape shape{};
  int64_t bytes_moved = 0;  // algorithmic HBM bytes of this launch (read + write)
  std::string label;
  const void *func = nullptr;  // kernel handle, set by Executable::resolve_kernels()

  static constexpr size_t kArgBytes = 1024;
  // The step's argument struct with its in/out (and aux) pointers set;
  // returns its size.
  size_t pack_args(void *src, void *dst, unsigned char *buf, void *aux = nullptr) const;
  // The same launch for `batch` sequences of a 1D plan built for `built` (every
  // per-batch count of a 1D plan is proportional to the batch).
  Step rebatched(int64_t batch, int64_t built) const;
};

class Executable {
 public:
  std::vector<Step> steps;
  std::vector<size_t> ws_bytes;                 // per workspace id (kBufWs0 + i)
  std::vector<std::shared_ptr<DevBuf>> ws;
  std::map<std::string, std::shared_ptr<DevBuf>> tables;

  Executable() = default;
  ~Executable();
  Executable(const Executable &) = delete;
  Executable &operator=(const Executable &) = delete;

  // Enqueue every step on `st`; in/out are device pointers.
  hipError_t run(const void *in, void *out, hipStream_t st) const;
  // Same launches as one hipGraph: recorded on first use, later calls patch
  // only the kernel nodes that read `in` or write `out`, then launch it.
  hipError_t run_graph(const void *in, void *out, hipStream_t st);
  // Look up every step's kernel handle and raise LDS limits (after building).
  void resolve_kernels();
  // 1D plans: run for `batch` sequences with caller-provided workspaces
  // (ws[i] of ws_bytes[i] / built_batch * batch bytes).
  hipError_t run_batch(const void *in, void *out, void *const *ws, int64_t batch, hipStream_t st) const;
  int64_t built_batch = 1;
  // A launch reads input that another workgroup of the same launch
  // overwrites when the input and output are one buffer (split leaves): an
  // in-place call needs the executable built with PlanOptions::in_place.
  bool inplace_hazard = false;
  std::string describe() const;
  void allocate_workspaces();

  // Per-launch HIP-event timing (flagfftxSetProfiling / flagfftxGetProfile).
  void set_profiling(bool on);
  hipError_t harvest_profile() const;  // waits for recorded events, accumulates
  bool profiling = false;
  mutable std::vector<double> prof_ms;
  mutable std::vector<int64_t> prof_calls;

 private:
  struct EvPair {
    hipEvent_t start = nullptr, stop = nullptr;
  };
  mutable std::vector<EvPair> ev_free_;
  mutable std::vector<std::pair<size_t, EvPair>> ev_pending_;
  hipError_t take_pair(EvPair &p) const;
  void *buffer(int id, const void *in, void *out) const;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t graph_exec_ = nullptr;
  std::vector<hipGraphNode_t> graph_nodes_;
  const void *graph_in_ = nullptr;
  void *graph_out_ = nullptr;
};

// Geometry of a batch of equal-length sequences in one buffer.
struct SeqSet {
  int64_t nbatch = 1, ncol = 1;
  Geometry gi{1, 0, 0}, go{1, 0, 0};
  bool col_fast = false;
};

struct Modes {
  int src_mode = kSrcComplex, dst_mode = kDstComplex;
  bool conj_in = false, conj_out = false;
  int64_t nlog = 0;
  double scale = 1.0;
};

class Builder {
 public:
  Builder(Prec p, const PlanOptions &o, Executable &ex, hipStream_t setup);
  ~Builder();

  void emit_core(const Core &c, const SeqSet &s, int src, int dst, const Modes &m);
  void emit_plan1d(const Plan1D &pl, const SeqSet &s, int src, int dst, bool inverse);
  int pick_ws(std::initializer_list<int> avoid, size_t bytes);
  void finish();  // wait for table generation, drop temporaries
  // DC / Nyquist pair mode of the real-row leaf and the column pass emitted
  // next (PlanTree::dc_nyq_pair; 1 forward, 2 inverse, 0 off); returns the
  // number of launches emitted in that mode since the last call.
  int set_dc_nyq_pair(int mode);

 private:
  Prec prec_;
  PlanOptions opt_;
  Executable &ex_;
  hipStream_t st_;
  std::vector<std::shared_ptr<DevBuf>> temps_;
  std::vector<std::unique_ptr<Executable>> temp_execs_;
  int pair_ = 0, pair_steps_ = 0;

  bool dbl() const { return prec_ == Prec::F64; }
  int eb() const { return elem_bytes(prec_); }
  const void *roots(int64_t count, int64_t n, int64_t mult);
  const void *chirp(int64_t n);
  const void *bhat_bluestein(int64_t n, int64_t M);
  const void *bhat_bluestein_digit_reversed(int64_t n, int64_t M, const std::vector<int> &radices);
  int cu_count();
  const void *rader_idx(int64_t p, int64_t g);
  const void *bhat_rader(int64_t p, int64_t g);
  const void *fp64_transform_then_cast(const std::string &key, const std::shared_ptr<DevBuf> &src64, int64_t len,
                                       double scale);

  struct TwoLevel {  // w_n^j = hi[j >> bits] * lo[j & (2^bits - 1)]
    int bits = 0;
    const void *lo = nullptr, *hi = nullptr;
  };
  TwoLevel two_level(int64_t n);
  PassArgs chain_pass(const std::vector<int> &radices, int R, int64_t p, int64_t n, const SeqSet &s,
                      const Geometry &gi, const Geometry &go, const TwoLevel &tl);
  // Push a specialised super-stage step for `a` (mode 0 pass, 1 Bluestein
  // convolution); false when no instance or the layout does not fit it.
  bool try_stage(const PassArgs &a, const std::vector<int> &radices, int mode, int src, int dst, const SeqSet &s,
                 int64_t n, int pass, int npasses);
  void emit_pass_step(const PassArgs &a, int src, int dst, const SeqSet &s, int64_t n, int pass, int npasses);
  void emit_leaf(const Core &c, const SeqSet &s, int src, int dst, const Modes &m);
  void emit_multipass(const Core &c, const SeqSet &s, int src, int dst, const Modes &first, const Modes &last,
                      std::initializer_list<int> avoid_extra);
  void emit_rader(const Core &c, const SeqSet &s, int src, int dst, const Modes &m);
  void emit_bluestein(const Core &c, const SeqSet &s, int src, int dst, const Modes &m);
  void emit_bluestein_split(const Core &c, const SeqSet &s, int src, int dst, const Modes &m, int split);
  bool try_leaf_split(const Core &c, const SeqSet &s, int src, int dst, const Modes &m);
  bool try_leaf_split_real(const Plan1D &pl, const SeqSet &s, int src, int dst);
  void emit_leaf_split(int id, int mode, int64_t n, int64_t nseq, int64_t idist, int64_t odist, int src, int dst,
                       bool conj_in, bool conj_out, double scale);
  void emit_matrix_core(const Core &c, const SeqSet &s, int src, int dst, const Modes &m);
  void fill_stages(int32_t *rad, FastDiv *fdq, int32_t *twstep, int32_t &nst, const std::vector<int> &radices,
                   int64_t R);
};

// General complex layout of rank 1 / 2 (advanced PlanMany): element y of
// row x of batch entry b at b * dist + x * row + y * stride (row = embed[1] *
// stride; rank 1 has one row).
struct Strides {
  int64_t istride = 1, ostride = 1, irow = 0, orow = 0;
};

// Whole-plan builders (forward / inverse executables of a PlanTree).
// allocate_ws = false leaves the workspaces to the caller (Executable::run_batch).
std::unique_ptr<Executable> build_executable(const PlanTree &t, bool inverse, const PlanOptions &o,
                                             int64_t idist, int64_t odist, bool allocate_ws = true,
                                             const Strides *strides = nullptr);

}  // namespace flagfft
