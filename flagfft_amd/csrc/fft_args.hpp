// fft_args.hpp — kernel argument blocks shared by the host executor and the
// gfx950 kernels.  Plain structs, passed by value in the kernarg segment.
#pragma once

#include <cstdint>

namespace flagfft {

constexpr int kMaxStages = 16;       // small-radix stages inside one LDS transform
constexpr int kLdsBytesMax = 160 * 1024;  // gfx950 LDS per CU

// Unsigned 32-bit division by a run-time invariant (Granlund-Montgomery):
// q = (umulhi(n, mul) + n) >> shift, valid for n < 2^31.
struct FastDiv {
  uint32_t d;
  uint32_t mul;
  uint32_t shift;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f{d, 0, 0};
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  f.shift = l;
  // mul = floor(2^32 * (2^l - d) / d) + 1
  f.mul = static_cast<uint32_t>(((static_cast<uint64_t>(1) << 32) * ((static_cast<uint64_t>(1) << l) - d)) / d + 1);
  return f;
}

// Source of logical element e of an input sequence (first pass / fused kernels).
enum SrcMode : int32_t {
  kSrcComplex = 0,   // complex at in[b*bs + c*cs + e*es]
  kSrcReal = 1,      // real at in[...], imaginary part 0
  kSrcHermitian = 2, // n/2+1 complex bins expanded to n (Im of DC/Nyquist dropped)
  kSrcChirpFlag = 4, // Bluestein: e < n -> base(e) * chirp[e], n <= e < M -> 0
};

// Sink of logical output element e (last pass / fused kernels).
enum DstMode : int32_t {
  kDstComplex = 0,   // complex at out[...]
  kDstHalf = 1,      // complex, only e <= n/2
  kDstReal = 2,      // real part only
  kDstPointwiseFlag = 8,  // Bluestein middle: v = conj(v * bhat[e]), complex store
  kDstChirpFlag = 16,     // Bluestein end: e < n only, v = conj(v) * chirp[e] * scale, then base sink
};

struct Geometry {
  // element e of sequence (b, c) lives at base + b*bs + c*cs + e*es
  // (units of complex elements, or of reals for real sources / sinks)
  int64_t es, cs, bs;
};

// One Stockham super-stage (or a whole transform when tcnt == 1) over a batch
// of sequences.  A "unit" is one length-R DFT: unit u -> (b, c, i), i < tcnt.
struct PassArgs {
  const void *in;
  void *out;
  Geometry gi, go;
  int64_t units;       // nbatch * ncol * tcnt
  int64_t nseq;        // logical sequence length (R * tcnt) — Bluestein: conv length M
  int64_t nlog;        // user length n (real / Hermitian / chirp modes)
  int32_t R;           // DFT length computed in LDS by one unit
  int32_t tcnt;        // units per sequence = nseq / R
  int32_t p;           // prior product (1 on the first pass)
  int32_t ncol;
  int32_t col_fast;    // unit order: 1 -> c fastest, 0 -> i fastest
  int32_t C;           // units per workgroup
  int32_t tpu;         // threads per unit
  int32_t lds_stride;  // padded complex slots per unit in LDS
  int32_t load_slot_fast, store_slot_fast;  // coalescing order of the staged global accesses
  int32_t nst;
  int32_t rad[kMaxStages];
  FastDiv fd_q[kMaxStages];  // divide by the prior product q_s of each LDS stage
  FastDiv fd_C, fd_R, fd_ncol, fd_tcnt, fd_p;
  int32_t twq_step[kMaxStages];  // R / (q_s * rad_s)
  const void *tw_small;  // w_R^j, j < R
  const void *tw_lo;     // w_nseq^l, l < 2^tw_bits
  const void *tw_hi;     // w_nseq^(h 2^tw_bits)
  int32_t tw_bits;
  int32_t twp_mult;      // tcnt / p: super-stage twiddle index = e * k * twp_mult
  int32_t src_mode, dst_mode;
  int32_t conj_in, conj_out;
  double scale;          // applied on the final sink (1 unless stated)
  const void *chirp;     // Bluestein chirp (nlog entries)
  const void *bhat;      // Bluestein / Rader FFT of the convolution kernel
  int32_t xcd_remap;     // map consecutive tiles to one XCD
  int32_t nblocks;
  int32_t direct;        // single-pass contiguous sequences: first/last stage talk to HBM directly
};

// Fused single-kernel transforms that keep a whole sequence in LDS.
struct FusedArgs {
  const void *in;
  void *out;
  Geometry gi, go;
  int64_t nseqs;       // number of sequences (ncol * nbatch); sequence q -> (c = q % ncol, b = q / ncol)
  int32_t ncol;
  int32_t load_slot_fast, store_slot_fast;  // coalescing order of the global accesses
  int32_t n;           // user length (prime p for Rader)
  int32_t m;           // inner transform length (p-1 for Rader, M for Bluestein, n/2 for real-even)
  int32_t C;           // sequences per workgroup
  int32_t tpu;
  int32_t lds_stride;  // complex slots per sequence in LDS
  int32_t nst;
  int32_t rad[kMaxStages];
  FastDiv fd_q[kMaxStages];
  int32_t twq_step[kMaxStages];
  FastDiv fd_C, fd_ncol, fd_m;
  int32_t lext, sext;    // elements per sequence read in the load phase / written in the store phase
  FastDiv fd_lext, fd_sext;
  const void *tw_small;  // w_m^j, j < m
  const void *aux;       // Rader: int32 index table g^q mod p; real-even: w_n^k, k <= n/2
  const void *bhat;      // FFT_m of the convolution kernel, pre-scaled by 1/m
  const void *chirp;     // Bluestein chirp (n entries)
  int32_t src_mode, dst_mode;
  int32_t conj_in, conj_out;
  double scale;
};

// Compile-time specialised single-pass transform of contiguous complex rows.
struct LeafArgs {
  const void *in;
  void *out;
  int64_t nseq;          // transforms
  int64_t idist, odist;  // complex elements between consecutive transforms
  const void *tw;        // w_N^j, j < N
  const void *aux;       // real rows: w_n^k, k <= N (n = 2N)
  int32_t conj_in, conj_out;
  double scale;
  int32_t C;             // transforms per workgroup
  // real rows of a 2D transform with one column set: bin N's real part rides
  // in the imaginary part of bin 0 (R2C writes it there and leaves bin N
  // unwritten, C2R reads it there; both bins are real on that side)
  int32_t dc_nyq_pair;
};

// Single-pass transform of length N = P L spread over P workgroups per
// sequence (contiguous complex rows): workgroup (q, h) computes
// X[P k + h], k < L.
struct LeafSplitArgs {
  const void *in;
  void *out;
  int64_t nseq;
  int64_t idist, odist;  // complex elements between consecutive transforms
  const void *tw_n;      // w_N^j, j < N
  const void *tw_l;      // w_L^j, j < L
  const void *aux;       // real rows (2N reals): w_2N^k, k <= N
  int32_t conj_in, conj_out;
  double scale;
};

// Compile-time specialised Stockham super-stage over complex sequences
// (unit u -> (b = u / tcnt, i = u % tcnt)).  Sequences are contiguous
// (ncol = 1), or the ncol columns of row-major rows of ncol elements: the
// column index is then the fastest digit of i, tcnt and p count
// column-interleaved units (ncol times the per-column figures), and element
// e of column c sits at c + e ncol.
struct StageArgs {
  const void *in;
  void *out;
  int64_t bs_in, bs_out;  // complex elements between sequences
  int32_t tcnt, p, twp_mult, tw_bits;
  FastDiv fd_tcnt, fd_p;
  int32_t ncol;
  FastDiv fd_ncol;
  const void *tw_small;   // w_R^j
  const void *tw_lo;      // w_nseq^l, l < 2^tw_bits
  const void *tw_hi;      // w_nseq^(h 2^tw_bits)
  int32_t conj_in, conj_out;
  double scale;
  int32_t nblocks, xcd_remap;
  int64_t units;          // nbatch * tcnt
  int32_t block_store;    // the C units' outputs are one contiguous block (first pass, contiguous rows)
  int32_t src_mode, dst_mode;  // complex, or the Bluestein chirp source / pointwise and chirp sinks
  int64_t nlog;           // user length for the chirp modes
  const void *chirp, *bhat;
  // single column pass of a 2D real transform whose DC column A and Nyquist
  // column B (column `units`) share unit 0 as A + i B:
  //   1: forward, A and B real: split unit 0's output Z into
  //      FFT A = (Z_e + conj Z_-e) / 2 (column 0) and FFT B = -i (Z_e - conj Z_-e) / 2
  //   2: inverse: unit 0 reads Herm(A) + i Herm(B), Herm(V)_e = (V_e + conj V_-e) / 2,
  //      whose inverse transform is Re IFFT A + i Re IFFT B
  int32_t dc_nyq_pair;
};

// Bluestein with the whole convolution in LDS (one sequence per workgroup).
struct BluesteinLdsArgs {
  const void *in;
  void *out;
  Geometry gi, go;
  int32_t ncol;
  FastDiv fd_ncol;
  int64_t n;             // transform length
  const void *tw;        // w_M^j, j < M
  const void *chirp;     // n entries
  const void *bhat;      // FFT_M(b) / M
  void *park;            // two-halves kernels: n elements per sequence
  const void *chirp_split;  // split kernels: chirp[k] w_M^(h k), P blocks of n
  int32_t src_mode, dst_mode;
  int32_t conj_in, conj_out;
  double scale;
};

// Last stage of a split LDS Bluestein convolution: W[e] = sum_h Z_h[e mod ml]
// w_M^(h e), y[e] = conj(W[e]) chirp[e] for e < n (M = P ml).
struct BluesteinCombineArgs {
  const void *in;  // the P blocks Z_h of every sequence (P ml elements each)
  void *out;
  Geometry go;
  int32_t ncol;
  FastDiv fd_ncol;
  int32_t nb;  // workgroups per sequence
  FastDiv fd_nb;
  int32_t n, ml;
  const void *tw;     // w_M^j, j < M
  const void *chirp;  // n entries
  int32_t dst_mode, conj_out;
  double scale;
};

// Split-precision matrix-core transform of length 1024 (contiguous complex rows).
enum ComputeMode : int32_t { kComputeVector = 0, kComputeBf16x3 = 1, kComputeFp16x3 = 2 };
struct MfmaArgs {
  const void *in;
  void *out;
  int64_t nseq;          // transforms
  int64_t idist, odist;  // complex elements between consecutive transforms
  const void *tw;        // FP32 w_1024^j, j < 1024
  int32_t conj_in, conj_out;
  double scale;
};


// Split-spectrum pre/post pass for even-length real transforms that do not fit one LDS transform.
struct RealPassArgs {
  const void *in;
  void *out;
  Geometry gi, go;
  int64_t nseqs;
  int32_t ncol;
  int32_t n;            // real length (even)
  // w_n^k = tw_hi[k >> tw_bits] tw_lo[k & (2^tw_bits - 1)] (two small tables
  // that stay in cache instead of one of n/2 + 1 entries read beside the data)
  const void *tw_lo, *tw_hi;
  int32_t tw_bits;
  int32_t conj_flag;
};

// One job of the persistent mixed-length kernel: `nblocks` workgroup tiles of
// one launch of a plan (a single-pass transform, or one pass of a multi-pass
// chain whose earlier passes ran in an earlier queue launch), with the
// arguments of the equivalent standalone launch (run-time-radix kinds: FP32
// with 16 elements per thread, FP64 with 8; register-resident passes: the
// queue instance `variant` of 256 threads).
enum QueueKind : int32_t {
  kQueuePass = 0,
  kQueueRader = 1,
  kQueueBluestein = 2,
  kQueueR2CEven = 3,
  kQueueC2REven = 4,
  kQueueCol = 5,
  kQueueRealPost = 6,
  kQueueRealPre = 7
};
struct QueueJob {
  int32_t kind;     // QueueKind
  int32_t dbl;
  int32_t first;    // index of the job's first tile in the launch
  int32_t nblocks;  // tiles
  int32_t variant;  // kQueueCol: queue instance of the register-resident pass
  int32_t pad_;
  union {
    PassArgs pa;      // kQueuePass
    FusedArgs fa;     // kQueueRader, kQueueBluestein, kQueueR2CEven, kQueueC2REven
    StageArgs sa;     // kQueueCol
    RealPassArgs ra;  // kQueueRealPost, kQueueRealPre
  };
};
struct QueueArgs {
  const QueueJob *jobs;  // sorted by `first`
  int32_t njobs;
  int32_t total;         // tiles over all jobs
  uint32_t *counter;     // next tile to hand out; 0 at launch
};
constexpr int kQueueThreads = 256;

}  // namespace flagfft
