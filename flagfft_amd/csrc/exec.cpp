// exec.cpp — build and run executable plans.
#include "exec.hpp"

#include <algorithm>
#include <cstring>
#include <limits>
#include <sstream>

namespace flagfft {

namespace {
// Workgroups of the matrix-core kernel: each of its 4 waves loops over
// transforms; 1024 groups keep every CU at its register-limited occupancy.
}  // namespace

void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) throw HipError(e, std::string(what) + ": " + hipGetErrorString(e));
}

DevBuf::DevBuf(size_t b) : bytes(b) {
  if (b == 0) return;
  hipError_t e = hipMalloc(&ptr, b);
  if (e != hipSuccess) {
    ptr = nullptr;
    (void)hipGetLastError();
    if (e == hipErrorOutOfMemory) throw std::bad_alloc();
    hip_check(e, "hipMalloc");
  }
}

DevBuf::~DevBuf() {
  if (ptr) (void)hipFree(ptr);
}

void Executable::allocate_workspaces() {
  ws.clear();
  for (size_t b : ws_bytes) ws.push_back(std::make_shared<DevBuf>(std::max<size_t>(b, 256)));
}

Executable::~Executable() {
  if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  for (auto &p : ev_pending_) {
    (void)hipEventDestroy(p.second.start);
    (void)hipEventDestroy(p.second.stop);
  }
  for (auto &p : ev_free_) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
}

void Executable::set_profiling(bool on) {
  (void)harvest_profile();
  profiling = on;
  prof_ms.assign(steps.size(), 0.0);
  prof_calls.assign(steps.size(), 0);
}

hipError_t Executable::take_pair(EvPair &p) const {
  if (!ev_free_.empty()) {
    p = ev_free_.back();
    ev_free_.pop_back();
    return hipSuccess;
  }
  hipError_t e = hipEventCreate(&p.start);
  if (e != hipSuccess) return e;
  return hipEventCreate(&p.stop);
}

hipError_t Executable::harvest_profile() const {
  hipError_t first_err = hipSuccess;
  if (prof_ms.size() != steps.size()) {
    prof_ms.assign(steps.size(), 0.0);
    prof_calls.assign(steps.size(), 0);
  }
  for (auto &pe : ev_pending_) {
    hipError_t e = hipEventSynchronize(pe.second.stop);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, pe.second.start, pe.second.stop);
    if (e == hipSuccess) {
      prof_ms[pe.first] += ms;
      prof_calls[pe.first] += 1;
    } else if (first_err == hipSuccess) {
      first_err = e;
    }
    ev_free_.push_back(pe.second);
  }
  ev_pending_.clear();
  return first_err;
}

size_t Step::pack_args(void *src, void *dst, unsigned char *buf, void *aux) const {
  auto put = [&](auto a) {
    static_assert(sizeof(a) <= kArgBytes, "kernel argument struct too large");
    a.in = src;
    a.out = dst;
    std::memcpy(buf, &a, sizeof(a));
    return sizeof(a);
  };
  switch (kind) {
    case Pass: return put(pa);
    case Leaf: return put(la);
    case Stage:
    case Col: return put(sa);
    case Mfma: return put(ma);
    case BluesteinLds: {
      BluesteinLdsArgs b = ba;
      b.park = aux;
      return put(b);
    }
    case BluesteinCombine: return put(ca);
    case LeafSplit: return put(lsa);
    case Rader:
    case Bluestein:
    case R2CEven:
    case C2REven: return put(fa);
    case RealPost:
    case RealPre: return put(ra);
  }
  return 0;
}

Step Step::rebatched(int64_t batch, int64_t built) const {
  Step s = *this;
  if (batch == built) return s;
  auto scale = [&](int64_t v) { return v / built * batch; };
  auto blocks = [](int64_t units, int64_t per) { return std::max<int64_t>(1, (units + per - 1) / per); };
  switch (kind) {
    case Pass:
      s.pa.units = scale(pa.units);
      s.pa.nblocks = static_cast<int32_t>(blocks(s.pa.units, pa.C));
      s.shape.grid = dim3(static_cast<unsigned>(s.pa.nblocks));
      break;
    case Col: {
      int per = 1, thr = 0, ebytes = 0;
      size_t lds = 0;
      col_kernel_shape(leaf_id, per, thr, lds, ebytes);
      // a tile may end partially only when the pass has a single sequence
      if (sa.tcnt % per != 0) throw std::logic_error("re-batched pass whose tiles would straddle sequences");
      s.sa.units = scale(sa.units);
      s.sa.nblocks = static_cast<int32_t>(blocks(s.sa.units, per));
      s.shape.grid = dim3(static_cast<unsigned>(s.sa.nblocks));
      break;
    }
    case Stage: {
      int per = 1, ept = 0;  // units per workgroup of the specialised kernel
      stage_kernel_shape(leaf_id, dbl, per, ept);
      s.sa.units = scale(sa.units);
      s.sa.nblocks = static_cast<int32_t>(blocks(s.sa.units, per));
      s.shape.grid = dim3(static_cast<unsigned>(s.sa.nblocks));
      break;
    }
    case Leaf:
      s.la.nseq = scale(la.nseq);
      s.shape.grid = dim3(static_cast<unsigned>(blocks(s.la.nseq, la.C)));
      break;
    case Mfma:
      s.ma.nseq = scale(ma.nseq);
      s.shape.grid = dim3(static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks(s.ma.nseq, 4), groups_max))));
      break;
    case BluesteinLds:
    case BluesteinCombine:  // workgroups per sequence are fixed
      s.shape.grid = dim3(static_cast<unsigned>(scale(static_cast<int64_t>(shape.grid.x))));
      break;
    case LeafSplit:
      s.lsa.nseq = scale(lsa.nseq);
      s.shape.grid = dim3(static_cast<unsigned>(scale(static_cast<int64_t>(shape.grid.x))));
      break;
    case Rader:
    case Bluestein:
    case R2CEven:
    case C2REven:
      s.fa.nseqs = scale(fa.nseqs);
      s.shape.grid = dim3(static_cast<unsigned>(blocks(s.fa.nseqs, fa.C)));
      break;
    case RealPost:
      s.ra.nseqs = scale(ra.nseqs);
      s.shape = real_post_shape(s.ra);
      break;
    case RealPre:
      s.ra.nseqs = scale(ra.nseqs);
      s.shape = real_pre_shape(s.ra);
      break;
  }
  return s;
}

hipError_t Executable::run_batch(const void *in, void *out, void *const *wsp, int64_t batch, hipStream_t st) const {
  alignas(16) unsigned char buf[Step::kArgBytes];
  auto res = [&](int id) -> void * {
    if (id == kBufIn) return const_cast<void *>(in);
    if (id == kBufOut) return out;
    return wsp[id - kBufWs0];
  };
  for (const Step &s0 : steps) {
    const Step s = s0.rebatched(batch, built_batch);
    s.pack_args(res(s.src), res(s.dst), buf, s.aux >= 0 ? res(s.aux) : nullptr);
    void *kp[] = {buf};
    hipError_t e = hipLaunchKernel(s.func, s.shape.grid, s.shape.block, kp, s.shape.lds_bytes, st);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

void Executable::resolve_kernels() {
  for (Step &s : steps) {
    switch (s.kind) {
      case Step::Pass: s.func = pass_kernel(s.dbl, s.ept); break;
      case Step::Leaf: s.func = leaf_kernel(s.dbl, s.leaf_id, s.leaf_mode); break;
      case Step::Stage: s.func = stage_kernel(s.dbl, s.leaf_id, s.leaf_mode); break;
      case Step::Col: s.func = s.leaf_mode != 0 ? col_pair_kernel(s.leaf_id) : col_kernel(s.leaf_id); break;
      case Step::Mfma: s.func = mfma_kernel(s.compute, s.leaf_id); break;
      case Step::BluesteinLds: s.func = bluestein_lds_kernel(s.leaf_id, s.leaf_mode); break;
      case Step::BluesteinCombine: s.func = bluestein_combine_kernel(s.dbl, s.split); break;
      case Step::LeafSplit: s.func = leaf_split_kernel(s.leaf_id, s.dbl, s.leaf_mode); break;
      case Step::Rader: s.func = rader_kernel(s.dbl, s.ept); break;
      case Step::Bluestein: s.func = bluestein_kernel(s.dbl, s.ept); break;
      case Step::R2CEven: s.func = r2c_even_kernel(s.dbl, s.ept); break;
      case Step::C2REven: s.func = c2r_even_kernel(s.dbl, s.ept); break;
      case Step::RealPost:
        s.func = real_post_kernel(s.dbl);
        s.shape = real_post_shape(s.ra);
        break;
      case Step::RealPre:
        s.func = real_pre_kernel(s.dbl);
        s.shape = real_pre_shape(s.ra);
        break;
    }
    if (s.func == nullptr) throw std::runtime_error("no compiled kernel for step: " + s.label);
    hip_check(allow_lds(s.func, s.shape.lds_bytes), "hipFuncSetAttribute");
  }
}

void *Executable::buffer(int id, const void *in, void *out) const {
  if (id == kBufIn) return const_cast<void *>(in);
  if (id == kBufOut) return out;
  return ws[static_cast<size_t>(id - kBufWs0)]->ptr;
}

hipError_t Executable::run(const void *in, void *out, hipStream_t st) const {
  alignas(16) unsigned char buf[Step::kArgBytes];
  for (size_t si = 0; si < steps.size(); ++si) {
    const Step &s = steps[si];
    s.pack_args(buffer(s.src, in, out), buffer(s.dst, in, out), buf, s.aux >= 0 ? buffer(s.aux, in, out) : nullptr);
    void *kp[] = {buf};
    hipError_t e = hipSuccess;
    EvPair ev;
    if (profiling) {
      e = take_pair(ev);
      if (e == hipSuccess) e = hipEventRecord(ev.start, st);
      if (e != hipSuccess) return e;
    }
    e = hipLaunchKernel(s.func, s.shape.grid, s.shape.block, kp, s.shape.lds_bytes, st);
    if (e != hipSuccess) return e;
    if (profiling) {
      e = hipEventRecord(ev.stop, st);
      if (e != hipSuccess) return e;
      ev_pending_.emplace_back(si, ev);
    }
  }
  return hipSuccess;
}

hipError_t Executable::run_graph(const void *in, void *out, hipStream_t st) {
  alignas(16) unsigned char buf[Step::kArgBytes];
  auto params = [&](const Step &s, void **kp) {
    s.pack_args(buffer(s.src, in, out), buffer(s.dst, in, out), buf, s.aux >= 0 ? buffer(s.aux, in, out) : nullptr);
    kp[0] = buf;
    hipKernelNodeParams p{};
    p.func = const_cast<void *>(s.func);
    p.gridDim = s.shape.grid;
    p.blockDim = s.shape.block;
    p.sharedMemBytes = static_cast<unsigned>(s.shape.lds_bytes);
    p.kernelParams = kp;
    p.extra = nullptr;
    return p;
  };
  hipError_t e = hipSuccess;
  if (graph_exec_ == nullptr) {
    e = hipGraphCreate(&graph_, 0);
    if (e != hipSuccess) return e;
    graph_nodes_.clear();
    for (const Step &s : steps) {
      void *kp[1];
      hipKernelNodeParams p = params(s, kp);
      hipGraphNode_t node;
      const hipGraphNode_t *dep = graph_nodes_.empty() ? nullptr : &graph_nodes_.back();
      e = hipGraphAddKernelNode(&node, graph_, dep, dep ? 1 : 0, &p);
      if (e != hipSuccess) return e;
      graph_nodes_.push_back(node);
    }
    e = hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      graph_exec_ = nullptr;
      return e;
    }
  } else if (in != graph_in_ || out != graph_out_) {
    for (size_t i = 0; i < steps.size(); ++i) {
      const Step &s = steps[i];
      if (s.src != kBufIn && s.src != kBufOut && s.dst != kBufIn && s.dst != kBufOut) continue;
      void *kp[1];
      hipKernelNodeParams p = params(s, kp);
      e = hipGraphExecKernelNodeSetParams(graph_exec_, graph_nodes_[i], &p);
      if (e != hipSuccess) return e;
    }
  }
  graph_in_ = in;
  graph_out_ = out;
  return hipGraphLaunch(graph_exec_, st);
}

namespace {

const char *buf_name(int id) {
  static const char *names[] = {"in", "out", "ws0", "ws1", "ws2", "ws3", "ws4", "ws5"};
  return id >= 0 && id < 8 ? names[id] : "ws?";
}

int ceil_log2(int64_t v) {
  int p = 0;
  while ((int64_t{1} << p) < v) ++p;
  return p;
}

// Address stride between consecutive units u, u+1 of a pass.
int64_t unit_stride(bool col_fast, int64_t tcnt, int64_t ncol, const Geometry &g, int64_t i_stride) {
  if (col_fast) return ncol > 1 ? g.cs : (tcnt > 1 ? i_stride : g.bs);
  return tcnt > 1 ? i_stride : (ncol > 1 ? g.cs : g.bs);
}

bool slot_fast(int64_t elem_stride, int64_t unit_str) { return elem_stride != 1 && unit_str == 1; }

int odd_stride(int ext) { return (ext % 2 == 0) ? ext + 1 : ext; }

// Algorithmic bytes read from / written to HBM for one sequence of a pass.
int64_t src_bytes(int mode, int64_t seq_len, int64_t nlog, int eb) {
  const int bm = mode & 3;
  const int64_t elems = (mode & kSrcChirpFlag) ? nlog : seq_len;
  if (bm == kSrcReal) return elems * eb / 2;
  if (bm == kSrcHermitian) return (nlog / 2 + 1) * eb;
  return elems * eb;
}

int64_t dst_bytes(int mode, int64_t seq_len, int64_t nlog, int eb) {
  if (mode & kDstPointwiseFlag) return seq_len * eb;
  const int bm = mode & 3;
  const int64_t elems = (mode & kDstChirpFlag) ? nlog : seq_len;
  if (bm == kDstHalf) return (nlog / 2 + 1) * eb;
  if (bm == kDstReal) return elems * eb / 2;
  return elems * eb;
}

}  // namespace

std::string Executable::describe() const {
  std::ostringstream o;
  o << "kernels=" << steps.size() << " workspaces=" << ws_bytes.size() << "\n";
  for (size_t i = 0; i < steps.size(); ++i) {
    const Step &s = steps[i];
    o << "  [" << i << "] " << s.label << " " << buf_name(s.src) << "->" << buf_name(s.dst) << " grid=" << s.shape.grid.x
      << " block=" << s.shape.block.x << " lds=" << s.shape.lds_bytes << "\n";
  }
  return o.str();
}

Builder::Builder(Prec p, const PlanOptions &o, Executable &ex, hipStream_t setup)
    : prec_(p), opt_(o), ex_(ex), st_(setup) {}

Builder::~Builder() {
  if (!temps_.empty() || !temp_execs_.empty()) (void)hipStreamSynchronize(st_);
}

int Builder::set_dc_nyq_pair(int mode) {
  const int n = pair_steps_;
  pair_ = mode;
  pair_steps_ = 0;
  return n;
}

void Builder::finish() {
  hip_check(hipStreamSynchronize(st_), "plan table generation");
  temps_.clear();
  temp_execs_.clear();
}

int Builder::cu_count() {
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
    n = 256;
  return n;
}

int Builder::pick_ws(std::initializer_list<int> avoid, size_t bytes) {
  for (int id = kBufWs0;; ++id) {
    if (std::find(avoid.begin(), avoid.end(), id) != avoid.end()) continue;
    size_t idx = static_cast<size_t>(id - kBufWs0);
    if (ex_.ws_bytes.size() <= idx) ex_.ws_bytes.resize(idx + 1, 0);
    ex_.ws_bytes[idx] = std::max(ex_.ws_bytes[idx], bytes);
    return id;
  }
}

const void *Builder::roots(int64_t count, int64_t n, int64_t mult) {
  std::ostringstream k;
  k << "roots:" << eb() << ":" << count << ":" << n << ":" << mult;
  auto it = ex_.tables.find(k.str());
  if (it != ex_.tables.end()) return it->second->ptr;
  auto buf = std::make_shared<DevBuf>(static_cast<size_t>(std::max<int64_t>(count, 1)) * eb());
  hip_check(gen_roots(buf->ptr, dbl(), std::max<int64_t>(count, 1), std::max<int64_t>(n, 1), mult, st_), "gen_roots");
  ex_.tables[k.str()] = buf;
  return buf->ptr;
}

const void *Builder::chirp(int64_t n) {
  std::string k = "chirp:" + std::to_string(eb()) + ":" + std::to_string(n);
  auto it = ex_.tables.find(k);
  if (it != ex_.tables.end()) return it->second->ptr;
  auto buf = std::make_shared<DevBuf>(static_cast<size_t>(n) * eb());
  hip_check(gen_chirp(buf->ptr, dbl(), n, st_), "gen_chirp");
  ex_.tables[k] = buf;
  return buf->ptr;
}

const void *Builder::fp64_transform_then_cast(const std::string &key, const std::shared_ptr<DevBuf> &src64,
                                              int64_t len, double scale) {
  Core c64 = plan_core(len, Prec::F64, opt_);
  if (c64.kind != CoreKind::Leaf && c64.kind != CoreKind::MultiPass)
    throw std::runtime_error("convolution length must be smooth");
  auto ex64 = std::make_unique<Executable>();
  {
    Builder b64(Prec::F64, opt_, *ex64, st_);
    SeqSet s;
    s.gi = Geometry{1, 0, len};
    s.go = Geometry{1, 0, len};
    Modes m;
    m.nlog = len;
    b64.emit_core(c64, s, kBufIn, kBufOut, m);
    b64.finish();
  }
  ex64->resolve_kernels();
  ex64->allocate_workspaces();
  auto out64 = std::make_shared<DevBuf>(static_cast<size_t>(len) * 16);
  hip_check(ex64->run(src64->ptr, out64->ptr, st_), "FP64 convolution kernel transform");
  auto dst = std::make_shared<DevBuf>(static_cast<size_t>(len) * eb());
  hip_check(scale_cast(static_cast<const double2 *>(out64->ptr), dst->ptr, dbl(), len, scale, st_), "scale_cast");
  temps_.push_back(src64);
  temps_.push_back(out64);
  temp_execs_.push_back(std::move(ex64));
  ex_.tables[key] = dst;
  return dst->ptr;
}

const void *Builder::bhat_bluestein(int64_t n, int64_t M) {
  std::string k = "bhatB:" + std::to_string(eb()) + ":" + std::to_string(n) + ":" + std::to_string(M);
  auto it = ex_.tables.find(k);
  if (it != ex_.tables.end()) return it->second->ptr;
  auto b64 = std::make_shared<DevBuf>(static_cast<size_t>(M) * 16);
  hip_check(gen_bluestein_b(static_cast<double2 *>(b64->ptr), n, M, st_), "gen_bluestein_b");
  return fp64_transform_then_cast(k, b64, M, 1.0 / static_cast<double>(M));
}

// FFT_M(b) / M with entry k moved to slot P(k) = sum_s digit_s(k) * m_s
// (digits of k in the mixed radix of the stages, least significant first;
// m_s = M / (r_0 ... r_s)): the order a decimation-in-frequency transform
// leaves its spectrum in.
const void *Builder::bhat_bluestein_digit_reversed(int64_t n, int64_t M, const std::vector<int> &radices) {
  std::string k = "bhatBrev:" + std::to_string(eb()) + ":" + std::to_string(n) + ":" + std::to_string(M);
  for (int r : radices) k += ":" + std::to_string(r);
  auto it = ex_.tables.find(k);
  if (it != ex_.tables.end()) return it->second->ptr;
  const void *nat = bhat_bluestein(n, M);
  const size_t eb_ = static_cast<size_t>(eb());
  std::vector<unsigned char> h(static_cast<size_t>(M) * eb_), p(static_cast<size_t>(M) * eb_);
  hip_check(hipMemcpyAsync(h.data(), nat, h.size(), hipMemcpyDeviceToHost, st_), "bhat download");
  hip_check(hipStreamSynchronize(st_), "bhat download");
  for (int64_t kk = 0; kk < M; ++kk) {
    int64_t pos = 0, w = 1, span = M;
    for (int r : radices) {
      span /= r;
      pos += ((kk / w) % r) * span;
      w *= r;
    }
    std::memcpy(&p[static_cast<size_t>(pos) * eb_], &h[static_cast<size_t>(kk) * eb_], eb_);
  }
  auto buf = std::make_shared<DevBuf>(p.size());
  hip_check(hipMemcpyAsync(buf->ptr, p.data(), p.size(), hipMemcpyHostToDevice, st_), "bhat upload");
  hip_check(hipStreamSynchronize(st_), "bhat upload");
  ex_.tables[k] = buf;
  return buf->ptr;
}

const void *Builder::rader_idx(int64_t p, int64_t g) {
  std::string k = "raderidx:" + std::to_string(p);
  auto it = ex_.tables.find(k);
  if (it != ex_.tables.end()) return it->second->ptr;
  std::vector<int32_t> idx(static_cast<size_t>(p - 1));
  int64_t v = 1;
  for (int64_t q = 0; q < p - 1; ++q) {
    idx[static_cast<size_t>(q)] = static_cast<int32_t>(v);
    v = (v * g) % p;
  }
  auto buf = std::make_shared<DevBuf>(idx.size() * sizeof(int32_t));
  hip_check(hipMemcpyAsync(buf->ptr, idx.data(), idx.size() * sizeof(int32_t), hipMemcpyHostToDevice, st_),
            "rader index upload");
  hip_check(hipStreamSynchronize(st_), "rader index upload");
  ex_.tables[k] = buf;
  return buf->ptr;
}

const void *Builder::bhat_rader(int64_t p, int64_t g) {
  std::string k = "bhatR:" + std::to_string(eb()) + ":" + std::to_string(p);
  auto it = ex_.tables.find(k);
  if (it != ex_.tables.end()) return it->second->ptr;
  const void *idx = rader_idx(p, g);
  auto b64 = std::make_shared<DevBuf>(static_cast<size_t>(p - 1) * 16);
  hip_check(gen_rader_b(static_cast<double2 *>(b64->ptr), static_cast<const int32_t *>(idx), p, st_), "gen_rader_b");
  return fp64_transform_then_cast(k, b64, p - 1, 1.0 / static_cast<double>(p - 1));
}

void Builder::fill_stages(int32_t *rad, FastDiv *fdq, int32_t *twstep, int32_t &nst, const std::vector<int> &radices,
                          int64_t R) {
  if (radices.size() > static_cast<size_t>(kMaxStages)) throw std::runtime_error("too many radix stages");
  nst = static_cast<int32_t>(radices.size());
  int64_t q = 1;
  for (size_t s = 0; s < radices.size(); ++s) {
    rad[s] = radices[s];
    fdq[s] = make_fastdiv(static_cast<uint32_t>(q));
    twstep[s] = static_cast<int32_t>(R / (q * radices[s]));
    q *= radices[s];
  }
  if (q != R) throw std::runtime_error("radix product mismatch");
}

namespace {

struct ShapeChoice {
  int C = 1, tpu = 1, ept = 16;
  size_t lds = 0;
};

ShapeChoice pick_shape(const PlanOptions &o, int eb, int64_t work_len, int LS, bool chunked, int64_t nunits) {
  // Workgroups wanted in flight: two per CU of the 256-CU part.
  constexpr int64_t kTargetGroups = 512;
  const bool dbl = eb == 16;
  const int epts_f[3] = {8, 16, 32};
  const int epts_d[2] = {8, 16};
  const int *epts = dbl ? epts_d : epts_f;
  const int nept = dbl ? 2 : 3;
  const int ept_default = dbl ? 8 : 16;
  struct Cand {
    ShapeChoice s;
    int64_t groups;
  };
  std::vector<Cand> cands;
  for (int i = 0; i < nept; ++i) {
    const int ept = epts[i];
    if (o.ept && ept != o.ept) continue;
    int tpu = static_cast<int>((work_len + ept - 1) / ept);
    if (tpu < 1) tpu = 1;
    if (tpu > o.max_threads) continue;
    // strided passes: at least chunk_bytes of adjacent units per row segment,
    // and (short units of few threads) at least target_threads threads
    int desired = chunked ? std::max({1, o.chunk_bytes / eb, o.target_threads / tpu})
                          : std::max(1, o.target_threads / tpu);
    if (!chunked && (nunits + desired - 1) / desired < kTargetGroups)
      desired = static_cast<int>(std::max<int64_t>(1, nunits / kTargetGroups));
    if (o.forced_units > 0) desired = o.forced_units;
    if (desired > nunits) desired = static_cast<int>(std::max<int64_t>(nunits, 1));
    const size_t cap = static_cast<size_t>(chunked ? o.lds_cap_bytes : o.lds_pref_bytes);
    int C = desired;
    while (C > 1 && (C * tpu > o.max_threads || static_cast<size_t>(C) * LS * eb > cap)) --C;
    if (static_cast<size_t>(C) * LS * eb > static_cast<size_t>(o.lds_cap_bytes)) continue;
    ShapeChoice sc;
    sc.C = C;
    sc.tpu = tpu;
    sc.ept = ept;
    sc.lds = static_cast<size_t>(C) * LS * eb;
    cands.push_back({sc, (nunits + C - 1) / C});
  }
  if (cands.empty()) throw std::runtime_error("no launch shape fits the LDS / block limits");
  // Few workgroups: the most threads per transform wins; otherwise the default
  // elements-per-thread with the largest tile.
  const Cand *best = &cands[0];
  for (const Cand &c : cands) {
    const bool starved = c.groups < kTargetGroups;
    if (starved) {
      if (c.groups * c.s.C * c.s.tpu > best->groups * best->s.C * best->s.tpu) best = &c;
    } else if (c.s.C > best->s.C || (c.s.C == best->s.C && c.s.ept == ept_default)) {
      best = &c;
    }
  }
  return best->s;
}

}  // namespace

void Builder::emit_leaf(const Core &c, const SeqSet &s, int src, int dst, const Modes &m) {
  // Compile-time specialised kernel for contiguous complex rows of a listed radix sequence.
  const int leaf_id = opt_.generic_kernels ? -1 : leaf_kernel_id(c.radices);
  if (leaf_id >= 0 && !s.col_fast && s.ncol == 1 && s.gi.es == 1 && s.go.es == 1 && m.src_mode == kSrcComplex &&
      m.dst_mode == kDstComplex && c.n > 1) {
    Step st;
    st.kind = Step::Leaf;
    st.dbl = dbl();
    st.src = src;
    st.dst = dst;
    st.leaf_id = leaf_id;
    LeafArgs &a = st.la;
    const int R = static_cast<int>(c.n);
    const int ept = leaf_kernel_ept(leaf_id, dbl());
    const int tpu = (R + ept - 1) / ept;
    const int LS = odd_stride(lds_pad(R - 1) + 1);
    a.nseq = s.nbatch;
    a.idist = s.gi.bs;
    a.odist = s.go.bs;
    a.tw = roots(R, R, 1);
    a.aux = nullptr;
    a.conj_in = m.conj_in;
    a.conj_out = m.conj_out;
    a.scale = m.scale;
    int C = std::max(1, opt_.target_threads / tpu);
    if ((s.nbatch + C - 1) / C < 512) C = static_cast<int>(std::max<int64_t>(1, s.nbatch / 512));
    if (opt_.forced_units > 0) C = opt_.forced_units;
    while (C > 1 && (C * tpu > kMaxBlockThreads || static_cast<size_t>(C) * LS * eb() > static_cast<size_t>(opt_.lds_pref_bytes)))
      --C;
    a.C = C;
    const int64_t nblocks = (s.nbatch + C - 1) / C;
    st.ept = ept;
    st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(C * tpu)),
                           static_cast<size_t>(C) * LS * eb()};
    st.bytes_moved = s.nbatch * R * eb() * 2;
    std::ostringstream lab;
    lab << "k_leaf<" << (dbl() ? "fp64" : "fp32") << "," << ept << ",0>[";
    for (size_t i = 0; i < c.radices.size(); ++i) lab << (i ? "x" : "") << c.radices[i];
    lab << "] R=" << R << " C=" << C << " tpu=" << tpu << " transforms=" << s.nbatch;
    st.label = lab.str();
    ex_.steps.push_back(st);
    return;
  }
  if (s.col_fast && m.src_mode == kSrcComplex && m.dst_mode == kDstComplex && c.n > 1) {
    // columns of contiguous rows: a specialised super-stage with the column
    // index as the fastest unit digit (no inter-pass twiddle)
    TwoLevel tl;
    tl.lo = tl.hi = roots(c.n, c.n, 1);
    PassArgs a = chain_pass(c.radices, static_cast<int>(c.n), 1, c.n, s, s.gi, s.go, tl);
    a.conj_in = m.conj_in;
    a.conj_out = m.conj_out;
    a.scale = m.scale;
    a.nlog = m.nlog ? m.nlog : c.n;
    if (try_stage(a, c.radices, 0, src, dst, s, c.n, 0, 1)) return;
  }
  Step st;
  st.kind = Step::Pass;
  st.dbl = dbl();
  st.src = src;
  st.dst = dst;
  PassArgs &a = st.pa;
  const int R = static_cast<int>(c.n);
  a.gi = s.gi;
  a.go = s.go;
  a.R = R;
  a.tcnt = 1;
  a.p = 1;
  a.nseq = c.n;
  a.nlog = m.nlog ? m.nlog : c.n;
  a.ncol = static_cast<int32_t>(s.ncol);
  a.col_fast = s.col_fast ? 1 : 0;
  a.units = s.nbatch * s.ncol;
  if (a.units >= (int64_t{1} << 31)) throw std::runtime_error("too many sequences for one launch");
  fill_stages(a.rad, a.fd_q, a.twq_step, a.nst, c.radices, R);
  a.tw_small = roots(R, R, 1);
  a.tw_lo = a.tw_small;
  a.tw_hi = a.tw_small;
  a.tw_bits = 0;
  a.twp_mult = 1;
  const int LS = odd_stride(lds_pad(R - 1) + 1);
  const bool lsf = slot_fast(s.gi.es, unit_stride(s.col_fast, 1, s.ncol, s.gi, s.gi.es));
  const bool ssf = slot_fast(s.go.es, unit_stride(s.col_fast, 1, s.ncol, s.go, s.go.es));
  ShapeChoice sh = pick_shape(opt_, eb(), R, LS, lsf || ssf, a.units);
  a.C = sh.C;
  a.tpu = sh.tpu;
  a.lds_stride = LS;
  a.load_slot_fast = lsf;
  a.store_slot_fast = ssf;
  a.fd_C = make_fastdiv(static_cast<uint32_t>(sh.C));
  a.fd_R = make_fastdiv(static_cast<uint32_t>(R));
  a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
  a.fd_tcnt = make_fastdiv(1);
  a.fd_p = make_fastdiv(1);
  a.src_mode = m.src_mode;
  a.dst_mode = m.dst_mode;
  a.conj_in = m.conj_in;
  a.conj_out = m.conj_out;
  a.scale = m.scale;
  a.chirp = nullptr;
  a.bhat = nullptr;
  const int64_t nblocks = (a.units + sh.C - 1) / sh.C;
  a.nblocks = static_cast<int32_t>(nblocks);
  a.xcd_remap = 0;
  a.direct = (!lsf && !ssf && !s.col_fast && s.gi.es == 1 && s.go.es == 1) ? 1 : 0;
  st.ept = sh.ept;
  st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(sh.C * sh.tpu)), sh.lds};
  st.bytes_moved = a.units * (src_bytes(a.src_mode, R, a.nlog, eb()) + dst_bytes(a.dst_mode, R, a.nlog, eb()));
  std::ostringstream lab;
  lab << "k_pass<" << (dbl() ? "fp64" : "fp32") << "," << sh.ept << "> single-pass" << (a.direct ? "/direct" : "/staged")
      << " R=" << R << " C=" << sh.C
      << " tpu=" << sh.tpu << " units=" << a.units;
  st.label = lab.str();
  ex_.steps.push_back(st);
}

PassArgs Builder::chain_pass(const std::vector<int> &radices, int R, int64_t p, int64_t n, const SeqSet &s,
                             const Geometry &gi, const Geometry &go, const TwoLevel &tl) {
  const int64_t tcnt = n / R;
  PassArgs a{};
  a.gi = gi;
  a.go = go;
  a.R = R;
  a.tcnt = static_cast<int32_t>(tcnt);
  a.p = static_cast<int32_t>(p);
  a.nseq = n;
  a.ncol = static_cast<int32_t>(s.ncol);
  a.col_fast = s.col_fast ? 1 : 0;
  a.units = s.nbatch * s.ncol * tcnt;
  if (a.units >= (int64_t{1} << 31)) throw std::runtime_error("too many units for one launch");
  fill_stages(a.rad, a.fd_q, a.twq_step, a.nst, radices, R);
  a.tw_small = roots(R, R, 1);
  a.tw_lo = tl.lo;
  a.tw_hi = tl.hi;
  a.tw_bits = tl.bits;
  a.twp_mult = static_cast<int32_t>(tcnt / p);
  a.src_mode = kSrcComplex;
  a.dst_mode = kDstComplex;
  a.conj_in = a.conj_out = 0;
  a.scale = 1.0;
  a.nlog = n;
  a.chirp = nullptr;
  a.bhat = nullptr;
  return a;
}

Builder::TwoLevel Builder::two_level(int64_t n) {
  TwoLevel t;
  t.bits = (ceil_log2(n) + 1) / 2;
  const int64_t lo_count = int64_t{1} << t.bits;
  const int64_t hi_count = (n + lo_count - 1) / lo_count;
  t.lo = roots(lo_count, n, 1);
  t.hi = roots(hi_count, n, lo_count);
  return t;
}

bool Builder::try_stage(const PassArgs &a, const std::vector<int> &radices, int mode, int src, int dst,
                        const SeqSet &s, int64_t n, int pass, int npasses) {
  const int sid = opt_.generic_kernels ? -1 : stage_kernel_id(radices, dbl(), opt_.forced_units);
  // register-resident two-stage pass (plain complex passes, and Bluestein
  // chain ends: chirp source, pointwise and chirp sinks)
  const bool col_allowed = mode == 0 && opt_.stage_form != 2 && (!opt_.generic_kernels || opt_.queue_form) &&
                           (a.src_mode & 3) == kSrcComplex && (a.dst_mode & 3) == kDstComplex;
  int cid = !col_allowed        ? -1
            : opt_.queue_form ? col_queue_kernel_id(a.R, dbl())
                              : col_kernel_id(a.R, dbl(), opt_.forced_units, a.p == 1);
  // DC / Nyquist pair: one transform's single column pass over the n1 / 2
  // columns of the half spectrum (row pitch n1 / 2 + 1) but its Nyquist
  // column, which goes with column 0; the other side is a workspace of n1 / 2
  // columns per row (row pitch at least that)
  const bool pair = pair_ != 0;
  const int64_t pair_units = pair ? (pair_ == 1 ? a.go.es : a.gi.es) - 1 : 0;
  if (pair) {
    if (!col_allowed || !s.col_fast || s.nbatch != 1 || npasses != 1 || a.tcnt != 1 || a.gi.cs != 1 ||
        a.go.cs != 1 || (pair_ == 1 ? a.gi.es : a.go.es) < pair_units || pair_units < 1)
      throw std::logic_error("DC/Nyquist pair: not a single column pass of one transform");
    cid = col_pair_kernel_id(a.R, dbl());
    if (cid < 0) throw std::logic_error("DC/Nyquist pair: no paired column instance");
  }
  if ((sid < 0 && cid < 0) || (a.src_mode & 3) != kSrcComplex || (a.dst_mode & 3) != kDstComplex) return false;
  // contiguous sequences, or the columns of contiguous rows of ncol elements
  const bool rows = !s.col_fast && s.ncol == 1 && a.gi.es == 1 && a.go.es == 1;
  const bool cols = pair || (s.col_fast && s.ncol > 1 && a.gi.es == s.ncol && a.gi.cs == 1 && a.go.es == s.ncol &&
                             a.go.cs == 1 && mode == 0 && a.src_mode == kSrcComplex && a.dst_mode == kDstComplex);
  if (!rows && !cols) return false;
  // pair: units read at input pitch tcnt and written at output pitch p (one
  // transform, so every unit index is below both)
  const int64_t ncol = pair ? a.go.es : cols ? s.ncol : 1;
  const int R = a.R;
  const int64_t tcnt = pair ? a.gi.es : a.tcnt * ncol, p = pair ? a.go.es : a.p * ncol;
  const int64_t units = pair ? pair_units : a.units;
  if (std::max(tcnt, p) >= (int64_t{1} << 31)) return false;
  int cunits = 0, cthreads = 0, cebytes = 0;
  size_t clds = 0;
  if (cid >= 0) col_kernel_shape(cid, cunits, cthreads, clds, cebytes);
  // the default instance's tile does not divide the units of a sequence:
  // another instance of the radix whose tile does
  if (cid >= 0 && !pair && tcnt % cunits != 0 && a.units != tcnt && opt_.forced_units == 0 && !opt_.queue_form) {
    cid = col_kernel_id_dividing(R, dbl(), tcnt);
    if (cid >= 0) col_kernel_shape(cid, cunits, cthreads, clds, cebytes);
  }
  // the register-resident pass addresses a sequence from a uniform base with
  // 32-bit offsets: its tiles may not straddle sequences, nor a sequence 4 GiB
  // (one sequence: the last tile may be partial, its lanes past the units idle)
  const bool col_ok = cid >= 0 && (pair || tcnt % cunits == 0 || a.units == tcnt) &&
                      std::max(tcnt, p) * R * eb() < (int64_t{1} << 32);
  if (pair && !col_ok) throw std::logic_error("DC/Nyquist pair: the column pass does not fit");
  if (!col_ok && sid < 0) return false;
  if (col_ok) {
    Step cst;
    cst.kind = Step::Col;
    cst.dbl = dbl();
    cst.src = src;
    cst.dst = dst;
    cst.leaf_id = cid;
    StageArgs &g = cst.sa;
    g.bs_in = a.gi.bs;
    g.bs_out = a.go.bs;
    g.tcnt = static_cast<int32_t>(tcnt);
    g.p = static_cast<int32_t>(p);
    g.twp_mult = a.twp_mult;
    g.tw_bits = a.tw_bits;
    g.fd_tcnt = make_fastdiv(static_cast<uint32_t>(tcnt));
    g.fd_p = make_fastdiv(static_cast<uint32_t>(p));
    g.ncol = static_cast<int32_t>(ncol);
    g.fd_ncol = make_fastdiv(static_cast<uint32_t>(ncol));
    g.tw_small = a.tw_small;
    g.tw_lo = a.tw_lo;
    g.tw_hi = a.tw_hi;
    g.conj_in = a.conj_in;
    g.conj_out = a.conj_out;
    g.scale = a.scale;
    g.units = units;
    g.block_store = (rows && p == 1 && a.go.bs == n && a.dst_mode == kDstComplex) ? 1 : 0;
    g.src_mode = a.src_mode;
    g.dst_mode = a.dst_mode;
    g.nlog = a.nlog;
    g.chirp = a.chirp;
    g.bhat = a.bhat;
    g.dc_nyq_pair = pair_;
    if (pair) {
      cst.leaf_mode = 1;  // the paired instance (Executable::resolve_kernels)
      ++pair_steps_;
    }
    const int64_t nblocks = (g.units + cunits - 1) / cunits;
    g.nblocks = static_cast<int32_t>(nblocks);
    g.xcd_remap = opt_.xcd_remap;
    cst.ept = 0;
    cst.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(cthreads)), clds};
    // (pair: the n1 / 2 + 1 half-spectrum columns' bytes on either side)
    cst.bytes_moved = s.nbatch * (pair ? units + 1 : ncol) *
                      (src_bytes(a.src_mode, n, a.nlog, eb()) + dst_bytes(a.dst_mode, n, a.nlog, eb()));
    std::ostringstream lab;
    lab << "k_col<" << (dbl() ? "fp64" : "fp32") << "," << col_kernel_targs(cid, pair) << "> pass " << pass << "/"
        << npasses << " R=" << R << " p=" << p << " tcnt=" << tcnt << (g.block_store ? " block-store" : "")
        << (cols ? " columns=" : "");
    if (cols) lab << ncol;
    if (pair) lab << (pair_ == 1 ? " dc-nyquist-unpack" : " dc-nyquist-pack");
    cst.label = lab.str();
    ex_.steps.push_back(cst);
    return true;
  }
  int cu = 0, ept = 0;
  stage_kernel_shape(sid, dbl(), cu, ept);
  const int tpu = (R + ept - 1) / ept;
  const int LS = odd_stride(lds_pad(R - 1) + 1);
  // the instance's launch bounds and the LDS of one CU
  if (cu * tpu > kMaxBlockThreads || static_cast<size_t>(cu) * LS * eb() > static_cast<size_t>(kLdsBytesMax))
    return false;
  Step sst;
  sst.kind = Step::Stage;
  sst.dbl = dbl();
  sst.src = src;
  sst.dst = dst;
  sst.leaf_id = sid;
  sst.leaf_mode = mode;
  StageArgs &g = sst.sa;
  g.bs_in = a.gi.bs;
  g.bs_out = a.go.bs;
  g.tcnt = static_cast<int32_t>(tcnt);
  g.p = static_cast<int32_t>(p);
  g.twp_mult = a.twp_mult;
  g.tw_bits = a.tw_bits;
  g.fd_tcnt = make_fastdiv(static_cast<uint32_t>(tcnt));
  g.fd_p = make_fastdiv(static_cast<uint32_t>(p));
  g.ncol = static_cast<int32_t>(ncol);
  g.fd_ncol = make_fastdiv(static_cast<uint32_t>(ncol));
  g.tw_small = a.tw_small;
  g.tw_lo = a.tw_lo;
  g.tw_hi = a.tw_hi;
  g.conj_in = a.conj_in;
  g.conj_out = a.conj_out;
  g.scale = a.scale;
  g.units = a.units;
  g.block_store = (mode == 1 || (rows && p == 1 && a.dst_mode == kDstComplex && a.go.bs == n)) ? 1 : 0;
  if (mode == 1 && a.go.bs != n) throw std::runtime_error("convolution pass needs contiguous sequences");
  g.src_mode = a.src_mode;
  g.dst_mode = a.dst_mode;
  g.nlog = a.nlog;
  g.chirp = a.chirp;
  g.bhat = a.bhat;
  const int64_t nblocks = (a.units + cu - 1) / cu;
  g.nblocks = static_cast<int32_t>(nblocks);
  g.xcd_remap = opt_.xcd_remap;
  sst.ept = ept;
  sst.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(cu * tpu)),
                          static_cast<size_t>(cu) * LS * eb()};
  sst.bytes_moved =
      s.nbatch * ncol * (src_bytes(a.src_mode, n, a.nlog, eb()) + dst_bytes(a.dst_mode, n, a.nlog, eb()));
  std::ostringstream lab;
  lab << "k_stage<" << (dbl() ? "fp64" : "fp32") << "," << ept << "," << cu << "," << mode << ">[";
  for (size_t ri = 0; ri < radices.size(); ++ri) lab << (ri ? "x" : "") << radices[ri];
  lab << "] " << (mode == 1 ? "convolution" : "pass") << " " << pass << "/" << npasses << " R=" << R << " p=" << p
      << " tcnt=" << tcnt << (g.block_store ? " block-store" : "") << (cols ? " columns=" : "");
  if (cols) lab << ncol;
  sst.label = lab.str();
  ex_.steps.push_back(sst);
  return true;
}

void Builder::emit_pass_step(const PassArgs &a0, int src, int dst, const SeqSet &s, int64_t n, int pass,
                             int npasses) {
  Step st;
  st.kind = Step::Pass;
  st.dbl = dbl();
  st.src = src;
  st.dst = dst;
  PassArgs a = a0;
  const int R = a.R;
  const int64_t tcnt = a.tcnt, p = a.p;
  const int LS = odd_stride(lds_pad(R - 1) + 1);
  const int64_t in_i_stride = a.gi.es;
  const int64_t out_i_stride = (p > 1) ? a.go.es : static_cast<int64_t>(R) * a.go.es;
  const bool lsf = slot_fast(tcnt * a.gi.es, unit_stride(s.col_fast, tcnt, s.ncol, a.gi, in_i_stride));
  const bool ssf = slot_fast(p * a.go.es, unit_stride(s.col_fast, tcnt, s.ncol, a.go, out_i_stride));
  ShapeChoice sh = pick_shape(opt_, eb(), R, LS, true, a.units);
  a.C = sh.C;
  a.tpu = sh.tpu;
  a.lds_stride = LS;
  a.load_slot_fast = lsf;
  a.store_slot_fast = ssf;
  a.fd_C = make_fastdiv(static_cast<uint32_t>(sh.C));
  a.fd_R = make_fastdiv(static_cast<uint32_t>(R));
  a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
  a.fd_tcnt = make_fastdiv(static_cast<uint32_t>(tcnt));
  a.fd_p = make_fastdiv(static_cast<uint32_t>(p));
  const int64_t nblocks = (a.units + sh.C - 1) / sh.C;
  a.nblocks = static_cast<int32_t>(nblocks);
  a.xcd_remap = opt_.xcd_remap;
  a.direct = 0;
  st.pa = a;
  st.ept = sh.ept;
  st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(sh.C * sh.tpu)), sh.lds};
  st.bytes_moved = s.nbatch * s.ncol *
                   (src_bytes(a.src_mode, n, a.nlog, eb()) + dst_bytes(a.dst_mode, n, a.nlog, eb()));
  std::ostringstream lab;
  lab << "k_pass<" << (dbl() ? "fp64" : "fp32") << "," << sh.ept << "> stockham-pass " << pass << "/" << npasses
      << " R=" << R << " p=" << p << " tcnt=" << tcnt << " C=" << sh.C << " tpu=" << sh.tpu;
  st.label = lab.str();
  ex_.steps.push_back(st);
}

void Builder::emit_multipass(const Core &c, const SeqSet &s, int src, int dst, const Modes &first, const Modes &last,
                             std::initializer_list<int> avoid_extra) {
  const size_t k = c.passes.size();
  const int64_t n = c.n;
  const Geometry gws = s.col_fast ? Geometry{s.ncol, 1, n * s.ncol} : Geometry{1, n, n * s.ncol};
  const size_t wsbytes = static_cast<size_t>(s.nbatch * s.ncol * n) * eb();
  std::vector<int> avoid = {src, dst};
  avoid.insert(avoid.end(), avoid_extra.begin(), avoid_extra.end());
  auto pick = [&](std::vector<int> av) {
    for (int id = kBufWs0;; ++id) {
      if (std::find(av.begin(), av.end(), id) != av.end()) continue;
      size_t idx = static_cast<size_t>(id - kBufWs0);
      if (ex_.ws_bytes.size() <= idx) ex_.ws_bytes.resize(idx + 1, 0);
      ex_.ws_bytes[idx] = std::max(ex_.ws_bytes[idx], wsbytes);
      return id;
    }
  };
  const int wA = pick(avoid);
  int wB = -1;
  if (k > 2) {
    std::vector<int> av2 = avoid;
    av2.push_back(wA);
    wB = pick(av2);
  }
  const TwoLevel tl = two_level(n);
  int64_t p = 1;
  for (size_t j = 0; j < k; ++j) {
    const int R = static_cast<int>(c.passes[j]);
    const int st_src = (j == 0) ? src : ((j % 2 == 1) ? wA : wB);
    const int st_dst = (j + 1 == k) ? dst : (((j + 1) % 2 == 1) ? wA : wB);
    PassArgs a = chain_pass(c.pass_radices[j], R, p, n, s, (j == 0) ? s.gi : gws, (j + 1 == k) ? s.go : gws, tl);
    a.src_mode = (j == 0) ? first.src_mode : kSrcComplex;
    a.conj_in = (j == 0) ? first.conj_in : 0;
    a.dst_mode = (j + 1 == k) ? last.dst_mode : kDstComplex;
    a.conj_out = (j + 1 == k) ? last.conj_out : 0;
    a.scale = (j + 1 == k) ? last.scale : 1.0;
    const int64_t nl = (j == 0) ? first.nlog : last.nlog;
    a.nlog = nl ? nl : n;
    if ((a.src_mode & kSrcChirpFlag) || (a.dst_mode & kDstChirpFlag)) a.chirp = chirp(a.nlog);
    // the pointwise pass of a Bluestein chain gets its table from emit_bluestein
    if (!try_stage(a, c.pass_radices[j], 0, st_src, st_dst, s, n, static_cast<int>(j), static_cast<int>(k)))
      emit_pass_step(a, st_src, st_dst, s, n, static_cast<int>(j), static_cast<int>(k));
    p *= R;
  }
}

void Builder::emit_rader(const Core &c, const SeqSet &s, int src, int dst, const Modes &m) {
  Step st;
  st.kind = Step::Rader;
  st.dbl = dbl();
  st.src = src;
  st.dst = dst;
  FusedArgs &a = st.fa;
  const int p = static_cast<int>(c.n), mm = static_cast<int>(c.m);
  a.gi = s.gi;
  a.go = s.go;
  a.nseqs = s.nbatch * s.ncol;
  a.ncol = static_cast<int32_t>(s.ncol);
  a.n = p;
  a.m = mm;
  fill_stages(a.rad, a.fd_q, a.twq_step, a.nst, c.radices, mm);
  a.tw_small = roots(mm, mm, 1);
  a.aux = rader_idx(p, c.root);
  a.bhat = bhat_rader(p, c.root);
  a.chirp = nullptr;
  const int LS = odd_stride(lds_pad(p) + 1 + lds_pad(mm) + 1 + 1);
  const int64_t seq_stride_in = s.ncol > 1 ? s.gi.cs : s.gi.bs;
  const int64_t seq_stride_out = s.ncol > 1 ? s.go.cs : s.go.bs;
  const bool lsf = slot_fast(s.gi.es, seq_stride_in);
  const bool ssf = slot_fast(s.go.es, seq_stride_out);
  ShapeChoice sh = pick_shape(opt_, eb(), mm, LS, lsf || ssf, a.nseqs);
  a.C = sh.C;
  a.tpu = sh.tpu;
  a.lds_stride = LS;
  a.load_slot_fast = lsf;
  a.store_slot_fast = ssf;
  a.fd_C = make_fastdiv(static_cast<uint32_t>(sh.C));
  a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
  a.fd_m = make_fastdiv(static_cast<uint32_t>(mm));
  a.lext = p;
  a.sext = p;
  a.fd_lext = make_fastdiv(static_cast<uint32_t>(p));
  a.fd_sext = make_fastdiv(static_cast<uint32_t>(p));
  a.src_mode = m.src_mode & 3;
  a.dst_mode = m.dst_mode & 3;
  a.conj_in = m.conj_in;
  a.conj_out = m.conj_out;
  a.scale = m.scale;
  const int64_t nblocks = (a.nseqs + sh.C - 1) / sh.C;
  st.ept = sh.ept;
  st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(sh.C * sh.tpu)), sh.lds};
  st.bytes_moved = a.nseqs * (src_bytes(a.src_mode, p, p, eb()) + dst_bytes(a.dst_mode, p, p, eb()));
  std::ostringstream lab;
  lab << "k_rader<" << (dbl() ? "fp64" : "fp32") << "," << sh.ept << "> p=" << p << " m=" << mm << " C=" << sh.C
      << " tpu=" << sh.tpu;
  st.label = lab.str();
  ex_.steps.push_back(st);
}

void Builder::emit_matrix_core(const Core &c, const SeqSet &s, int src, int dst, const Modes &m) {
  if (c.n != 1024 || dbl() || s.col_fast || s.ncol != 1 || s.gi.es != 1 || s.go.es != 1 || m.src_mode != kSrcComplex ||
      m.dst_mode != kDstComplex)
    throw std::invalid_argument("matrix-core transform needs contiguous complex FP32 rows of length 1024");
  Step st;
  st.kind = Step::Mfma;
  st.src = src;
  st.dst = dst;
  st.compute = c.compute;
  MfmaArgs &a = st.ma;
  a.nseq = s.nbatch;
  a.idist = s.gi.bs;
  a.odist = s.go.bs;
  a.tw = roots(1024, 1024, 1);
  a.conj_in = m.conj_in;
  a.conj_out = m.conj_out;
  a.scale = m.scale;
  // 4 waves per workgroup (one per SIMD), each wave loops over transforms;
  // one workgroup per CU, all resident from the start (no workgroup starts
  // after the first ones finish).  A tuned plan's variant picks the kernel
  // instance, its units the workgroup count.
  if (opt_.kernel_variant > mfma_variants()) throw std::invalid_argument("no such matrix-core kernel variant");
  const int variant = opt_.kernel_variant > 0 ? opt_.kernel_variant - 1 : 0;
  st.leaf_id = variant;
  st.groups_max = opt_.forced_units > 0 ? opt_.forced_units : cu_count();
  const int64_t groups = std::max<int64_t>(1, std::min<int64_t>((s.nbatch + 3) / 4, st.groups_max));
  st.shape = LaunchShape{dim3(static_cast<unsigned>(groups)), dim3(256), 0};
  st.bytes_moved = s.nbatch * 1024 * 8 * 2;
  std::ostringstream lab;
  lab << "k_mfma1024<" << (c.compute == kComputeFp16x3 ? "true," : "false,") << mfma_kernel_targs(variant) << "> "
      << (c.compute == kComputeFp16x3 ? "fp16x3" : "bf16x3") << " R=1024 transforms=" << s.nbatch << " groups=" << groups;
  st.label = lab.str();
  ex_.steps.push_back(st);
}

// Convolution of length M over `split` workgroups per sequence: the split
// kernel convolves block h of the radix-split first stage and parks it, the
// combine kernel applies the radix-split last stage, the chirp and the sink.
void Builder::emit_bluestein_split(const Core &c, const SeqSet &s, int src, int dst, const Modes &m, int split) {
  const int64_t n = c.n, M = c.m;
  const int id = bluestein_split_kernel_id(M, dbl(), split);
  const std::vector<int> rad = bluestein_lds_kernel_radices(id);
  const int64_t ml = M / split;
  if (2 * n > M) throw std::runtime_error("split LDS Bluestein needs n <= M / 2");
  const int64_t nseqs = s.nbatch * s.ncol;
  const int park = pick_ws({src, dst}, static_cast<size_t>(nseqs * M) * eb());
  std::vector<int> digits = rad;
  digits.insert(digits.begin(), split);
  const void *tw = roots(M, M, 1);
  const void *ch = chirp(n);
  Step st;
  st.kind = Step::BluesteinLds;
  st.dbl = dbl();
  st.src = src;
  st.dst = dst;
  st.aux = park;
  st.split = split;
  st.leaf_id = id;
  st.leaf_mode = 0;
  BluesteinLdsArgs &a = st.ba;
  a.gi = s.gi;
  a.go = s.go;
  a.ncol = static_cast<int32_t>(s.ncol);
  a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
  a.n = n;
  a.tw = tw;
  a.chirp = ch;
  a.bhat = bhat_bluestein_digit_reversed(n, M, digits);
  {
    const std::string key = "chirpsplit:" + std::to_string(eb()) + ":" + std::to_string(n) + ":" +
                            std::to_string(M) + ":" + std::to_string(split);
    auto it = ex_.tables.find(key);
    if (it == ex_.tables.end()) {
      auto buf = std::make_shared<DevBuf>(static_cast<size_t>(n * split) * eb());
      hip_check(gen_chirp_split(buf->ptr, dbl(), n, M, split, st_), "gen_chirp_split");
      it = ex_.tables.emplace(key, buf).first;
    }
    a.chirp_split = it->second->ptr;
  }
  a.src_mode = m.src_mode & 3;
  a.dst_mode = m.dst_mode & 3;
  a.conj_in = m.conj_in;
  a.conj_out = m.conj_out;
  a.scale = m.scale;
  const int tpu = bluestein_lds_threads(id, 0);
  const int64_t last = (ml - 1) + ((ml - 1) >> 6);
  st.ept = static_cast<int>((ml + tpu - 1) / tpu);
  st.shape = LaunchShape{dim3(static_cast<unsigned>(nseqs * split)), dim3(static_cast<unsigned>(tpu)),
                         static_cast<size_t>(last + 1) * eb()};
  st.bytes_moved = nseqs * src_bytes(a.src_mode, n, n, eb());
  const char *pr = dbl() ? "fp64" : "fp32";
  std::ostringstream lab;
  lab << "k_bluestein_split<" << pr << "," << tpu << "," << split << ">[";
  for (size_t i = 0; i < rad.size(); ++i) lab << (i ? "x" : "") << rad[i];
  lab << "] n=" << n << " M=" << M << " transforms=" << nseqs;
  st.label = lab.str();
  ex_.steps.push_back(st);

  Step cb;
  cb.kind = Step::BluesteinCombine;
  cb.dbl = dbl();
  cb.src = park;
  cb.dst = dst;
  cb.split = split;
  BluesteinCombineArgs &b = cb.ca;
  b.go = s.go;
  b.ncol = static_cast<int32_t>(s.ncol);
  b.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
  b.nb = static_cast<int32_t>((n + 255) / 256);
  b.fd_nb = make_fastdiv(static_cast<uint32_t>(b.nb));
  b.n = static_cast<int32_t>(n);
  b.ml = static_cast<int32_t>(ml);
  b.tw = tw;
  b.chirp = ch;
  b.dst_mode = m.dst_mode & 3;
  b.conj_out = m.conj_out;
  b.scale = m.scale;
  cb.shape = LaunchShape{dim3(static_cast<unsigned>(nseqs * b.nb)), dim3(256), 0};
  cb.bytes_moved = nseqs * dst_bytes(b.dst_mode, n, n, eb());
  std::ostringstream lb;
  lb << "k_bluestein_combine<" << pr << "," << split << "> n=" << n << " M=" << M << " transforms=" << nseqs;
  cb.label = lb.str();
  ex_.steps.push_back(cb);
}

void Builder::emit_bluestein(const Core &c, const SeqSet &s, int src, int dst, const Modes &m) {
  const int64_t n = c.n, M = c.m;
  if (c.lds_conv) {
    const int64_t nseqs_all = s.nbatch * s.ncol;
    // Few sequences: spread each convolution over 2 or 4 workgroups (blocks
    // of a radix-P first stage) plus a combine pass, so the launch fills the
    // GPU.  Tuning (units): 5 one workgroup per sequence, 6 split in 2, 7
    // split in 4.
    int split = 0;
    if (opt_.forced_units == 6) {
      split = 2;
    } else if (opt_.forced_units == 7) {
      split = 4;
    } else if (opt_.forced_units < 1 || opt_.forced_units > 5) {
      // measured on MI355X: a whole convolution per workgroup stays ahead
      // until fewer than 1/8 of the CUs have a sequence; the two-halves form
      // (two convolutions in a row per workgroup) splits in 2 below 1/2
      // (in 4 below 1/8: the combine pass reads P blocks, so P = 2 costs
      // less device time when the GPU is shared)
      const int ncu = cu_count();
      if (nseqs_all * 8 <= ncu)
        split = 4;
      else if (c.lds_halves && nseqs_all * 2 <= ncu)
        split = 2;
    }
    if (split > 0 && bluestein_split_kernel_id(M, dbl(), split) < 0) split = split == 4 ? 2 : 0;
    if (split > 0 && bluestein_split_kernel_id(M, dbl(), split) < 0) split = 0;
    if (split > 0) {
      emit_bluestein_split(c, s, src, dst, m, split);
      return;
    }
    const int id = bluestein_lds_kernel_id(c.radices, dbl(), c.lds_halves);
    if (id < 0) throw std::runtime_error("no LDS Bluestein instance for this convolution length");
    Step st;
    st.kind = Step::BluesteinLds;
    st.dbl = dbl();
    st.src = src;
    st.dst = dst;
    st.leaf_id = id;
    // tuning (units), FP32 instances with Stockham kernels: 1 Stockham with
    // prefetched base twiddles, 2 Stockham, 3 in-place with table twiddles,
    // 4 (default) in-place with binary-powered twiddles
    int variant = 3;
    if (bluestein_lds_has_stockham(id))
      variant = opt_.forced_units == 1 ? 1 : opt_.forced_units == 2 ? 0 : opt_.forced_units == 3 ? 2 : 3;
    const bool in_place = variant >= 2;
    if (in_place && 2 * n > M) throw std::runtime_error("in-place LDS Bluestein needs n <= M / 2");
    st.leaf_mode = variant;
    const int64_t nseqs = s.nbatch * s.ncol;
    BluesteinLdsArgs &a = st.ba;
    a.gi = s.gi;
    a.go = s.go;
    a.ncol = static_cast<int32_t>(s.ncol);
    a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
    a.n = n;
    a.tw = roots(M, M, 1);
    a.chirp = chirp(n);
    if (!in_place) {
      a.bhat = bhat_bluestein(n, M);
    } else {
      std::vector<int> digits = c.radices;
      if (c.lds_halves) digits.insert(digits.begin(), 2);
      a.bhat = bhat_bluestein_digit_reversed(n, M, digits);
    }
    if (c.lds_halves) st.aux = pick_ws({src, dst}, static_cast<size_t>(nseqs * n) * eb());
    a.src_mode = m.src_mode & 3;
    a.dst_mode = m.dst_mode & 3;
    a.conj_in = m.conj_in;
    a.conj_out = m.conj_out;
    a.scale = m.scale;
    const int tpu = bluestein_lds_threads(id, variant);
    const int64_t ml = c.lds_halves ? M / 2 : M;  // points held in LDS
    const int64_t last = (ml - 1) + ((ml - 1) >> 6);
    st.ept = in_place ? static_cast<int>((ml + tpu - 1) / tpu) : bluestein_lds_ept(id);
    st.shape = LaunchShape{dim3(static_cast<unsigned>(nseqs)), dim3(static_cast<unsigned>(tpu)),
                           static_cast<size_t>(last + 1) * eb()};
    st.bytes_moved = nseqs * (src_bytes(a.src_mode, n, n, eb()) + dst_bytes(a.dst_mode, n, n, eb()));
    std::ostringstream lab;
    const char *pr = dbl() ? "fp64" : "fp32";
    if (c.lds_halves)
      lab << "k_bluestein_halves<" << pr << "," << tpu << ">[";
    else if (in_place)
      lab << "k_bluestein_ip<" << pr << "," << tpu << "," << (variant == 2 ? "true" : "false") << ">[";
    else
      lab << "k_bluestein_lds<" << pr << "," << st.ept << "," << (variant == 1 ? "true" : "false") << ">[";
    for (size_t i = 0; i < c.radices.size(); ++i) lab << (i ? "x" : "") << c.radices[i];
    lab << "] n=" << n << " M=" << M << " transforms=" << nseqs;
    st.label = lab.str();
    ex_.steps.push_back(st);
    return;
  }
  if (c.conv && c.conv->passes.size() == 2 && !opt_.generic_kernels && !s.col_fast && s.ncol == 1 &&
      s.gi.es == 1 && s.go.es == 1 && (m.src_mode & 3) == kSrcComplex && (m.dst_mode & 3) == kDstComplex &&
      stage_kernel_id(c.conv->pass_radices[0]) >= 0 && stage_kernel_id(c.conv->pass_radices[1]) >= 0) {
    // Three launches over M = R0 * R1:
    //   A  chirp * x, forward pass 0 (R0)                        user -> wA
    //   B  forward pass 1 (R1), conj(X * bhat), inverse pass 0 (R1, the
    //      inverse chain runs its super-radices in the opposite order)  wA -> wB
    //   C  inverse pass 1 (R0), conj, * chirp, first n              wB -> user
    const Core &cv = *c.conv;
    const int R0 = static_cast<int>(cv.passes[0]), R1 = static_cast<int>(cv.passes[1]);
    const Geometry gws{1, M, M};
    const size_t wbytes = static_cast<size_t>(s.nbatch * M) * eb();
    const int wA = pick_ws({src, dst}, wbytes);
    const int wB = pick_ws({src, dst, wA}, wbytes);
    const TwoLevel tl = two_level(M);
    const void *ch = chirp(n);
    const void *bh = bhat_bluestein(n, M);
    SeqSet sw = s;
    PassArgs a = chain_pass(cv.pass_radices[0], R0, 1, M, sw, s.gi, gws, tl);
    a.src_mode = (m.src_mode & 3) | kSrcChirpFlag;
    a.conj_in = m.conj_in;
    a.nlog = n;
    a.chirp = ch;
    if (!try_stage(a, cv.pass_radices[0], 0, src, wA, sw, M, 0, 3)) throw std::runtime_error("bluestein pass A");
    PassArgs b = chain_pass(cv.pass_radices[1], R1, R0, M, sw, gws, gws, tl);
    b.bhat = bh;
    if (!try_stage(b, cv.pass_radices[1], 1, wA, wB, sw, M, 1, 3)) throw std::runtime_error("bluestein pass B");
    PassArgs cc = chain_pass(cv.pass_radices[0], R0, R1, M, sw, gws, s.go, tl);
    cc.dst_mode = (m.dst_mode & 3) | kDstChirpFlag;
    cc.conj_out = m.conj_out;
    cc.scale = m.scale;
    cc.nlog = n;
    cc.chirp = ch;
    if (!try_stage(cc, cv.pass_radices[0], 0, wB, dst, sw, M, 2, 3)) throw std::runtime_error("bluestein pass C");
    return;
  }
  if (c.conv) {
    // two multi-pass length-M transforms; chirp / pointwise / finalize fused into their end passes
    const Geometry gws = s.col_fast ? Geometry{s.ncol, 1, M * s.ncol} : Geometry{1, M, M * s.ncol};
    const int wP = pick_ws({src, dst}, static_cast<size_t>(s.nbatch * s.ncol * M) * eb());
    SeqSet sa = s;
    sa.go = gws;
    SeqSet sb = s;
    sb.gi = gws;
    Modes first = m;
    first.src_mode = (m.src_mode & 3) | kSrcChirpFlag;
    first.nlog = n;
    Modes mid;
    mid.dst_mode = kDstPointwiseFlag;
    mid.nlog = n;
    emit_multipass(*c.conv, sa, src, wP, first, mid, {dst});
    const void *bh = bhat_bluestein(n, M);
    Step &pw = ex_.steps.back();
    if (pw.kind == Step::Stage || pw.kind == Step::Col)
      pw.sa.bhat = bh;
    else
      pw.pa.bhat = bh;
    Modes mfirst;
    mfirst.nlog = n;
    Modes fin = m;
    fin.dst_mode = (m.dst_mode & 3) | kDstChirpFlag;
    fin.nlog = n;
    emit_multipass(*c.conv, sb, wP, dst, mfirst, fin, {src});
    return;
  }
  Step st;
  st.kind = Step::Bluestein;
  st.dbl = dbl();
  st.src = src;
  st.dst = dst;
  FusedArgs &a = st.fa;
  a.gi = s.gi;
  a.go = s.go;
  a.nseqs = s.nbatch * s.ncol;
  a.ncol = static_cast<int32_t>(s.ncol);
  a.n = static_cast<int32_t>(n);
  a.m = static_cast<int32_t>(M);
  fill_stages(a.rad, a.fd_q, a.twq_step, a.nst, c.radices, M);
  a.tw_small = roots(M, M, 1);
  a.chirp = chirp(n);
  a.bhat = bhat_bluestein(n, M);
  a.aux = nullptr;
  const int LS = odd_stride(lds_pad(static_cast<int>(M) - 1) + 1);
  const int64_t seq_stride_in = s.ncol > 1 ? s.gi.cs : s.gi.bs;
  const int64_t seq_stride_out = s.ncol > 1 ? s.go.cs : s.go.bs;
  const bool lsf = slot_fast(s.gi.es, seq_stride_in);
  const bool ssf = slot_fast(s.go.es, seq_stride_out);
  ShapeChoice sh = pick_shape(opt_, eb(), M, LS, lsf || ssf, a.nseqs);
  a.C = sh.C;
  a.tpu = sh.tpu;
  a.lds_stride = LS;
  a.load_slot_fast = lsf;
  a.store_slot_fast = ssf;
  a.fd_C = make_fastdiv(static_cast<uint32_t>(sh.C));
  a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
  a.fd_m = make_fastdiv(static_cast<uint32_t>(M));
  a.lext = static_cast<int32_t>(n);
  a.sext = static_cast<int32_t>(n);
  a.fd_lext = make_fastdiv(static_cast<uint32_t>(n));
  a.fd_sext = make_fastdiv(static_cast<uint32_t>(n));
  a.src_mode = m.src_mode & 3;
  a.dst_mode = m.dst_mode & 3;
  a.conj_in = m.conj_in;
  a.conj_out = m.conj_out;
  a.scale = m.scale;
  const int64_t nblocks = (a.nseqs + sh.C - 1) / sh.C;
  st.ept = sh.ept;
  st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(sh.C * sh.tpu)), sh.lds};
  st.bytes_moved = a.nseqs * (src_bytes(a.src_mode, n, n, eb()) + dst_bytes(a.dst_mode, n, n, eb()));
  std::ostringstream lab;
  lab << "k_bluestein<" << (dbl() ? "fp64" : "fp32") << "," << sh.ept << "> n=" << n << " M=" << M << " C=" << sh.C
      << " tpu=" << sh.tpu;
  st.label = lab.str();
  ex_.steps.push_back(st);
}

// Few contiguous complex sequences of a length with a split instance: one
// launch of P workgroups per sequence instead of one workgroup per sequence
// (or a multi-pass chain), while the P x batch workgroups fit twice the CUs
// (measured: 16384 x 96 split 8 ways 26.8 us against 22.0 for the two-pass
// chain, 10935 x 110 split 5 ways 23.2 us against 16.7 for one workgroup
// per sequence).  Any explicit tuning override keeps the planner's form, so
// the autotuner weighs both.
bool Builder::try_leaf_split(const Core &c, const SeqSet &s, int src, int dst, const Modes &m) {
  if (opt_.generic_kernels || opt_.forced_units != 0 || opt_.ept != 0 || !opt_.forced_passes.empty() ||
      !opt_.forced_radices.empty())
    return false;
  if ((c.kind != CoreKind::Leaf && c.kind != CoreKind::MultiPass) || s.col_fast || s.ncol != 1 || s.gi.es != 1 ||
      s.go.es != 1 || m.src_mode != kSrcComplex || m.dst_mode != kDstComplex)
    return false;
  if (opt_.in_place && src == kBufIn && dst == kBufOut) return false;
  const int id = leaf_split_kernel_id(c.n);
  if (id < 0) return false;
  int sp = 1, sl = 0, se = 0;
  std::vector<int> srad;
  leaf_split_shape(id, dbl(), sp, sl, se, srad);
  if (s.nbatch * sp > 2 * cu_count()) return false;
  emit_leaf_split(id, kRowComplex, c.n, s.nbatch, s.gi.bs, s.go.bs, src, dst, m.conj_in, m.conj_out, m.scale);
  return true;
}

// Real rows of even length n whose n/2-point core has a two-way split
// instance, in small batches: the split leaf with the pack / unpack fused.
bool Builder::try_leaf_split_real(const Plan1D &pl, const SeqSet &s, int src, int dst) {
  if (opt_.generic_kernels || opt_.forced_units != 0 || opt_.ept != 0 || !opt_.forced_radices.empty())
    return false;
  const int64_t h = pl.n / 2;
  if (s.ncol != 1 || s.col_fast || s.gi.es != 1 || s.go.es != 1 || h < 2 ||
      (pl.real_fwd ? s.gi.bs % 2 != 0 : s.go.bs % 2 != 0) || (opt_.in_place && src == kBufIn && dst == kBufOut))
    return false;
  const int id = leaf_split_kernel_id(h);
  if (id < 0 || leaf_split_kernel(id, dbl(), pl.real_fwd ? kRowR2C : kRowC2R) == nullptr ||
      s.nbatch * 2 > cu_count())
    return false;
  emit_leaf_split(id, pl.real_fwd ? kRowR2C : kRowC2R, h, s.nbatch, pl.real_fwd ? s.gi.bs / 2 : s.gi.bs,
                  pl.real_fwd ? s.go.bs : s.go.bs / 2, src, dst, false, !pl.real_fwd, 1.0);
  return true;
}

void Builder::emit_leaf_split(int id, int mode, int64_t n, int64_t nseq, int64_t idist, int64_t odist, int src,
                              int dst, bool conj_in, bool conj_out, double scale) {
  int p = 1, l = 1, ept = 0;
  std::vector<int> rad;
  leaf_split_shape(id, dbl(), p, l, ept, rad);
  // each of the P workgroups of a sequence reads all of it
  if (src == kBufIn && dst == kBufOut) ex_.inplace_hazard = true;
  const int tpu = (l + ept - 1) / ept;
  Step st;
  st.kind = Step::LeafSplit;
  st.dbl = dbl();
  st.src = src;
  st.dst = dst;
  st.leaf_id = id;
  st.leaf_mode = mode;
  st.split = p;
  LeafSplitArgs &a = st.lsa;
  a.nseq = nseq;
  a.idist = idist;
  a.odist = odist;
  a.tw_n = roots(n, n, 1);
  a.tw_l = roots(l, l, 1);
  a.aux = mode == kRowComplex ? nullptr : roots(n + 1, 2 * n, 1);
  a.conj_in = conj_in;
  a.conj_out = conj_out;
  a.scale = scale;
  const int64_t lds = (l - 1) + ((l - 1) >> 5) + 1;
  st.ept = ept;
  st.shape = LaunchShape{dim3(static_cast<unsigned>(nseq * p)), dim3(static_cast<unsigned>(tpu)),
                         static_cast<size_t>(lds) * eb()};
  st.bytes_moved = mode == kRowComplex ? nseq * n * eb() * 2 : nseq * (n * eb() + (n + 1) * eb());
  std::ostringstream lab;
  lab << "k_leaf_split<" << (dbl() ? "fp64" : "fp32") << "," << ept << "," << p << "," << mode << ">[";
  for (size_t i = 0; i < rad.size(); ++i) lab << (i ? "x" : "") << rad[i];
  lab << "] " << (mode == kRowComplex ? "C2C" : mode == kRowR2C ? "R2C" : "C2R") << " n=" << n
      << " transforms=" << nseq;
  st.label = lab.str();
  ex_.steps.push_back(st);
}

void Builder::emit_core(const Core &c, const SeqSet &s, int src, int dst, const Modes &m) {
  if (try_leaf_split(c, s, src, dst, m)) return;
  switch (c.kind) {
    case CoreKind::Leaf: emit_leaf(c, s, src, dst, m); break;
    case CoreKind::MultiPass: emit_multipass(c, s, src, dst, m, m, {}); break;
    case CoreKind::Rader: emit_rader(c, s, src, dst, m); break;
    case CoreKind::Bluestein: emit_bluestein(c, s, src, dst, m); break;
    case CoreKind::MatrixCore: emit_matrix_core(c, s, src, dst, m); break;
  }
}

void Builder::emit_plan1d(const Plan1D &pl, const SeqSet &s, int src, int dst, bool inverse) {
  const int64_t n = pl.n;
  if (pl.real_form == RealForm::None) {
    Modes m;
    m.conj_in = inverse;
    m.conj_out = inverse;
    m.nlog = n;
    emit_core(pl.core, s, src, dst, m);
    return;
  }
  if (pl.real_form == RealForm::Full) {
    Modes m;
    m.nlog = n;
    if (pl.real_fwd) {
      m.src_mode = kSrcReal;
      m.dst_mode = kDstHalf;
    } else {
      m.src_mode = kSrcHermitian;
      m.dst_mode = kDstReal;
      m.conj_in = true;
      m.conj_out = true;
    }
    emit_core(pl.core, s, src, dst, m);
    return;
  }
  const int64_t h = n / 2;
  if (pl.real_form == RealForm::EvenFused && pair_ == 0 && try_leaf_split_real(pl, s, src, dst)) return;
  if (pl.real_form == RealForm::EvenFused) {
    const int lid = opt_.generic_kernels ? -1 : leaf_kernel_id(pl.core.radices);
    if (lid >= 0 && s.ncol == 1 && !s.col_fast && s.gi.es == 1 && s.go.es == 1 && h > 1 &&
        (pl.real_fwd ? s.gi.bs % 2 == 0 : s.go.bs % 2 == 0)) {
      // specialised leaf with the split-spectrum pack / unpack fused
      Step st;
      st.kind = Step::Leaf;
      st.leaf_mode = pl.real_fwd ? 1 : 2;
      st.leaf_id = lid;
      st.dbl = dbl();
      st.src = src;
      st.dst = dst;
      LeafArgs &a = st.la;
      const int H = static_cast<int>(h);
      const int ept = leaf_kernel_ept(lid, dbl());
      const int tpu = (H + ept - 1) / ept;
      const int LS = odd_stride(lds_pad(H - 1) + 1);
      a.nseq = s.nbatch;
      a.idist = pl.real_fwd ? s.gi.bs / 2 : s.gi.bs;  // complex-pair units
      a.odist = pl.real_fwd ? s.go.bs : s.go.bs / 2;
      a.tw = roots(h, h, 1);
      a.aux = roots(h + 1, n, 1);
      a.conj_in = 0;
      a.conj_out = pl.real_fwd ? 0 : 1;
      a.scale = 1.0;
      a.dc_nyq_pair = pair_ != 0 ? 1 : 0;
      if (pair_ != 0) ++pair_steps_;
      int C = std::max(1, opt_.target_threads / tpu);
      if ((s.nbatch + C - 1) / C < 512) C = static_cast<int>(std::max<int64_t>(1, s.nbatch / 512));
      if (opt_.forced_units > 0) C = opt_.forced_units;
      while (C > 1 && (C * tpu > opt_.max_threads ||
                       static_cast<size_t>(C) * LS * eb() > static_cast<size_t>(opt_.lds_pref_bytes)))
        --C;
      a.C = C;
      const int64_t nblocks = (s.nbatch + C - 1) / C;
      st.ept = ept;
      st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(C * tpu)),
                             static_cast<size_t>(C) * LS * eb()};
      st.bytes_moved = s.nbatch * (n * eb() / 2 + (h + 1) * eb());
      std::ostringstream lab;
      lab << "k_leaf<" << (dbl() ? "fp64" : "fp32") << "," << ept << "," << (pl.real_fwd ? 1 : 2) << ">[";
      for (size_t i = 0; i < pl.core.radices.size(); ++i) lab << (i ? "x" : "") << pl.core.radices[i];
      lab << "] " << (pl.real_fwd ? "R2C" : "C2R") << " n=" << n << " half=" << h << " C=" << C << " tpu=" << tpu
          << (pair_ != 0 ? " dc-nyquist-pair" : "");
      st.label = lab.str();
      ex_.steps.push_back(st);
      return;
    }
    Step st;
    st.kind = pl.real_fwd ? Step::R2CEven : Step::C2REven;
    st.dbl = dbl();
    st.src = src;
    st.dst = dst;
    FusedArgs &a = st.fa;
    a.gi = s.gi;
    a.go = s.go;
    a.nseqs = s.nbatch * s.ncol;
    a.ncol = static_cast<int32_t>(s.ncol);
    a.n = static_cast<int32_t>(n);
    a.m = static_cast<int32_t>(h);
    fill_stages(a.rad, a.fd_q, a.twq_step, a.nst, pl.core.radices, h);
    a.tw_small = roots(h, h, 1);
    a.aux = roots(h + 1, n, 1);
    a.bhat = nullptr;
    a.chirp = nullptr;
    const int LS = odd_stride(lds_pad(static_cast<int>(h) - 1) + 1);
    ShapeChoice sh = pick_shape(opt_, eb(), h, LS, false, a.nseqs);
    a.C = sh.C;
    a.tpu = sh.tpu;
    a.lds_stride = LS;
    a.load_slot_fast = 0;
    a.store_slot_fast = 0;
    a.fd_C = make_fastdiv(static_cast<uint32_t>(sh.C));
    a.fd_ncol = make_fastdiv(static_cast<uint32_t>(s.ncol));
    a.fd_m = make_fastdiv(static_cast<uint32_t>(h));
    a.lext = static_cast<int32_t>(h);
    a.sext = static_cast<int32_t>(pl.real_fwd ? h + 1 : h);
    a.fd_lext = make_fastdiv(static_cast<uint32_t>(a.lext));
    a.fd_sext = make_fastdiv(static_cast<uint32_t>(a.sext));
    a.src_mode = 0;
    a.dst_mode = 0;
    a.conj_in = 0;
    a.conj_out = 0;
    a.scale = 1.0;
    const int64_t nblocks = (a.nseqs + sh.C - 1) / sh.C;
    st.ept = sh.ept;
    st.shape = LaunchShape{dim3(static_cast<unsigned>(nblocks)), dim3(static_cast<unsigned>(sh.C * sh.tpu)), sh.lds};
    st.bytes_moved = a.nseqs * (n * eb() / 2 + (h + 1) * eb());  // n reals one way, h+1 bins the other
    std::ostringstream lab;
    lab << (pl.real_fwd ? "k_r2c_even<" : "k_c2r_even<") << (dbl() ? "fp64" : "fp32") << "," << sh.ept << "> n=" << n
        << " half=" << h << " C=" << sh.C << " tpu=" << sh.tpu;
    st.label = lab.str();
    ex_.steps.push_back(st);
    return;
  }
  // EvenMulti: n/2-point core (multi-pass chain, Rader or Bluestein) plus a split-spectrum pass
  const Geometry gws{1, h, h * s.ncol};
  const size_t wbytes = static_cast<size_t>(s.nbatch * s.ncol * h) * eb();
  const TwoLevel tw = two_level(n);
  if (pl.real_fwd) {
    const int W = pick_ws({src, dst}, wbytes);
    SeqSet sc = s;
    sc.gi = Geometry{1, s.gi.cs / 2, s.gi.bs / 2};
    sc.go = gws;
    Modes m;
    emit_core(pl.core, sc, src, W, m);
    Step st;
    st.kind = Step::RealPost;
    st.dbl = dbl();
    st.src = W;
    st.dst = dst;
    st.ra.gi = gws;
    st.ra.go = s.go;
    st.ra.nseqs = s.nbatch * s.ncol;
    st.ra.ncol = static_cast<int32_t>(s.ncol);
    st.ra.n = static_cast<int32_t>(n);
    st.ra.tw_lo = tw.lo;
    st.ra.tw_hi = tw.hi;
    st.ra.tw_bits = tw.bits;
    st.bytes_moved = st.ra.nseqs * (h + (h + 1)) * eb();
    st.label = "k_real_post n=" + std::to_string(n);
    ex_.steps.push_back(st);
  } else {
    const int W = pick_ws({src, dst}, wbytes);
    Step st;
    st.kind = Step::RealPre;
    st.dbl = dbl();
    st.src = src;
    st.dst = W;
    st.ra.gi = s.gi;
    st.ra.go = gws;
    st.ra.nseqs = s.nbatch * s.ncol;
    st.ra.ncol = static_cast<int32_t>(s.ncol);
    st.ra.n = static_cast<int32_t>(n);
    st.ra.tw_lo = tw.lo;
    st.ra.tw_hi = tw.hi;
    st.ra.tw_bits = tw.bits;
    st.bytes_moved = st.ra.nseqs * ((h + 1) + h) * eb();
    st.label = "k_real_pre n=" + std::to_string(n);
    ex_.steps.push_back(st);
    SeqSet sc = s;
    sc.gi = gws;
    sc.go = Geometry{1, s.go.cs / 2, s.go.bs / 2};
    Modes m;
    m.conj_out = true;
    emit_core(pl.core, sc, W, dst, m);
  }
}

std::unique_ptr<Executable> build_executable(const PlanTree &t, bool inverse, const PlanOptions &o, int64_t idist,
                                             int64_t odist, bool allocate_ws, const Strides *st) {
  auto ex = std::make_unique<Executable>();
  Builder b(t.desc.prec, o, *ex, nullptr);
  const int64_t batch = t.desc.batch;
  if (st != nullptr && t.desc.rank == 2 && (t.desc.real_fwd || t.desc.real_inv)) {
    // real 2D with padded rows (row pitch irow / orow in the side's own
    // elements, batches n0 rows apart, unit element stride), e.g. the
    // in-place layout of 2 (n1 / 2 + 1) reals per row
    const int64_t n0 = t.desc.n0, h1 = t.desc.n1 / 2 + 1;
    const int W = b.pick_ws({kBufIn, kBufOut}, static_cast<size_t>(batch * n0 * h1) * elem_bytes(t.desc.prec));
    SeqSet r;
    r.nbatch = batch * n0;
    SeqSet c;
    c.nbatch = batch;
    c.ncol = h1;
    c.col_fast = true;
    if (t.desc.real_fwd) {
      r.gi = Geometry{1, 0, st->irow};
      r.go = Geometry{1, 0, h1};
      b.emit_plan1d(t.rows, r, kBufIn, W, false);
      c.gi = Geometry{h1, 1, n0 * h1};
      c.go = Geometry{st->orow, 1, n0 * st->orow};
      b.emit_plan1d(t.cols, c, W, kBufOut, false);
    } else {
      c.gi = Geometry{st->irow, 1, n0 * st->irow};
      c.go = Geometry{h1, 1, n0 * h1};
      b.emit_plan1d(t.cols, c, kBufIn, W, true);
      r.gi = Geometry{1, 0, h1};
      r.go = Geometry{1, 0, st->orow};
      b.emit_plan1d(t.rows, r, W, kBufOut, true);
    }
  } else if (st != nullptr) {
    // general complex layouts: the run-time-radix and LDS kernels take any
    // element / row / batch stride; the specialised kernels take the
    // contiguous cases only
    if (t.desc.real_fwd || t.desc.real_inv || t.desc.rank > 2) throw std::runtime_error("strided layouts: C2C/Z2Z rank 1-2");
    SeqSet r;
    r.nbatch = batch;
    r.ncol = t.desc.rank == 2 ? t.desc.n0 : 1;
    r.gi = Geometry{st->istride, st->irow, idist};
    r.go = Geometry{st->ostride, st->orow, odist};
    b.emit_plan1d(t.rows, r, kBufIn, kBufOut, inverse);
    if (t.desc.rank == 2) {
      SeqSet c;
      c.nbatch = batch;
      c.ncol = t.desc.n1;
      c.gi = Geometry{st->orow, st->ostride, odist};
      c.go = c.gi;
      c.col_fast = true;
      b.emit_plan1d(t.cols, c, kBufOut, kBufOut, inverse);
    }
  } else if (t.desc.rank == 1) {
    SeqSet s;
    s.nbatch = batch;
    s.ncol = 1;
    s.gi = Geometry{1, 0, idist};
    s.go = Geometry{1, 0, odist};
    b.emit_plan1d(t.rows, s, kBufIn, kBufOut, inverse);
  } else if (t.desc.rank == 3) {
    // n0 x n1 x n2, n2 contiguous: rows, then the n1-point columns of every
    // n0 slab (the 2D column form with n0 batch entries), then the n0-point
    // columns over n1 x n2' contiguous columns
    const int64_t n0 = t.desc.n0, n1 = t.desc.n1, n2 = t.desc.n2;
    const bool real = t.desc.real_fwd || t.desc.real_inv;
    const int64_t w = real ? n2 / 2 + 1 : n2;  // complex row length between the column passes
    auto rows_of = [&](int64_t in_len, int64_t out_len) {
      SeqSet r;
      r.nbatch = batch * n0 * n1;
      r.gi = Geometry{1, 0, in_len};
      r.go = Geometry{1, 0, out_len};
      return r;
    };
    SeqSet m;
    m.nbatch = batch * n0;
    m.ncol = w;
    m.gi = Geometry{w, 1, n1 * w};
    m.go = m.gi;
    m.col_fast = true;
    SeqSet c;
    c.nbatch = batch;
    c.ncol = n1 * w;
    c.gi = Geometry{n1 * w, 1, n0 * n1 * w};
    c.go = c.gi;
    c.col_fast = true;
    if (!real) {
      b.emit_plan1d(t.rows, rows_of(n2, n2), kBufIn, kBufOut, inverse);
      b.emit_plan1d(t.mid, m, kBufOut, kBufOut, inverse);
      b.emit_plan1d(t.cols, c, kBufOut, kBufOut, inverse);
    } else if (t.desc.real_fwd) {
      // rows into a workspace (see the 2D R2C case), the column passes from it
      const int W = b.pick_ws({kBufIn, kBufOut}, static_cast<size_t>(batch * n0 * n1 * w) * elem_bytes(t.desc.prec));
      b.emit_plan1d(t.rows, rows_of(n2, w), kBufIn, W, false);
      b.emit_plan1d(t.mid, m, W, W, false);
      b.emit_plan1d(t.cols, c, W, kBufOut, false);
    } else {
      const int W = b.pick_ws({kBufIn, kBufOut}, static_cast<size_t>(batch * n0 * n1 * w) * elem_bytes(t.desc.prec));
      b.emit_plan1d(t.cols, c, kBufIn, W, true);
      b.emit_plan1d(t.mid, m, W, W, true);
      b.emit_plan1d(t.rows, rows_of(w, n2), W, kBufOut, true);
    }
  } else {
    const int64_t n0 = t.desc.n0, n1 = t.desc.n1, h1 = n1 / 2 + 1;
    if (!t.desc.real_fwd && !t.desc.real_inv) {
      SeqSet r;
      r.nbatch = batch * n0;
      r.gi = Geometry{1, 0, n1};
      r.go = Geometry{1, 0, n1};
      b.emit_plan1d(t.rows, r, kBufIn, kBufOut, inverse);
      SeqSet c;
      c.nbatch = batch;
      c.ncol = n1;
      c.gi = Geometry{n1, 1, n0 * n1};
      c.go = c.gi;
      c.col_fast = true;
      b.emit_plan1d(t.cols, c, kBufOut, kBufOut, inverse);
    } else if (t.desc.real_fwd) {
      // rows into a workspace (an in-place call would otherwise rewrite the
      // buffer at the complex row pitch while other rows are still read),
      // columns from it into the output: the same two sweeps.  With the DC /
      // Nyquist pair the workspace rows hold n1 / 2 bins (Re X_N in Im X_0)
      // at a pitch of n1 / 2 + 128 bytes: whole 128-byte lines that the
      // column pass reads as aligned segments, rows not a power of two apart
      const int64_t wp = t.dc_nyq_pair ? n1 / 2 + 128 / elem_bytes(t.desc.prec) : h1;
      const int W1 = b.pick_ws({kBufIn, kBufOut}, static_cast<size_t>(batch * n0 * wp) * elem_bytes(t.desc.prec));
      SeqSet r;
      r.nbatch = batch * n0;
      r.gi = Geometry{1, 0, n1};
      r.go = Geometry{1, 0, wp};
      b.set_dc_nyq_pair(t.dc_nyq_pair ? 1 : 0);
      b.emit_plan1d(t.rows, r, kBufIn, W1, false);
      SeqSet c;
      c.nbatch = batch;
      c.ncol = h1;
      c.gi = Geometry{wp, 1, n0 * wp};
      c.go = Geometry{h1, 1, n0 * h1};
      c.col_fast = true;
      b.emit_plan1d(t.cols, c, W1, kBufOut, false);
      if (b.set_dc_nyq_pair(0) != (t.dc_nyq_pair ? 2 : 0)) throw std::logic_error("DC/Nyquist pair: unpaired launch");
    } else {
      // (the paired column pass writes n1 / 2 columns per workspace row)
      const int64_t wp = t.dc_nyq_pair ? n1 / 2 + 128 / elem_bytes(t.desc.prec) : h1;
      const int W2 = b.pick_ws({kBufIn, kBufOut}, static_cast<size_t>(batch * n0 * wp) * elem_bytes(t.desc.prec));
      SeqSet c;
      c.nbatch = batch;
      c.ncol = h1;
      c.gi = Geometry{h1, 1, n0 * h1};
      c.go = Geometry{wp, 1, n0 * wp};
      c.col_fast = true;
      b.set_dc_nyq_pair(t.dc_nyq_pair ? 2 : 0);
      b.emit_plan1d(t.cols, c, kBufIn, W2, true);
      SeqSet r;
      r.nbatch = batch * n0;
      r.gi = Geometry{1, 0, wp};
      r.go = Geometry{1, 0, n1};
      b.emit_plan1d(t.rows, r, W2, kBufOut, true);
      if (b.set_dc_nyq_pair(0) != (t.dc_nyq_pair ? 2 : 0)) throw std::logic_error("DC/Nyquist pair: unpaired launch");
    }
  }
  b.finish();
  ex->built_batch = t.desc.rank == 1 ? t.desc.batch : 1;
  ex->resolve_kernels();
  if (allocate_ws) ex->allocate_workspaces();
  return ex;
}

}  // namespace flagfft
