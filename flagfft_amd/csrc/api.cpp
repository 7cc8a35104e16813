// api.cpp — the extern "C" entry points of include/flagfft.h and include/flagfft_ext.h.
//
// Validation order and result codes follow the reference's C API:
//   flagfftPlanMany            src/exec/c_api_plan.cpp:274-336
//   descriptor acceptance      src/exec/c_api_internal.cpp:5-120
//   build_plan error mapping   src/exec/c_api_plan.cpp:43-245
//   Exec* / SetStream / Destroy / GetPlanDescription  src/exec/c_api_exec.cpp:5-236
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "exec.hpp"
#include "flagfft.h"
#include "flagfft_ext.h"
#include "planner.hpp"
#include "tunedb.hpp"

struct flagfftPlan_t {
  void *impl = nullptr;
};

namespace flagfft {
namespace {

struct Desc {
  int rank = 0;
  std::vector<int64_t> n, inembed, onembed;
  int64_t istride = 1, idist = 0, ostride = 1, odist = 0, batch = 1;
  flagfftType type = FLAGFFT_C2C;
  Prec prec = Prec::F32;
  bool real_fwd = false, real_inv = false;
  int compute = 0;
  bool general = false;  // a C2C/Z2Z layout outside the reference's accepted ones (flagfftxPlanMany)
  Strides strides;
};

struct PlanImpl {
  Desc desc;
  int device = 0;
  PlanTree tree;
  std::unique_ptr<Executable> fwd, inv;
  std::unique_ptr<Executable> fwd_ip, inv_ip;  // in-place calls of executables with an in-place hazard
  hipStream_t stream = nullptr;
  bool graph_mode = false;  // Exec replays one hipGraph per direction
  std::string description;
  std::mutex mu;
};

std::mutex g_handles_mu;
// live handles; a call holds its plan's reference until it returns, so a
// Destroy that overlaps another call on the same handle frees the plan after
// that call instead of under it
std::map<flagfftHandle, std::shared_ptr<PlanImpl>> g_handles;
thread_local PlanOptions t_options;
thread_local bool t_explicit = false;  // flagfftxSetTuning overrides the tuned-plan database

bool overrides_to_options(const TuneOverrides &t, PlanOptions &o) {
  if (t.leaf_max > 0) {
    o.leaf_max_f32 = t.leaf_max;
    o.leaf_max_f64 = t.leaf_max;
  }
  if (t.pass_max > 0) {
    o.pass_max_f32 = t.pass_max;
    o.pass_max_f64 = t.pass_max;
  }
  if (t.units > 0) o.forced_units = t.units;
  if (t.ept == 8 || t.ept == 16 || t.ept == 32)
    o.ept = t.ept;
  else if (t.ept != 0)
    return false;
  if (t.chunk_bytes > 0) o.chunk_bytes = t.chunk_bytes;
  if (t.stage_form < 0 || t.stage_form > 2) return false;
  o.stage_form = t.stage_form;
  if (t.variant < 0) return false;
  o.kernel_variant = t.variant;
  if (t.passes.size() > 4 || t.radices.size() > 16) return false;
  o.forced_passes = t.passes;
  o.forced_radices = t.radices;
  return true;
}

TuneKey tune_key(flagfftType type, int64_t n, int64_t batch) {
  TuneKey k;
  k.n = n;
  k.batch = batch;
  switch (type) {
    case FLAGFFT_C2C: k.dtype = "complex64"; k.layout = "contiguous"; k.direction = "forward"; break;
    case FLAGFFT_Z2Z: k.dtype = "complex128"; k.layout = "contiguous"; k.direction = "forward"; break;
    case FLAGFFT_R2C: k.dtype = "float32"; k.layout = "real"; k.direction = "forward"; break;
    case FLAGFFT_D2Z: k.dtype = "float64"; k.layout = "real"; k.direction = "forward"; break;
    case FLAGFFT_C2R: k.dtype = "complex64"; k.layout = "hermitian"; k.direction = "inverse"; break;
    case FLAGFFT_Z2D: k.dtype = "complex128"; k.layout = "hermitian"; k.direction = "inverse"; break;
  }
  return k;
}

// Options for a new plan: the thread's explicit tuning, else the newest tuned
// row of the database (1D plans; C2C/Z2Z use the forward row for both
// directions), else the defaults.  `source` names the database row used.
PlanOptions plan_options(const Desc &d, std::string &source) {
  source.clear();
  if (t_explicit || d.rank != 1 || d.compute != 0 || d.general) return t_options;
  std::string js;
  if (!tuned_lookup(tune_key(d.type, d.n[0], d.batch), js)) return t_options;
  TuneOverrides t;
  PlanOptions o;
  if (!tuning_from_json(js, t) || !overrides_to_options(t, o)) return t_options;
  source = tuned_db_path() + " " + js;
  return o;
}

flagfftResult type_metadata(flagfftType type, Desc &d) {
  switch (type) {
    case FLAGFFT_C2C: d.prec = Prec::F32; d.real_fwd = false; d.real_inv = false; return FLAGFFT_SUCCESS;
    case FLAGFFT_Z2Z: d.prec = Prec::F64; d.real_fwd = false; d.real_inv = false; return FLAGFFT_SUCCESS;
    case FLAGFFT_R2C: d.prec = Prec::F32; d.real_fwd = true; d.real_inv = false; return FLAGFFT_SUCCESS;
    case FLAGFFT_D2Z: d.prec = Prec::F64; d.real_fwd = true; d.real_inv = false; return FLAGFFT_SUCCESS;
    case FLAGFFT_C2R: d.prec = Prec::F32; d.real_fwd = false; d.real_inv = true; return FLAGFFT_SUCCESS;
    case FLAGFFT_Z2D: d.prec = Prec::F64; d.real_fwd = false; d.real_inv = true; return FLAGFFT_SUCCESS;
  }
  return FLAGFFT_INVALID_TYPE;
}

// src/exec/c_api_internal.cpp:60-90
bool supported_1d(const Desc &d) {
  if (d.rank != 1 || d.batch <= 0 || d.n.size() != 1 || d.n[0] <= 0) return false;
  const int64_t n = d.n[0];
  const int64_t compact = n / 2 + 1, padded = 2 * compact;
  const int64_t in_len = d.real_inv ? compact : n;
  const int64_t out_len = d.real_fwd ? compact : n;
  const bool idist_ok = d.idist == in_len || (d.real_fwd && d.idist == padded);
  const bool odist_ok = d.odist == out_len || (d.real_inv && d.odist == padded);
  const bool iemb_ok = d.inembed.size() == 1 && (d.inembed[0] == in_len || (d.real_fwd && d.inembed[0] == padded));
  const bool oemb_ok = d.onembed.size() == 1 && (d.onembed[0] == out_len || (d.real_inv && d.onembed[0] == padded));
  return d.istride == 1 && d.ostride == 1 && idist_ok && odist_ok && iemb_ok && oemb_ok;
}

// src/exec/c_api_internal.cpp:92-120
bool supported_2d(const Desc &d) {
  if (d.rank != 2 || d.batch <= 0 || d.n.size() != 2 || d.n[0] <= 0 || d.n[1] <= 0) return false;
  if (d.istride != 1 || d.ostride != 1) return false;
  const int64_t n0 = d.n[0], n1 = d.n[1], h1 = n1 / 2 + 1;
  const int64_t in_logical = d.real_inv ? n0 * h1 : n0 * n1;
  const int64_t out_logical = d.real_fwd ? n0 * h1 : n0 * n1;
  return d.idist == in_logical && d.odist == out_logical;
}

// flagfftPlanMany argument checks up to the descriptor (src/exec/c_api_plan.cpp:285-333).
flagfftResult make_desc(int rank, const int *n, const int *inembed, int istride, int idist, const int *onembed,
                        int ostride, int odist, flagfftType type, int batch, Desc &d) {
  if (rank <= 0 || n == nullptr || batch <= 0) return FLAGFFT_INVALID_VALUE;
  for (int i = 0; i < rank; ++i)
    if (n[i] <= 0) return FLAGFFT_INVALID_SIZE;
  d.rank = rank;
  d.n.assign(n, n + rank);
  d.istride = istride;
  d.idist = idist;
  d.ostride = ostride;
  d.odist = odist;
  d.batch = batch;
  d.type = type;
  flagfftResult r = type_metadata(type, d);
  if (r != FLAGFFT_SUCCESS) return r;
  if (inembed != nullptr) d.inembed.assign(inembed, inembed + rank);
  else if (d.real_inv && rank == 1) d.inembed = {d.n[0] / 2 + 1};
  else if (d.real_inv && rank == 2) d.inembed = {d.n[0], d.n[1] / 2 + 1};
  else d.inembed = d.n;
  if (onembed != nullptr) d.onembed.assign(onembed, onembed + rank);
  else if (d.real_fwd && rank == 1) d.onembed = {d.n[0] / 2 + 1};
  else if (d.real_fwd && rank == 2) d.onembed = {d.n[0], d.n[1] / 2 + 1};
  else d.onembed = d.n;
  if (!supported_1d(d) && !supported_2d(d)) return FLAGFFT_NOT_SUPPORTED;
  return FLAGFFT_SUCCESS;
}

PlanDesc plan_desc(const Desc &d) {
  PlanDesc p;
  p.rank = d.rank;
  p.n0 = d.n[0];
  p.n1 = d.rank >= 2 ? d.n[1] : 0;
  p.n2 = d.rank == 3 ? d.n[2] : 0;
  p.type = static_cast<int>(d.type);
  p.batch = d.batch;
  p.prec = d.prec;
  p.real_fwd = d.real_fwd;
  p.real_inv = d.real_inv;
  p.compute = d.compute;
  return p;
}

std::shared_ptr<PlanImpl> checked(flagfftHandle h) {
  if (h == nullptr) return nullptr;
  std::lock_guard<std::mutex> lk(g_handles_mu);
  auto it = g_handles.find(h);
  return it == g_handles.end() ? nullptr : it->second;
}

flagfftResult build_plan(flagfftHandle *out, const Desc &d) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    (void)hipGetLastError();
    return FLAGFFT_INVALID_DEVICE;
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    return FLAGFFT_INVALID_DEVICE;
  }
  try {
    auto impl = std::make_unique<PlanImpl>();
    impl->desc = d;
    impl->device = dev;
    std::string tuned;
    const PlanOptions opts = plan_options(d, tuned);
    impl->tree = plan_tree(plan_desc(d), opts);
    const Strides *st = d.general ? &d.strides : nullptr;
    if (d.real_fwd) {
      impl->fwd = build_executable(impl->tree, false, opts, d.idist, d.odist, true, st);
    } else if (d.real_inv) {
      impl->inv = build_executable(impl->tree, true, opts, d.idist, d.odist, true, st);
    } else {
      impl->fwd = build_executable(impl->tree, false, opts, d.idist, d.odist, true, st);
      impl->inv = build_executable(impl->tree, true, opts, d.idist, d.odist, true, st);
    }
    PlanOptions ip = opts;
    ip.in_place = true;
    if (impl->fwd && impl->fwd->inplace_hazard)
      impl->fwd_ip = build_executable(impl->tree, false, ip, d.idist, d.odist, true, st);
    if (impl->inv && impl->inv->inplace_hazard)
      impl->inv_ip = build_executable(impl->tree, true, ip, d.idist, d.odist, true, st);
    std::ostringstream o;
    o << impl->tree.describe();
    if (d.general)
      o << "layout: istride=" << d.istride << " idist=" << d.idist << " input row=" << d.strides.irow
        << " ostride=" << d.ostride << " odist=" << d.odist << " output row=" << d.strides.orow << "\n";
    if (!tuned.empty()) o << "tuned-plan: " << tuned << "\n";
    o << "\n-- Forward execution --\n" << (impl->fwd ? impl->fwd->describe() : std::string("(not compiled)\n"));
    if (impl->inv) o << "\n-- Inverse execution --\n" << impl->inv->describe();
    if (impl->fwd_ip) o << "\n-- Forward execution, in place --\n" << impl->fwd_ip->describe();
    if (impl->inv_ip) o << "\n-- Inverse execution, in place --\n" << impl->inv_ip->describe();
    impl->description = o.str();
    std::shared_ptr<PlanImpl> shared(std::move(impl));
    auto handle = std::make_unique<flagfftPlan_t>();
    handle->impl = shared.get();
    flagfftHandle h = handle.release();
    {
      std::lock_guard<std::mutex> lk(g_handles_mu);
      g_handles.emplace(h, std::move(shared));
    }
    *out = h;
    return FLAGFFT_SUCCESS;
  } catch (const std::bad_alloc &) {
    return FLAGFFT_ALLOC_FAILED;
  } catch (const HipError &e) {
    std::fprintf(stderr, "[flagfft] plan setup failed: %s\n", e.what());
    return e.code == hipErrorOutOfMemory ? FLAGFFT_ALLOC_FAILED : FLAGFFT_SETUP_FAILED;
  } catch (const std::exception &e) {
    std::fprintf(stderr, "[flagfft] plan setup failed: %s\n", e.what());
    return FLAGFFT_SETUP_FAILED;
  }
}

// The executable one Exec of `direction` runs (real plans have one), with
// the in-place variant when the buffers are one and the plan has a hazard.
Executable *select_exec(PlanImpl &p, const void *in, const void *out, int direction) {
  Executable *ex = nullptr;
  if (p.desc.real_fwd) ex = p.fwd.get();
  else if (p.desc.real_inv) ex = p.inv.get();
  else ex = direction == FLAGFFT_INVERSE ? p.inv.get() : p.fwd.get();
  if (ex != nullptr && in == out && ex->inplace_hazard) ex = ex == p.fwd.get() ? p.fwd_ip.get() : p.inv_ip.get();
  return ex;
}

flagfftResult exec_common(flagfftHandle handle, const void *in, void *out, flagfftType expect, int direction,
                          bool has_direction) {
  std::shared_ptr<PlanImpl> p = checked(handle);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  if (in == nullptr || out == nullptr) return FLAGFFT_INVALID_VALUE;
  if (p->desc.type != expect) return FLAGFFT_INVALID_TYPE;
  if (has_direction && direction != FLAGFFT_FORWARD && direction != FLAGFFT_INVERSE) return FLAGFFT_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(p->mu);
  Executable *ex = select_exec(*p, in, out, direction);
  if (ex == nullptr) return FLAGFFT_INTERNAL_ERROR;
  int cur = p->device;
  if (hipGetDevice(&cur) == hipSuccess && cur != p->device) (void)hipSetDevice(p->device);
  hipError_t e = (p->graph_mode && !ex->profiling) ? ex->run_graph(in, out, p->stream) : ex->run(in, out, p->stream);
  if (cur != p->device) (void)hipSetDevice(cur);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[flagfft] exec failed: %s\n", hipGetErrorString(e));
    return FLAGFFT_EXEC_FAILED;
  }
  return FLAGFFT_SUCCESS;
}

}  // namespace
}  // namespace flagfft

using flagfft::exec_common;

extern "C" flagfftResult flagfftPlanMany(flagfftHandle *plan, int rank, int *n, int *inembed, int istride, int idist,
                                         int *onembed, int ostride, int odist, flagfftType type, int batch) {
  if (plan == nullptr) return FLAGFFT_INVALID_VALUE;
  *plan = nullptr;
  flagfft::Desc d;
  flagfftResult r = flagfft::make_desc(rank, n, inembed, istride, idist, onembed, ostride, odist, type, batch, d);
  if (r != FLAGFFT_SUCCESS) return r;
  return flagfft::build_plan(plan, d);
}

extern "C" flagfftResult flagfftPlan1d(flagfftHandle *plan, int nx, flagfftType type, int batch) {
  int n[1] = {nx};
  const bool real_fwd = type == FLAGFFT_R2C || type == FLAGFFT_D2Z;
  const bool real_inv = type == FLAGFFT_C2R || type == FLAGFFT_Z2D;
  const int half = nx / 2 + 1;
  return flagfftPlanMany(plan, 1, n, nullptr, 1, real_inv ? half : nx, nullptr, 1, real_fwd ? half : nx, type, batch);
}

extern "C" flagfftResult flagfftxPlan1dCompute(flagfftHandle *plan, int nx, flagfftType type, int batch,
                                              int compute) {
  if (plan == nullptr) return FLAGFFT_INVALID_VALUE;
  *plan = nullptr;
  if (compute < FLAGFFTX_COMPUTE_VECTOR || compute > FLAGFFTX_COMPUTE_FP16X3) return FLAGFFT_INVALID_VALUE;
  if (compute == FLAGFFTX_COMPUTE_VECTOR) return flagfftPlan1d(plan, nx, type, batch);
  int n[1] = {nx};
  flagfft::Desc d;
  flagfftResult r = flagfft::make_desc(1, n, nullptr, 1, nx, nullptr, 1, nx, type, batch, d);
  if (r != FLAGFFT_SUCCESS) return r;
  if (type != FLAGFFT_C2C || nx != 1024) return FLAGFFT_NOT_SUPPORTED;
  d.compute = compute;
  return flagfft::build_plan(plan, d);
}

// flagfftPlanMany with the advanced layouts the reference answers
// FLAGFFT_NOT_SUPPORTED for (src/exec/c_api_internal.cpp:60-120): any
// positive element stride, embeds and distances for C2C / Z2Z of rank 1 and 2.
extern "C" flagfftResult flagfftxPlanMany(flagfftHandle *plan, int rank, int *n, int *inembed, int istride, int idist,
                                         int *onembed, int ostride, int odist, flagfftType type, int batch) {
  if (plan == nullptr) return FLAGFFT_INVALID_VALUE;
  *plan = nullptr;
  flagfft::Desc d;
  flagfftResult r = flagfft::make_desc(rank, n, inembed, istride, idist, onembed, ostride, odist, type, batch, d);
  if (r == FLAGFFT_SUCCESS) return flagfft::build_plan(plan, d);
  if (r != FLAGFFT_NOT_SUPPORTED) return r;
  if (rank == 2 && (d.real_fwd || d.real_inv)) {
    // real 2D with padded rows: unit strides, batches n0 rows apart
    const int64_t n0 = d.n[0], n1 = d.n[1], h1 = n1 / 2 + 1;
    const int64_t iemb = d.inembed[1], oemb = d.onembed[1];
    const int64_t imin = d.real_fwd ? n1 : h1, omin = d.real_fwd ? h1 : n1;
    if (istride != 1 || ostride != 1) return FLAGFFT_NOT_SUPPORTED;
    if (iemb < imin || oemb < omin || idist != n0 * iemb || odist != n0 * oemb) return FLAGFFT_NOT_SUPPORTED;
    // pairs of reals are read / written as complex values: even real pitch
    if (n1 % 2 == 0 && (d.real_fwd ? iemb : oemb) % 2 != 0) return FLAGFFT_NOT_SUPPORTED;
    d.general = true;
    d.strides.irow = iemb;
    d.strides.orow = oemb;
    return flagfft::build_plan(plan, d);
  }
  if ((type != FLAGFFT_C2C && type != FLAGFFT_Z2Z) || rank > 2) return FLAGFFT_NOT_SUPPORTED;
  if (istride <= 0 || ostride <= 0 || idist <= 0 || odist <= 0) return FLAGFFT_INVALID_VALUE;
  const int64_t last = d.n[static_cast<size_t>(rank - 1)];
  const int64_t iemb = rank == 2 ? d.inembed[1] : last, oemb = rank == 2 ? d.onembed[1] : last;
  if (iemb < last || oemb < last) return FLAGFFT_INVALID_VALUE;
  d.general = true;
  d.strides.istride = istride;
  d.strides.ostride = ostride;
  d.strides.irow = iemb * istride;
  d.strides.orow = oemb * ostride;
  return flagfft::build_plan(plan, d);
}

// Rank 3, contiguous batches (flagfftPlan3d / rank-3 PlanMany keep the
// reference's FLAGFFT_NOT_SUPPORTED, src/exec/c_api_plan.cpp:274-336).
extern "C" flagfftResult flagfftxPlan3d(flagfftHandle *plan, int nx, int ny, int nz, flagfftType type, int batch) {
  if (plan == nullptr) return FLAGFFT_INVALID_VALUE;
  *plan = nullptr;
  if (batch <= 0) return FLAGFFT_INVALID_VALUE;
  if (nx <= 0 || ny <= 0 || nz <= 0) return FLAGFFT_INVALID_SIZE;
  flagfft::Desc d;
  flagfftResult r = flagfft::type_metadata(type, d);
  if (r != FLAGFFT_SUCCESS) return r;
  const int64_t h = nz / 2 + 1;
  const int64_t full = static_cast<int64_t>(nx) * ny * nz, half = static_cast<int64_t>(nx) * ny * h;
  if (static_cast<int64_t>(batch) * full * (d.real_fwd || d.real_inv ? 1 : 2) >= (int64_t{1} << 40))
    return FLAGFFT_INVALID_SIZE;
  d.rank = 3;
  d.n = {nx, ny, nz};
  d.type = type;
  d.batch = batch;
  d.istride = d.ostride = 1;
  d.idist = d.real_inv ? half : full;
  d.odist = d.real_fwd ? half : full;
  d.inembed = {nx, ny, d.real_inv ? static_cast<int>(h) : nz};
  d.onembed = {nx, ny, d.real_fwd ? static_cast<int>(h) : nz};
  return flagfft::build_plan(plan, d);
}

extern "C" flagfftResult flagfftPlan2d(flagfftHandle *plan, int nx, int ny, flagfftType type) {
  int n[2] = {nx, ny};
  const bool real_fwd = type == FLAGFFT_R2C || type == FLAGFFT_D2Z;
  const bool real_inv = type == FLAGFFT_C2R || type == FLAGFFT_Z2D;
  const int half = ny / 2 + 1;
  return flagfftPlanMany(plan, 2, n, nullptr, 1, real_inv ? nx * half : nx * ny, nullptr, 1,
                         real_fwd ? nx * half : nx * ny, type, 1);
}

extern "C" flagfftResult flagfftPlan3d(flagfftHandle *plan, int nx, int ny, int nz, flagfftType type) {
  int n[3] = {nx, ny, nz};
  return flagfftPlanMany(plan, 3, n, nullptr, 1, nx * ny * nz, nullptr, 1, nx * ny * nz, type, 1);
}

extern "C" flagfftResult flagfftExecC2C(flagfftHandle plan, flagfftComplex *idata, flagfftComplex *odata,
                                        int direction) {
  return exec_common(plan, idata, odata, FLAGFFT_C2C, direction, true);
}

extern "C" flagfftResult flagfftExecZ2Z(flagfftHandle plan, flagfftDoubleComplex *idata, flagfftDoubleComplex *odata,
                                        int direction) {
  return exec_common(plan, idata, odata, FLAGFFT_Z2Z, direction, true);
}

extern "C" flagfftResult flagfftExecR2C(flagfftHandle plan, flagfftReal *idata, flagfftComplex *odata) {
  return exec_common(plan, idata, odata, FLAGFFT_R2C, FLAGFFT_FORWARD, false);
}

extern "C" flagfftResult flagfftExecD2Z(flagfftHandle plan, flagfftDoubleReal *idata, flagfftDoubleComplex *odata) {
  return exec_common(plan, idata, odata, FLAGFFT_D2Z, FLAGFFT_FORWARD, false);
}

extern "C" flagfftResult flagfftExecC2R(flagfftHandle plan, flagfftComplex *idata, flagfftReal *odata) {
  return exec_common(plan, idata, odata, FLAGFFT_C2R, FLAGFFT_INVERSE, false);
}

extern "C" flagfftResult flagfftExecZ2D(flagfftHandle plan, flagfftDoubleComplex *idata, flagfftDoubleReal *odata) {
  return exec_common(plan, idata, odata, FLAGFFT_Z2D, FLAGFFT_INVERSE, false);
}

// ---- allocation-free execution ----------------------------------------------
//
// Per-device workspace of flagfftxExecC2CNoAlloc: a fixed array in the code
// object, so the call needs no hipMalloc.  Every device that loads the
// library holds one copy until the process ends.  In host code `scratch`
// names a host-side stand-in, never the device memory: kernels get the
// address from hipGetSymbolAddress (per device, kept in g_scratch_ptr).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace flagfft {
namespace {

std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_ptr;  // device ordinal -> address of that device's `scratch`

// The device address of `scratch` on `device` (current when called);
// looked up once per device.
hipError_t scratch_on(int device, void **ptr) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) {
    *ptr = it->second;
    return hipSuccess;
  }
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || p == nullptr) {
    (void)hipGetLastError();
    return e != hipSuccess ? e : hipErrorInvalidSymbol;
  }
  g_scratch_ptr[device] = p;
  *ptr = p;
  return hipSuccess;
}

}  // namespace
}  // namespace flagfft

extern "C" flagfftResult flagfftxExecC2CNoAlloc(flagfftHandle plan, flagfftComplex *idata, flagfftComplex *odata,
                                                int direction) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(plan);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  if (idata == nullptr || odata == nullptr) return FLAGFFT_INVALID_VALUE;
  if (p->desc.type != FLAGFFT_C2C) return FLAGFFT_INVALID_TYPE;
  if (direction != FLAGFFT_FORWARD && direction != FLAGFFT_INVERSE) return FLAGFFT_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(p->mu);
  flagfft::Executable *ex = flagfft::select_exec(*p, idata, odata, direction);
  if (ex == nullptr) return FLAGFFT_INTERNAL_ERROR;
  // the plan's workspaces, each at a 256-byte aligned offset of the scratch array
  constexpr size_t kMaxWs = 8;
  size_t offset[kMaxWs] = {};
  size_t need = 0;
  if (ex->ws_bytes.size() > kMaxWs) return FLAGFFT_INVALID_SIZE;
  for (size_t i = 0; i < ex->ws_bytes.size(); ++i) {
    offset[i] = need;
    need += (std::max<size_t>(ex->ws_bytes[i], 256) + 255) / 256 * 256;
  }
  if (need > kNoAllocScratchBytes) {
    std::fprintf(stderr, "[flagfft] flagfftxExecC2CNoAlloc: the plan needs %zu bytes of workspace, above the %zu-byte "
                 "scratch array\n", need, kNoAllocScratchBytes);
    return FLAGFFT_INVALID_SIZE;
  }
  int cur = p->device;
  if (hipGetDevice(&cur) == hipSuccess && cur != p->device) (void)hipSetDevice(p->device);
  void *base = nullptr;
  hipError_t e = flagfft::scratch_on(p->device, &base);
  if (e == hipSuccess) {
    void *wsp[kMaxWs] = {};
    for (size_t i = 0; i < ex->ws_bytes.size(); ++i) wsp[i] = static_cast<unsigned char *>(base) + offset[i];
    e = ex->run_batch(idata, odata, wsp, ex->built_batch, p->stream);
  }
  if (cur != p->device) (void)hipSetDevice(cur);
  if (e != hipSuccess) {
    std::fprintf(stderr, "[flagfft] exec failed: %s\n", hipGetErrorString(e));
    return FLAGFFT_EXEC_FAILED;
  }
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftSetStream(flagfftHandle handle, flagfftStream_t stream) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  std::lock_guard<std::mutex> lk(p->mu);
  p->stream = static_cast<hipStream_t>(stream);
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftDestroy(flagfftHandle handle) {
  if (handle == nullptr) return FLAGFFT_INVALID_PLAN;
  std::shared_ptr<flagfft::PlanImpl> p;
  {
    std::lock_guard<std::mutex> lk(flagfft::g_handles_mu);
    auto it = flagfft::g_handles.find(handle);
    if (it == flagfft::g_handles.end()) return FLAGFFT_INVALID_PLAN;
    p = std::move(it->second);
    flagfft::g_handles.erase(it);
  }
  // The handle shell stays allocated (a tombstone with a null impl): its
  // address is never handed out again, so a stale handle cannot alias a plan
  // created later and keeps returning FLAGFFT_INVALID_PLAN.
  handle->impl = nullptr;
  // the plan (workspaces, tables, graphs) goes with the last reference: here,
  // or at the end of a call that looked the handle up before the erase
  p.reset();
  return FLAGFFT_SUCCESS;
}

extern "C" const char *flagfftGetPlanDescription(flagfftHandle handle) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return nullptr;
  std::lock_guard<std::mutex> lk(p->mu);
  return p->description.c_str();
}

// ---- extensions -----------------------------------------------------------

extern "C" const char *flagfftxVersion(void) { return "flagfft-gfx950 0.1"; }

extern "C" flagfftResult flagfftxPlanDescribe(int rank, const int *n, flagfftType type, int batch, char *buf,
                                              size_t len) {
  if (buf == nullptr || len == 0) return FLAGFFT_INVALID_VALUE;
  buf[0] = '\0';
  if (rank <= 0 || n == nullptr || batch <= 0) return FLAGFFT_INVALID_VALUE;
  for (int i = 0; i < rank; ++i)
    if (n[i] <= 0) return FLAGFFT_INVALID_SIZE;
  const bool real_fwd = type == FLAGFFT_R2C || type == FLAGFFT_D2Z;
  const bool real_inv = type == FLAGFFT_C2R || type == FLAGFFT_Z2D;
  int64_t in_len = 1, out_len = 1;
  for (int i = 0; i < rank; ++i) {
    const int64_t inner = (i == rank - 1) ? (n[i] / 2 + 1) : n[i];
    in_len *= real_inv ? inner : n[i];
    out_len *= real_fwd ? inner : n[i];
  }
  if (in_len > INT32_MAX || out_len > INT32_MAX) return FLAGFFT_INVALID_SIZE;
  flagfft::Desc d;
  flagfftResult r = flagfft::make_desc(rank, n, nullptr, 1, static_cast<int>(in_len), nullptr, 1,
                                       static_cast<int>(out_len), type, rank == 2 ? 1 : batch, d);
  if (r != FLAGFFT_SUCCESS) return r;
  try {
    std::string tuned;
    const flagfft::PlanOptions opts = flagfft::plan_options(d, tuned);
    flagfft::PlanTree t = flagfft::plan_tree(flagfft::plan_desc(d), opts);
    std::string s = t.describe();
    if (!tuned.empty()) s += "tuned-plan: " + tuned + "\n";
    std::strncpy(buf, s.c_str(), len - 1);
    buf[len - 1] = '\0';
    return FLAGFFT_SUCCESS;
  } catch (const std::exception &) {
    return FLAGFFT_SETUP_FAILED;
  }
}

extern "C" flagfftResult flagfftxPlanStats(flagfftHandle handle, int direction, int *launches, long long *pass_bytes) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  if (launches == nullptr || pass_bytes == nullptr) return FLAGFFT_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(p->mu);
  const flagfft::Executable *ex = nullptr;
  if (p->desc.real_fwd) ex = p->fwd.get();
  else if (p->desc.real_inv) ex = p->inv.get();
  else ex = direction == FLAGFFT_INVERSE ? p->inv.get() : p->fwd.get();
  if (ex == nullptr) return FLAGFFT_INVALID_VALUE;
  *launches = static_cast<int>(ex->steps.size());
  long long total = 0;
  for (const auto &s : ex->steps) total += s.bytes_moved;
  *pass_bytes = total;
  return FLAGFFT_SUCCESS;
}

namespace flagfft {
namespace {
Executable *exec_for(PlanImpl *p, int direction) {
  if (p->desc.real_fwd) return p->fwd.get();
  if (p->desc.real_inv) return p->inv.get();
  return direction == FLAGFFT_INVERSE ? p->inv.get() : p->fwd.get();
}
}  // namespace
}  // namespace flagfft

extern "C" flagfftResult flagfftxSetProfiling(flagfftHandle handle, int enable) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  std::lock_guard<std::mutex> lk(p->mu);
  if (p->fwd) p->fwd->set_profiling(enable != 0);
  if (p->inv) p->inv->set_profiling(enable != 0);
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftxGetProfile(flagfftHandle handle, int direction, int capacity, double *ms_total,
                                           long long *calls, long long *bytes, int *nsteps) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  if (nsteps == nullptr || capacity < 0) return FLAGFFT_INVALID_VALUE;
  std::lock_guard<std::mutex> lk(p->mu);
  flagfft::Executable *ex = flagfft::exec_for(p.get(), direction);
  if (ex == nullptr) return FLAGFFT_INVALID_VALUE;
  if (ex->harvest_profile() != hipSuccess) return FLAGFFT_EXEC_FAILED;
  *nsteps = static_cast<int>(ex->steps.size());
  for (int i = 0; i < capacity && i < *nsteps; ++i) {
    if (ms_total) ms_total[i] = ex->prof_ms[static_cast<size_t>(i)];
    if (calls) calls[i] = ex->prof_calls[static_cast<size_t>(i)];
    if (bytes) bytes[i] = ex->steps[static_cast<size_t>(i)].bytes_moved;
  }
  return FLAGFFT_SUCCESS;
}

extern "C" const char *flagfftxStepLabel(flagfftHandle handle, int direction, int step) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return nullptr;
  std::lock_guard<std::mutex> lk(p->mu);
  flagfft::Executable *ex = flagfft::exec_for(p.get(), direction);
  if (ex == nullptr || step < 0 || step >= static_cast<int>(ex->steps.size())) return nullptr;
  return ex->steps[static_cast<size_t>(step)].label.c_str();
}

extern "C" flagfftResult flagfftxSetGraphMode(flagfftHandle handle, int enable) {
  std::shared_ptr<flagfft::PlanImpl> p = flagfft::checked(handle);
  if (p == nullptr) return FLAGFFT_INVALID_PLAN;
  std::lock_guard<std::mutex> lk(p->mu);
  p->graph_mode = enable != 0;
  return FLAGFFT_SUCCESS;
}

namespace flagfft {
namespace {
bool tuning_from_c(const flagfftxTuning *t, TuneOverrides &o) {
  if (t->npasses < 0 || t->npasses > 4 || t->nradices < 0 || t->nradices > 16) return false;
  o.leaf_max = t->leaf_max;
  o.pass_max = t->pass_max;
  o.units = t->units;
  o.ept = t->ept;
  o.chunk_bytes = t->chunk_bytes;
  o.stage_form = t->stage_form;
  o.variant = t->variant;
  o.passes.assign(t->passes, t->passes + t->npasses);
  o.radices.assign(t->radices, t->radices + t->nradices);
  return true;
}
}  // namespace
}  // namespace flagfft

extern "C" flagfftResult flagfftxSetTuning(const flagfftxTuning *t) {
  flagfft::PlanOptions o;
  if (t != nullptr) {
    flagfft::TuneOverrides ov;
    if (!flagfft::tuning_from_c(t, ov) || !flagfft::overrides_to_options(ov, o)) return FLAGFFT_INVALID_VALUE;
  }
  flagfft::t_options = o;
  flagfft::t_explicit = t != nullptr;
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftxTunedStore(int nx, flagfftType type, int batch, const flagfftxTuning *t,
                                            double median_us, double gbps, double default_us) {
  if (t == nullptr || batch <= 0) return FLAGFFT_INVALID_VALUE;
  if (nx <= 0) return FLAGFFT_INVALID_SIZE;
  flagfft::Desc d;
  if (flagfft::type_metadata(type, d) != FLAGFFT_SUCCESS) return FLAGFFT_INVALID_TYPE;
  flagfft::TuneOverrides ov;
  flagfft::PlanOptions o;
  if (!flagfft::tuning_from_c(t, ov) || !flagfft::overrides_to_options(ov, o)) return FLAGFFT_INVALID_VALUE;
  std::string err;
  if (!flagfft::tuned_store(flagfft::tune_key(type, nx, batch), flagfft::tuning_to_json(ov), median_us, gbps,
                            default_us, err)) {
    std::fprintf(stderr, "[flagfft] tuned-plan store failed: %s\n", err.c_str());
    return FLAGFFT_SETUP_FAILED;
  }
  return FLAGFFT_SUCCESS;
}

extern "C" flagfftResult flagfftxTunedDbPath(char *buf, size_t len) {
  if (buf == nullptr || len == 0) return FLAGFFT_INVALID_VALUE;
  const std::string p = flagfft::tuned_db_path();
  std::strncpy(buf, p.c_str(), len - 1);
  buf[len - 1] = '\0';
  return FLAGFFT_SUCCESS;
}

