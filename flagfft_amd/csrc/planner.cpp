// planner.cpp — decomposition choice (host only).
#include "planner.hpp"

#include "kernels.hpp"

#include <algorithm>
#include <sstream>
#include <stdexcept>

namespace flagfft {

bool is_prime(int64_t n) {
  if (n < 2) return false;
  if (n % 2 == 0) return n == 2;
  for (int64_t d = 3; d * d <= n; d += 2)
    if (n % d == 0) return false;
  return true;
}

namespace {

int64_t modpow(int64_t b, int64_t e, int64_t m) {
  __int128 r = 1, x = b % m;
  while (e > 0) {
    if (e & 1) r = (r * x) % m;
    x = (x * x) % m;
    e >>= 1;
  }
  return static_cast<int64_t>(r);
}

std::vector<int64_t> prime_factors(int64_t n) {
  std::vector<int64_t> f;
  for (int64_t d = 2; d * d <= n; ++d)
    while (n % d == 0) {
      f.push_back(d);
      n /= d;
    }
  if (n > 1) f.push_back(n);
  return f;
}

std::string join(const std::vector<int> &v, const char *sep) {
  std::ostringstream o;
  for (size_t i = 0; i < v.size(); ++i) o << (i ? sep : "") << v[i];
  return o.str();
}

std::string join64(const std::vector<int64_t> &v, const char *sep) {
  std::ostringstream o;
  for (size_t i = 0; i < v.size(); ++i) o << (i ? sep : "") << v[i];
  return o.str();
}

}  // namespace

// Smallest primitive root, as src/plan/rader_utils.cpp:55-77 chooses it.
int64_t primitive_root(int64_t p) {
  if (!is_prime(p)) return -1;
  if (p == 2) return 1;
  std::vector<int64_t> f = prime_factors(p - 1);
  f.erase(std::unique(f.begin(), f.end()), f.end());
  for (int64_t g = 2; g < p; ++g) {
    bool ok = true;
    for (int64_t q : f)
      if (modpow(g, (p - 1) / q, p) == 1) {
        ok = false;
        break;
      }
    if (ok) return g;
  }
  return -1;
}

bool smooth13(int64_t n) {
  if (n < 1) return false;
  for (int64_t p : {2, 3, 5, 7, 11, 13})
    while (n % p == 0) n /= p;
  return n == 1;
}

// Lengths whose prime factors all have a radix codelet: 2, 3, 5, 7, 11, 13,
// 17, 19 (the reference's supported radices, src/utils/include/flagfft/base.hpp:47).
bool smooth19(int64_t n) {
  if (n < 1) return false;
  for (int64_t p : {2, 3, 5, 7, 11, 13, 17, 19})
    while (n % p == 0) n /= p;
  return n == 1;
}

// First-fit-decreasing packing of the prime factors into radices <= 16;
// the primes 17 and 19 are radices of their own.
bool radix_factor(int64_t n, std::vector<int> &out) {
  out.clear();
  if (n < 1) return false;
  if (n == 1) return true;
  std::vector<int64_t> f = prime_factors(n);
  for (int64_t p : f)
    if (p > 19) return false;
  std::sort(f.begin(), f.end(), std::greater<int64_t>());
  std::vector<int> bins;
  for (int64_t p : f) {
    bool placed = false;
    for (int &b : bins)
      if (b * p <= 16) {
        b *= static_cast<int>(p);
        placed = true;
        break;
      }
    if (!placed) bins.push_back(static_cast<int>(p));
  }
  std::sort(bins.begin(), bins.end(), std::greater<int>());
  out = bins;
  return true;
}

bool passes_specialised(const std::vector<int64_t> &passes, bool fast_only) {
  std::vector<int> rad;
  for (int64_t r : passes) {
    if (!radix_factor(r, rad)) return false;
    const int id = stage_kernel_id(rad);
    if (id < 0 || (fast_only && !stage_kernel_fast(id))) return false;
  }
  return true;
}

int64_t leaf_max(const PlanOptions &o, Prec p) { return p == Prec::F32 ? o.leaf_max_f32 : o.leaf_max_f64; }
int64_t pass_max(const PlanOptions &o, Prec p) { return p == Prec::F32 ? o.pass_max_f32 : o.pass_max_f64; }

namespace {

// All splits of n into k factors (non-increasing), each <= cap and radix-factorable.
void enum_splits(int64_t n, int k, int64_t cap, int64_t hi, std::vector<int64_t> &cur,
                 std::vector<std::vector<int64_t>> &out) {
  std::vector<int> tmp;
  if (k == 1) {
    if (n <= cap && n <= hi && n >= 2 && radix_factor(n, tmp)) {
      cur.push_back(n);
      out.push_back(cur);
      cur.pop_back();
    }
    return;
  }
  for (int64_t d = std::min(cap, hi); d >= 2; --d) {
    if (n % d != 0 || !radix_factor(d, tmp)) continue;
    cur.push_back(d);
    enum_splits(n / d, k - 1, cap, d, cur, out);
    cur.pop_back();
  }
}

// Split n into k passes.  Prefer splits whose every pass has a specialised
// kernel, then the one with the largest smallest factor (fewest HBM round
// trips per element are equal for a fixed k; balance keeps every pass's
// workgroups full).
bool split_passes(int64_t n, int k, int64_t cap, std::vector<int64_t> &best) {
  std::vector<std::vector<int64_t>> all;
  std::vector<int64_t> cur;
  enum_splits(n, k, cap, cap, cur, all);
  if (all.empty()) return false;
  auto score = [](const std::vector<int64_t> &v) {
    return std::make_pair(passes_specialised(v) ? 1 : 0, *std::min_element(v.begin(), v.end()));
  };
  const std::vector<int64_t> *b = &all[0];
  for (const auto &v : all)
    if (score(v) > score(*b)) b = &v;
  best = *b;
  return true;
}

// Row-segment bytes of the widest specialised column super-stage for a radix
// sequence (adjacent columns per workgroup x element bytes); 0 if none.
int column_segment_bytes(const std::vector<int> &rad, Prec p) {
  for (int units : {64, 32, 16, 8, 4, 2, 1})
    if (stage_kernel_id(rad, p == Prec::F64, units) >= 0) return units * elem_bytes(p);
  return 0;
}

// Relative time of one HBM sweep by a pass of super-radix r: the inverse of
// its measured rate on MI355X (TB/s of read + write bytes).  Register-resident
// two-stage passes run at 3.8 (first pass, from uncached input) to 4.6 with
// row segments of 128 bytes or more, 3.0 to 4.2 with 64;
// LDS super-stages at 2.2 to 3.5 by the row-segment bytes their tile reads;
// the run-time-radix pass at about 1.1.
// tcnt: units per sequence (n / r); a register-resident instance takes the
// pass only when its units per workgroup divide it (one tile, one sequence).
double pass_time(int64_t r, Prec p, bool first, const PlanOptions &o, int64_t tcnt) {
  std::vector<int> rad;
  if (o.generic_kernels) {
    // persistent-queue plans: register-resident queue instances, else the run-time-radix pass
    return (o.queue_form && col_queue_kernel_id(r, p == Prec::F64) >= 0) ? 1.0 / 4.0 : 1.0 / 1.1;
  }
  if (o.stage_form != 2 && col_kernel_id(r, p == Prec::F64, 0) >= 0 &&
      col_kernel_id_dividing(r, p == Prec::F64, tcnt) >= 0) {
    // 64-byte segments (the 4096-point instance: 8 units of 32 KiB in one
    // workgroup) run at 3.0 / 4.2
    if (col_run_bytes(r, p == Prec::F64, first) < 128) return first ? 1.0 / 3.0 : 1.0 / 4.2;
    return first ? 1.0 / 3.8 : 1.0 / 4.6;
  }
  if (radix_factor(r, rad)) {
    const int seg = column_segment_bytes(rad, p);
    if (seg >= 128) return 1.0 / 3.5;
    if (seg >= 64) return 1.0 / 3.0;
    if (seg > 0) return 1.0 / 2.2;
  }
  return 1.0 / 1.1;
}

// The split of n into 2-4 passes with the least total pass time; near ties
// (2 %) go to fewer passes, then to the larger smallest factor.
bool cheapest_split(int64_t n, Prec p, const PlanOptions &o, std::vector<int64_t> &best) {
  const int64_t cap = pass_max(o, p);
  double best_t = 0.0;
  bool have = false;
  for (int k = 2; k <= 4; ++k) {
    std::vector<std::vector<int64_t>> all;
    std::vector<int64_t> cur;
    enum_splits(n, k, cap, cap, cur, all);
    for (const auto &v : all) {
      double tt = 0.0;
      for (size_t i = 0; i < v.size(); ++i) tt += pass_time(v[i], p, i == 0, o, n / v[i]);
      const bool better = !have || tt < best_t * 0.98 ||
                          (tt <= best_t * 1.02 && v.size() == best.size() &&
                           *std::min_element(v.begin(), v.end()) > *std::min_element(best.begin(), best.end()));
      if (better) {
        best = v;
        best_t = tt;
        have = true;
      }
    }
  }
  return have;
}

bool multipass_core(int64_t n, Prec p, const PlanOptions &o, Core &c, bool by_cost = false) {
  std::vector<int64_t> passes;
  const int64_t cap = pass_max(o, p);
  bool ok = false;
  if (!o.forced_passes.empty()) {
    int64_t prod = 1;
    bool valid = true;
    std::vector<int> tmp;
    for (int64_t r : o.forced_passes) {
      prod *= r;
      valid = valid && r >= 2 && r <= leaf_max(o, p) && radix_factor(r, tmp);
    }
    if (valid && prod == n) {
      passes = o.forced_passes;
      ok = true;
    }
  }
  if (!ok && by_cost) ok = cheapest_split(n, p, o, passes);
  for (int k = 2; !ok && k <= 4; ++k) ok = split_passes(n, k, cap, passes);
  if (!ok) return false;
  c.kind = CoreKind::MultiPass;
  c.n = n;
  c.passes = passes;
  c.pass_radices.clear();
  for (int64_t r : passes) {
    std::vector<int> rad;
    radix_factor(r, rad);
    c.pass_radices.push_back(rad);
  }
  return true;
}

int rader_lds_bytes(int64_t p, Prec prec) {
  const int64_t m = p - 1;
  return static_cast<int>((lds_pad(static_cast<int>(p)) + 1 + lds_pad(static_cast<int>(m)) + 1 + 1) *
                          elem_bytes(prec));
}

}  // namespace

Core plan_core(int64_t n, Prec p, const PlanOptions &o) {
  if (n < 1) throw std::invalid_argument("transform length must be positive");
  Core c;
  c.n = n;
  std::vector<int> rad;
  const int64_t lmax = leaf_max(o, p);
  // an explicit super-radix split of exactly n overrides the single pass
  int64_t forced_prod = o.forced_passes.size() >= 2 ? 1 : 0;
  for (int64_t r : o.forced_passes) forced_prod *= r;
  bool leaf = n <= lmax && radix_factor(n, rad) && forced_prod != n;
  if (!leaf && !o.generic_kernels && o.forced_passes.empty() && n <= 32768 && radix_factor(n, rad) &&
      leaf_kernel_id(rad) >= 0 &&
      static_cast<int64_t>(lds_pad(static_cast<int>(n) - 1) + 2) * elem_bytes(p) <= o.lds_cap_bytes) {
    // a specialised instance holding the whole sequence in LDS: one HBM round
    // trip instead of two (one workgroup per sequence)
    leaf = true;
  }
  if (leaf) {
    c.kind = CoreKind::Leaf;
    if (!o.forced_radices.empty()) {
      int64_t prod = 1;
      for (int r : o.forced_radices) prod *= r;
      bool valid = prod == n;
      for (int r : o.forced_radices) valid = valid && r >= 2 && (r <= 17 || r == 19);
      if (valid) rad = o.forced_radices;
    }
    c.radices = rad;
    return c;
  }
  if (is_prime(n) && n > 16) {
    const int64_t m = n - 1;
    if (m <= lmax && radix_factor(m, rad) && rader_lds_bytes(n, p) <= o.lds_cap_bytes) {
      c.kind = CoreKind::Rader;
      c.m = m;
      c.root = primitive_root(n);
      c.radices = rad;
      return c;
    }
  }
  if (smooth19(n) && multipass_core(n, p, o, c, true)) return c;
  // Bluestein over a 13-smooth convolution length M >= 2n-1.  A length that
  // fits one LDS pass is taken as soon as it is found (the whole convolution
  // is one launch).  Otherwise every M up to twice the minimum is weighed by
  // its HBM sweeps: M x passes, x 2.5 when a pass has no fast specialised
  // kernel (the run-time-radix pass and the 5- / 10-smooth instances run
  // about that much slower than the power-of-two passes).
  Core best;
  double best_cost = 0.0;
  bool have = false;
  const int64_t m0 = 2 * n - 1;
  if (m0 > lmax && o.forced_passes.empty()) {
    // whole convolution in LDS by a specialised instance: one launch (the
    // two-halves instances hold M / 2 at a time)
    const BluesteinLdsShape *pick = nullptr;
    const std::vector<BluesteinLdsShape> cands = bluestein_lds_candidates(p == Prec::F64);
    for (const BluesteinLdsShape &sh : cands)
      if (sh.m >= m0 && sh.m <= 2 * m0 && (pick == nullptr || sh.m < pick->m)) pick = &sh;
    if (pick != nullptr) {
      c.kind = CoreKind::Bluestein;
      c.m = pick->m;
      c.radices = pick->radices;
      c.conv.reset();
      c.lds_conv = true;
      c.lds_halves = pick->halves;
      return c;
    }
  }
  for (int64_t M = m0;; ++M) {
    if (have && static_cast<double>(M) >= best_cost) break;
    if (M > 2 * m0 + 64) {
      if (have) break;
      if (M > 64 * n + 1024) throw std::runtime_error("no Bluestein convolution length found");
    }
    if (!smooth13(M)) continue;
    if (M <= lmax && radix_factor(M, rad)) {
      c.kind = CoreKind::Bluestein;
      c.m = M;
      c.radices = rad;
      c.conv.reset();
      return c;
    }
    Core sub;
    if (!multipass_core(M, p, o, sub, o.queue_form)) continue;
    if (!o.forced_passes.empty()) {
      best = sub;
      have = true;
      break;
    }
    const double cost = static_cast<double>(M) * static_cast<double>(sub.passes.size()) / 2.0 *
                        (passes_specialised(sub.passes, true) ? 1.0 : 2.5);
    if (!have || cost < best_cost) {
      best = sub;
      best_cost = cost;
      have = true;
    }
  }
  c.kind = CoreKind::Bluestein;
  c.m = best.n;
  c.radices.clear();
  c.conv = std::make_shared<Core>(best);
  return c;
}

std::string Core::describe(int indent) const {
  std::ostringstream o;
  std::string pad(static_cast<size_t>(indent), ' ');
  switch (kind) {
    case CoreKind::Leaf:
      o << pad << "Leaf(n=" << n << ", radices=[" << join(radices, ",") << "], passes=1)";
      break;
    case CoreKind::MultiPass:
      o << pad << "MultiPassStockham(n=" << n << ", super_radices=[" << join64(passes, ",") << "], passes="
        << passes.size() << ")";
      for (size_t i = 0; i < passes.size(); ++i)
        o << "\n" << pad << "  pass" << i << ": R=" << passes[i] << " radices=[" << join(pass_radices[i], ",") << "]";
      break;
    case CoreKind::Rader:
      o << pad << "Rader(p=" << n << ", root=" << root << ", conv_length=" << m << ", fused=1)\n"
        << pad << "  conv: Leaf(n=" << m << ", radices=[" << join(radices, ",") << "])";
      break;
    case CoreKind::MatrixCore:
      o << pad << "MatrixCoreSplit(n=" << n << ", factors=[" << join(radices, ",") << "], operands="
        << (compute == 2 ? "fp16" : "bf16") << " hi+lo, products=3, accumulate=fp32)";
      break;
    case CoreKind::Bluestein:
      o << pad << "Bluestein(n=" << n << ", conv_length=" << m << ", fused=" << (conv ? 0 : 1)
        << (lds_conv ? ", lds_resident=1" : "") << (lds_halves ? ", halves=1" : "") << ")\n";
      if (conv)
        o << conv->describe(indent + 2);
      else if (lds_halves)
        o << pad << "  conv: Leaf(n=" << m << ", radices=[2," << join(radices, ",") << "], in LDS: two halves of "
          << m / 2 << ")";
      else
        o << pad << "  conv: Leaf(n=" << m << ", radices=[" << join(radices, ",") << "])";
      break;
  }
  return o.str();
}

std::string Plan1D::describe(int indent) const {
  std::ostringstream o;
  std::string pad(static_cast<size_t>(indent), ' ');
  const char *form = "";
  switch (real_form) {
    case RealForm::None: form = "Complex"; break;
    case RealForm::EvenFused: form = "RealEvenFused"; break;
    case RealForm::EvenMulti: form = "RealEvenSplitPass"; break;
    case RealForm::Full: form = "RealPromoted"; break;
  }
  o << pad << form << "(n=" << n << (real_fwd ? ", R2C" : "") << (real_inv ? ", C2R" : "") << ")\n"
    << core.describe(indent + 2);
  return o.str();
}

namespace {

Plan1D plan_1d(int64_t n, bool real_fwd, bool real_inv, Prec p, const PlanOptions &o) {
  Plan1D pl;
  pl.n = n;
  pl.real_fwd = real_fwd;
  pl.real_inv = real_inv;
  if (!real_fwd && !real_inv) {
    pl.real_form = RealForm::None;
    pl.core = plan_core(n, p, o);
    return pl;
  }
  if (n % 2 == 0 && n >= 2) {
    // half-length complex core, pack / unpack fused (single pass) or as a
    // split-spectrum pass beside the core (any other core kind)
    Core half = plan_core(n / 2, p, o);
    pl.real_form = half.kind == CoreKind::Leaf ? RealForm::EvenFused : RealForm::EvenMulti;
    pl.core = half;
    return pl;
  }
  pl.real_form = RealForm::Full;
  pl.core = plan_core(n, p, o);
  return pl;
}

// A column super-stage reads row segments of (adjacent columns per
// workgroup) x (element bytes).  When the single pass's widest specialised
// tile gives segments under 64 B (a long column filling the LDS with few
// columns), two passes of wide tiles move the data twice at full segment
// width and win (4096-point FP32 columns: 66 us against 108 us per
// 2049 x 4096 half spectrum).
// ncol columns per transform, nseq transforms.
void widen_columns(Plan1D &cols, Prec prec, const PlanOptions &o, int64_t ncol, int64_t nseq) {
  const bool tuned = !o.forced_passes.empty() || !o.forced_radices.empty() || o.forced_units != 0 || o.generic_kernels;
  if (tuned || cols.core.kind != CoreKind::Leaf) return;
  // a register-resident instance of the whole column length reads 64-byte
  // segments or more in one pass (4096-point FP32 columns: 75 us against
  // 95 us for two 64-point passes), when its workgroups (one per CU) fill the
  // 256 CUs in whole rounds (2049 half-spectrum columns are 257 tiles: the
  // last one would run alone, 228 against 173 us for the 4096^2 R2C->C2R
  // step) and tile whole transforms
  const int run = o.stage_form != 2 ? col_run_bytes(cols.n, prec == Prec::F64, false) : 0;
  if (run >= 64) {
    constexpr int64_t kCus = 256;
    const int64_t c = run / elem_bytes(prec);
    const int64_t tiles = (ncol * nseq + c - 1) / c;
    const int64_t rounds = (tiles + kCus - 1) / kCus;
    if ((nseq == 1 || ncol % c == 0) && 4 * tiles >= 3 * rounds * kCus) return;
  }
  // (no LDS super-stage instance for the whole column either: the
  // run-time-radix pass would read it one element per column)
  const int seg = column_segment_bytes(cols.core.radices, prec);
  Core mc;
  if (seg >= 64 || !split_passes(cols.n, 2, pass_max(o, prec), mc.passes)) return;
  for (int64_t r : mc.passes) {
    std::vector<int> rad;
    if (!radix_factor(r, rad) || column_segment_bytes(rad, prec) < 64) return;
    mc.pass_radices.push_back(rad);
  }
  mc.kind = CoreKind::MultiPass;
  mc.n = cols.n;
  cols.core = mc;
}

}  // namespace

PlanTree plan_tree(const PlanDesc &d, const PlanOptions &o) {
  PlanTree t;
  t.desc = d;
  if (d.compute != 0) {
    if (d.rank != 1 || d.real_fwd || d.real_inv || d.prec != Prec::F32 || d.n0 != 1024)
      throw std::invalid_argument("matrix-core compute modes cover C2C n=1024 only");
    t.rows.n = d.n0;
    t.rows.real_form = RealForm::None;
    t.rows.core.kind = CoreKind::MatrixCore;
    t.rows.core.n = d.n0;
    t.rows.core.radices = {32, 32};
    t.rows.core.compute = d.compute;
    return t;
  }
  if (d.rank == 1) {
    t.rows = plan_1d(d.n0, d.real_fwd, d.real_inv, d.prec, o);
  } else if (d.rank == 2) {
    t.rows = plan_1d(d.n1, d.real_fwd, d.real_inv, d.prec, o);
    t.cols = plan_1d(d.n0, false, false, d.prec, o);
    const bool half = d.real_fwd || d.real_inv;
    // one real transform whose n1 / 2 + 1 half-spectrum columns take two
    // passes only because the last tile would run alone: the DC and Nyquist
    // columns (both real on the real side) share a unit of the single pass
    // (4096^2 R2C: one 41-43 us column pass instead of two of 24-25 us)
    if (half && d.batch == 1 && d.n1 % 2 == 0 && !o.generic_kernels && t.rows.real_form == RealForm::EvenFused &&
        leaf_kernel_id(t.rows.core.radices) >= 0 && t.cols.core.kind == CoreKind::Leaf &&
        col_pair_kernel_id(d.n0, d.prec == Prec::F64) >= 0) {
      Plan1D split = t.cols, single = t.cols;
      widen_columns(split, d.prec, o, d.n1 / 2 + 1, 1);
      widen_columns(single, d.prec, o, d.n1 / 2, 1);
      t.dc_nyq_pair = split.core.kind != CoreKind::Leaf && single.core.kind == CoreKind::Leaf;
    }
    if (!t.dc_nyq_pair) widen_columns(t.cols, d.prec, o, half ? d.n1 / 2 + 1 : d.n1, d.batch);
  } else {
    t.rows = plan_1d(d.n2, d.real_fwd, d.real_inv, d.prec, o);
    t.mid = plan_1d(d.n1, false, false, d.prec, o);
    t.cols = plan_1d(d.n0, false, false, d.prec, o);
    const int64_t n2c = (d.real_fwd || d.real_inv) ? d.n2 / 2 + 1 : d.n2;
    widen_columns(t.mid, d.prec, o, n2c, d.n0 * d.batch);
    widen_columns(t.cols, d.prec, o, d.n1 * n2c, d.batch);
  }
  return t;
}

std::string PlanTree::describe() const {
  std::ostringstream o;
  o << "=== FlagFFT Plan (gfx950) ===\n";
  o << "rank=" << desc.rank << " n=[" << desc.n0;
  if (desc.rank >= 2) o << "," << desc.n1;
  if (desc.rank == 3) o << "," << desc.n2;
  o << "] batch=" << desc.batch << " type=" << desc.type << " precision=" << (desc.prec == Prec::F32 ? "fp32" : "fp64")
    << "\n\n-- Plan tree --\n";
  if (desc.rank == 1) {
    o << rows.describe(0) << "\n";
  } else if (desc.rank == 2) {
    o << "TwoDim(n0=" << desc.n0 << ", n1=" << desc.n1 << ", order=rows-then-columns, column_pass=strided-tile"
      << (dc_nyq_pair ? ", dc_nyquist_pair" : "") << ")\n";
    o << "  rows (axis 1):\n" << rows.describe(4) << "\n";
    o << "  columns (axis 0):\n" << cols.describe(4) << "\n";
  } else {
    o << "ThreeDim(n0=" << desc.n0 << ", n1=" << desc.n1 << ", n2=" << desc.n2
      << ", order=rows-then-axis1-then-axis0, column_pass=strided-tile)\n";
    o << "  rows (axis 2):\n" << rows.describe(4) << "\n";
    o << "  columns (axis 1):\n" << mid.describe(4) << "\n";
    o << "  columns (axis 0):\n" << cols.describe(4) << "\n";
  }
  return o.str();
}

}  // namespace flagfft
