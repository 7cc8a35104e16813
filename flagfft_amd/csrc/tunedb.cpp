// tunedb.cpp — tuned-plan database (SQLite, loaded at run time).
//
// Same table and lookup as the reference's tuned-plan cache
// (src/utils/plan_json.cpp:170-207, src/exec/tuned_plan.cpp:21-58): table
// `tuned_measurements`, newest valid row for (schema_version, device_arch,
// fft_length, batch_bucket, dtype, direction, norm, input_layout, planner /
// codegen / runtime fingerprints), batch buckets 1 / 2-8 / 9-64 / 65-512 /
// 513+, FLAGFFT_TUNE_DB overrides the path and FLAGFFT_TUNE_DISABLE turns the
// lookup off.  The reference only reads this table; here the autotuner
// (flagfft_amd/autotune.py) writes it through flagfftxTunedStore.  The plan
// text is this library's own: the flagfftxTuning overrides as a flat JSON
// object.  libsqlite3 is opened with dlopen so the library has no link-time
// dependency on it; without it lookups find nothing and stores fail.
#include "tunedb.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <mutex>
#include <sstream>

namespace flagfft {

// Planner and kernel fingerprints: hashes of the planner sources and of the
// device sources, computed by the build (flagfft_amd/build.py), so a row
// measured with other kernels or another planner never applies.
#ifndef FLAGFFT_PLANNER_HASH
#error "FLAGFFT_PLANNER_HASH must be defined by the build"
#endif
#ifndef FLAGFFT_KERNEL_HASH
#error "FLAGFFT_KERNEL_HASH must be defined by the build"
#endif
constexpr const char *kPlannerFingerprint = "gfx950-planner-" FLAGFFT_PLANNER_HASH;
constexpr const char *kCodegenFingerprint = "gfx950-kernels-" FLAGFFT_KERNEL_HASH;
namespace {

struct sqlite3;
struct sqlite3_stmt;

struct Sqlite {
  int (*open_v2)(const char *, sqlite3 **, int, const char *) = nullptr;
  int (*close)(sqlite3 *) = nullptr;
  int (*prepare_v2)(sqlite3 *, const char *, int, sqlite3_stmt **, const char **) = nullptr;
  int (*bind_int64)(sqlite3_stmt *, int, long long) = nullptr;
  int (*bind_double)(sqlite3_stmt *, int, double) = nullptr;
  int (*bind_text)(sqlite3_stmt *, int, const char *, int, void (*)(void *)) = nullptr;
  int (*step)(sqlite3_stmt *) = nullptr;
  const unsigned char *(*column_text)(sqlite3_stmt *, int) = nullptr;
  int (*finalize)(sqlite3_stmt *) = nullptr;
  int (*exec)(sqlite3 *, const char *, int (*)(void *, int, char **, char **), void *, char **) = nullptr;
  void (*free_)(void *) = nullptr;
  int (*busy_timeout)(sqlite3 *, int) = nullptr;
  bool ok = false;
};

constexpr int kOpenReadOnly = 0x1, kOpenReadWrite = 0x2, kOpenCreate = 0x4;
constexpr int kRow = 100, kDone = 101;
void (*const kTransient)(void *) = reinterpret_cast<void (*)(void *)>(-1);

const Sqlite &sqlite() {
  static Sqlite s;
  static std::once_flag once;
  std::call_once(once, [] {
    void *h = dlopen("libsqlite3.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!h) h = dlopen("libsqlite3.so", RTLD_NOW | RTLD_LOCAL);
    if (!h) return;
    auto sym = [&](auto &fn, const char *name) {
      fn = reinterpret_cast<std::remove_reference_t<decltype(fn)>>(dlsym(h, name));
      return fn != nullptr;
    };
    s.ok = sym(s.open_v2, "sqlite3_open_v2") && sym(s.close, "sqlite3_close") &&
           sym(s.prepare_v2, "sqlite3_prepare_v2") && sym(s.bind_int64, "sqlite3_bind_int64") &&
           sym(s.bind_double, "sqlite3_bind_double") && sym(s.bind_text, "sqlite3_bind_text") &&
           sym(s.step, "sqlite3_step") && sym(s.column_text, "sqlite3_column_text") &&
           sym(s.finalize, "sqlite3_finalize") && sym(s.exec, "sqlite3_exec") && sym(s.free_, "sqlite3_free") &&
           sym(s.busy_timeout, "sqlite3_busy_timeout");
  });
  return s;
}

bool env_on(const char *v) {
  if (v == nullptr || *v == '\0') return false;
  std::string s(v);
  for (char &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return !(s == "0" || s == "false" || s == "off" || s == "no");
}

std::string library_dir() {
  Dl_info info;
  if (dladdr(reinterpret_cast<void *>(&library_dir), &info) != 0 && info.dli_fname != nullptr) {
    std::string p(info.dli_fname);
    const size_t slash = p.rfind('/');
    if (slash != std::string::npos) return p.substr(0, slash);
  }
  return ".";
}

const char *kCreate =
    "CREATE TABLE IF NOT EXISTS tuned_measurements ("
    " plan_json TEXT NOT NULL, schema_version INTEGER NOT NULL, status TEXT NOT NULL, rank INTEGER NOT NULL,"
    " device_arch TEXT NOT NULL, fft_length INTEGER NOT NULL, batch_bucket TEXT NOT NULL, dtype TEXT NOT NULL,"
    " direction TEXT NOT NULL, norm TEXT NOT NULL, input_layout TEXT NOT NULL,"
    " planner_fingerprint TEXT NOT NULL, codegen_fingerprint TEXT NOT NULL, runtime_fingerprint TEXT NOT NULL,"
    " measured_at TEXT NOT NULL, batch INTEGER, median_us REAL, gbps REAL, default_us REAL)";

}  // namespace

std::string batch_bucket(int64_t batch) {
  if (batch <= 1) return "1";
  if (batch <= 8) return "2-8";
  if (batch <= 64) return "9-64";
  if (batch <= 512) return "65-512";
  return "513+";
}

std::string tuned_db_path() {
  if (env_on(std::getenv("FLAGFFT_TUNE_DISABLE"))) return std::string();
  const char *p = std::getenv("FLAGFFT_TUNE_DB");
  if (p != nullptr && *p != '\0') return std::string(p);
  return library_dir() + "/tuned_plans.sqlite";
}

bool tuned_lookup(const TuneKey &k, std::string &plan_json) {
  const std::string path = tuned_db_path();
  if (path.empty()) return false;
  const Sqlite &s = sqlite();
  if (!s.ok) return false;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (f == nullptr) return false;
  std::fclose(f);
  sqlite3 *db = nullptr;
  if (s.open_v2(path.c_str(), &db, kOpenReadOnly, nullptr) != 0) {
    if (db) s.close(db);
    return false;
  }
  s.busy_timeout(db, 2000);
  sqlite3_stmt *st = nullptr;
  bool found = false;
  const char *q =
      "SELECT plan_json FROM tuned_measurements WHERE schema_version=? AND status='valid' AND rank=0 "
      "AND device_arch=? AND fft_length=? AND batch_bucket=? AND dtype=? AND direction=? AND norm=? "
      "AND input_layout=? AND planner_fingerprint=? AND codegen_fingerprint=? AND runtime_fingerprint=? "
      "ORDER BY measured_at DESC, rowid DESC LIMIT 1";
  if (s.prepare_v2(db, q, -1, &st, nullptr) == 0) {
    s.bind_int64(st, 1, kTuneSchemaVersion);
    s.bind_text(st, 2, kTuneDeviceArch, -1, kTransient);
    s.bind_int64(st, 3, k.n);
    s.bind_text(st, 4, batch_bucket(k.batch).c_str(), -1, kTransient);
    s.bind_text(st, 5, k.dtype.c_str(), -1, kTransient);
    s.bind_text(st, 6, k.direction.c_str(), -1, kTransient);
    s.bind_text(st, 7, "backward", -1, kTransient);
    s.bind_text(st, 8, k.layout.c_str(), -1, kTransient);
    s.bind_text(st, 9, kPlannerFingerprint, -1, kTransient);
    s.bind_text(st, 10, kCodegenFingerprint, -1, kTransient);
    s.bind_text(st, 11, kRuntimeFingerprint, -1, kTransient);
    if (s.step(st) == kRow) {
      const unsigned char *t = s.column_text(st, 0);
      if (t != nullptr) {
        plan_json = reinterpret_cast<const char *>(t);
        found = true;
      }
    }
    s.finalize(st);
  }
  s.close(db);
  return found;
}

bool tuned_store(const TuneKey &k, const std::string &plan_json, double median_us, double gbps, double default_us,
                 std::string &err) {
  const std::string path = tuned_db_path();
  if (path.empty()) {
    err = "tuned-plan database disabled (FLAGFFT_TUNE_DISABLE)";
    return false;
  }
  const Sqlite &s = sqlite();
  if (!s.ok) {
    err = "libsqlite3.so.0 not available";
    return false;
  }
  sqlite3 *db = nullptr;
  if (s.open_v2(path.c_str(), &db, kOpenReadWrite | kOpenCreate, nullptr) != 0) {
    err = "cannot open " + path;
    if (db) s.close(db);
    return false;
  }
  s.busy_timeout(db, 5000);
  char *msg = nullptr;
  if (s.exec(db, kCreate, nullptr, nullptr, &msg) != 0) {
    err = msg ? msg : "create table failed";
    if (msg) s.free_(msg);
    s.close(db);
    return false;
  }
  char stamp[32];
  const std::time_t now = std::time(nullptr);
  std::tm tmv;
  gmtime_r(&now, &tmv);
  std::strftime(stamp, sizeof(stamp), "%Y-%m-%dT%H:%M:%SZ", &tmv);
  sqlite3_stmt *st = nullptr;
  const char *ins =
      "INSERT INTO tuned_measurements (plan_json, schema_version, status, rank, device_arch, fft_length, "
      "batch_bucket, dtype, direction, norm, input_layout, planner_fingerprint, codegen_fingerprint, "
      "runtime_fingerprint, measured_at, batch, median_us, gbps, default_us) "
      "VALUES (?, ?, 'valid', 0, ?, ?, ?, ?, ?, 'backward', ?, ?, ?, ?, ?, ?, ?, ?, ?)";
  bool ok = false;
  if (s.prepare_v2(db, ins, -1, &st, nullptr) == 0) {
    s.bind_text(st, 1, plan_json.c_str(), -1, kTransient);
    s.bind_int64(st, 2, kTuneSchemaVersion);
    s.bind_text(st, 3, kTuneDeviceArch, -1, kTransient);
    s.bind_int64(st, 4, k.n);
    s.bind_text(st, 5, batch_bucket(k.batch).c_str(), -1, kTransient);
    s.bind_text(st, 6, k.dtype.c_str(), -1, kTransient);
    s.bind_text(st, 7, k.direction.c_str(), -1, kTransient);
    s.bind_text(st, 8, k.layout.c_str(), -1, kTransient);
    s.bind_text(st, 9, kPlannerFingerprint, -1, kTransient);
    s.bind_text(st, 10, kCodegenFingerprint, -1, kTransient);
    s.bind_text(st, 11, kRuntimeFingerprint, -1, kTransient);
    s.bind_text(st, 12, stamp, -1, kTransient);
    s.bind_int64(st, 13, k.batch);
    s.bind_double(st, 14, median_us);
    s.bind_double(st, 15, gbps);
    s.bind_double(st, 16, default_us);
    ok = s.step(st) == kDone;
    if (!ok) err = "insert failed";
    s.finalize(st);
  } else {
    err = "prepare failed";
  }
  s.close(db);
  return ok;
}

// ---- the plan text: {"leaf_max":..,"pass_max":..,"units":..,"ept":..,"chunk_bytes":..,
//      "passes":[..],"radices":[..]} -------------------------------------------
std::string tuning_to_json(const TuneOverrides &t) {
  std::ostringstream o;
  o << "{\"leaf_max\":" << t.leaf_max << ",\"pass_max\":" << t.pass_max << ",\"units\":" << t.units
    << ",\"ept\":" << t.ept << ",\"chunk_bytes\":" << t.chunk_bytes << ",\"stage_form\":" << t.stage_form << ",\"variant\":" << t.variant
    << ",\"passes\":[";
  for (size_t i = 0; i < t.passes.size(); ++i) o << (i ? "," : "") << t.passes[i];
  o << "],\"radices\":[";
  for (size_t i = 0; i < t.radices.size(); ++i) o << (i ? "," : "") << t.radices[i];
  o << "]}";
  return o.str();
}

bool tuning_from_json(const std::string &js, TuneOverrides &t) {
  auto find = [&](const char *key, size_t &pos) {
    const std::string k = std::string("\"") + key + "\"";
    pos = js.find(k);
    if (pos == std::string::npos) return false;
    pos = js.find(':', pos + k.size());
    if (pos == std::string::npos) return false;
    ++pos;
    while (pos < js.size() && (js[pos] == ' ' || js[pos] == '\t')) ++pos;
    return true;
  };
  auto num = [&](const char *key, int64_t &out) {
    size_t p;
    if (!find(key, p)) return true;  // absent: keep default
    char *end = nullptr;
    out = std::strtoll(js.c_str() + p, &end, 10);
    return end != js.c_str() + p;
  };
  auto arr = [&](const char *key, std::vector<int64_t> &out) {
    size_t p;
    out.clear();
    if (!find(key, p)) return true;
    if (p >= js.size() || js[p] != '[') return false;
    ++p;
    while (p < js.size() && js[p] != ']') {
      char *end = nullptr;
      const long long v = std::strtoll(js.c_str() + p, &end, 10);
      if (end == js.c_str() + p) return false;
      out.push_back(v);
      p = static_cast<size_t>(end - js.c_str());
      while (p < js.size() && (js[p] == ',' || js[p] == ' ')) ++p;
    }
    return p < js.size();
  };
  int64_t v;
  std::vector<int64_t> a;
  if (!num("leaf_max", v = 0)) return false;
  t.leaf_max = v;
  if (!num("pass_max", v = 0)) return false;
  t.pass_max = v;
  if (!num("units", v = 0)) return false;
  t.units = static_cast<int>(v);
  if (!num("ept", v = 0)) return false;
  t.ept = static_cast<int>(v);
  if (!num("chunk_bytes", v = 0)) return false;
  t.chunk_bytes = static_cast<int>(v);
  if (!num("stage_form", v = 0)) return false;
  t.stage_form = static_cast<int>(v);
  if (!num("variant", v = 0)) return false;
  t.variant = static_cast<int>(v);
  if (!arr("passes", a)) return false;
  t.passes = a;
  if (!arr("radices", a)) return false;
  t.radices.assign(a.begin(), a.end());
  return true;
}

}  // namespace flagfft
