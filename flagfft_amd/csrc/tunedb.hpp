// tunedb.hpp — tuned-plan database (see tunedb.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace flagfft {

constexpr int64_t kTuneSchemaVersion = 3;
constexpr const char *kTuneDeviceArch = "gfx950";
constexpr const char *kRuntimeFingerprint = "hip-runtime-1";

// Lookup key of one transform kind (1D).
struct TuneKey {
  int64_t n = 0;
  int64_t batch = 1;
  std::string dtype;      // input element type: complex64, complex128, float32, float64
  std::string direction;  // forward / inverse
  std::string layout;     // contiguous (complex), real (R2C / D2Z input), hermitian (C2R / Z2D input)
};

// Plan-time overrides a tuned row holds (flagfftxTuning without the C arrays).
struct TuneOverrides {
  int64_t leaf_max = 0, pass_max = 0;
  int units = 0, ept = 0, chunk_bytes = 0, stage_form = 0, variant = 0;
  std::vector<int64_t> passes;
  std::vector<int> radices;
};

std::string batch_bucket(int64_t batch);
// Database file: $FLAGFFT_TUNE_DB, else tuned_plans.sqlite beside libflagfft.so;
// empty when FLAGFFT_TUNE_DISABLE is set.
std::string tuned_db_path();
bool tuned_lookup(const TuneKey &k, std::string &plan_json);
bool tuned_store(const TuneKey &k, const std::string &plan_json, double median_us, double gbps, double default_us,
                 std::string &err);
std::string tuning_to_json(const TuneOverrides &t);
bool tuning_from_json(const std::string &js, TuneOverrides &t);

}  // namespace flagfft
