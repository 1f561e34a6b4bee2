// Tile-shape autotuner for the matrix-core GEMMs.
//
// The FP64 and split-bf16 products pick a workgroup tile and a split-K count
// from hand-set thresholds.  With tuning on, the first product of every new
// shape (kind, M, N, K, operand flags) instead times each candidate (the
// four tile shapes, each with split-K counts around its default that fit the
// caller's scratch) on the caller's stream with HIP events, keeps the
// fastest, and re-runs it so that the result is the chosen kernel's.  Later
// products of that shape use the table; a CTMRG iteration repeats the same
// shapes, so one warmup iteration tunes a whole run.  The table can be saved
// to and loaded from a text file.  Tuning is skipped under stream capture
// and for accumulating products (beta != 0, which cannot be re-run), which
// then use the table when it has the shape, else the thresholds.
#include "internal.h"

#include <cstdio>
#include <map>
#include <mutex>
#include <tuple>

namespace acetn {
namespace {

struct Tuner {
  std::mutex mu;
  int mode = 0;  // 0: thresholds only, 1: table + tune new shapes, 2: table only
  std::map<std::tuple<int, int, int, int, int>, TuneChoice> table;
};
Tuner& tuner() {
  static Tuner t;
  return t;
}
std::tuple<int, int, int, int, int> key_of(const TuneKey& k) {
  return std::make_tuple(k.kind, k.M, k.N, k.K, k.flags);
}

}  // namespace

void autotune_set_mode(int mode) {
  ACETN_REQUIRE(mode >= 0 && mode <= 2, "autotune: mode must be 0, 1 or 2");
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  t.mode = mode;
}

int autotune_mode() {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  return t.mode;
}

bool autotune_lookup(const TuneKey& k, TuneChoice* out) {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  if (t.mode == 0) return false;
  auto it = t.table.find(key_of(k));
  if (it == t.table.end()) return false;
  *out = it->second;
  return true;
}

bool autotune_should_tune(const TuneKey& k, const GemmDesc& g, hipStream_t s) {
  if (g.beta != 0.0 || autotune_mode() != 1) return false;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return false;
  TuneChoice c;
  return !autotune_lookup(k, &c);
}

TuneChoice autotune_run(const TuneKey& k, const std::vector<TuneChoice>& cands, hipStream_t s,
                        const std::function<void(const TuneChoice&)>& launch) {
  ACETN_REQUIRE(!cands.empty(), "autotune: no candidate");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ACETN_HIP_CHECK(hipEventCreate(&e0));
  ACETN_HIP_CHECK(hipEventCreate(&e1));
  constexpr int kReps = 3;
  float best = 0.f;
  size_t bi = 0;
  for (size_t i = 0; i < cands.size(); ++i) {
    launch(cands[i]);  // warm (first-launch and cache effects)
    ACETN_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < kReps; ++r) launch(cands[i]);
    ACETN_HIP_CHECK(hipEventRecord(e1, s));
    ACETN_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    ACETN_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    if (i == 0 || ms < best) {
      best = ms;
      bi = i;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  TuneChoice c = cands[bi];
  c.ms = best / kReps;
  {
    Tuner& t = tuner();
    std::lock_guard<std::mutex> g(t.mu);
    t.table[key_of(k)] = c;
  }
  launch(c);  // the result the caller keeps comes from the chosen kernel
  return c;
}

int64_t autotune_entries() {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  return (int64_t)t.table.size();
}

void autotune_clear() {
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  t.table.clear();
}

void autotune_save(const char* path) {
  ACETN_REQUIRE(path != nullptr, "autotune: null path");
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  FILE* f = std::fopen(path, "w");
  ACETN_REQUIRE(f != nullptr, std::string("autotune: cannot write ") + path);
  std::fprintf(f, "# acetn gemm tile table: kind M N K flags wm wn ks ms variant\n");
  for (const auto& kv : t.table) {
    const auto& k = kv.first;
    const TuneChoice& c = kv.second;
    std::fprintf(f, "%d %d %d %d %d %d %d %d %.6f %d\n", std::get<0>(k), std::get<1>(k),
                 std::get<2>(k), std::get<3>(k), std::get<4>(k), c.wm, c.wn, c.ks, c.ms, c.var);
  }
  std::fclose(f);
}

int64_t autotune_load(const char* path) {
  ACETN_REQUIRE(path != nullptr, "autotune: null path");
  FILE* f = std::fopen(path, "r");
  ACETN_REQUIRE(f != nullptr, std::string("autotune: cannot read ") + path);
  Tuner& t = tuner();
  std::lock_guard<std::mutex> g(t.mu);
  char line[256];
  int64_t n = 0;
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    int kind, M, N, K, flags, wm, wn, ks, var = 0;
    double ms = 0.0;
    if (std::sscanf(line, "%d %d %d %d %d %d %d %d %lf %d", &kind, &M, &N, &K, &flags, &wm, &wn,
                    &ks, &ms, &var) < 8)
      continue;
    const bool tile_ok = (wm == 1 && wn == 5) || (wm == 2 && wn == 5) || (wm == 2 && wn == 2) ||
                         (wm == 4 && wn == 4);
    // the wider FP64 variants exist for single-stride products: 2 x 5 tiles
    // in variants 1-3, 1 x 5 tiles in variant 3; two-piece split products
    // have 64-deep k-tiles as variant 1
    const bool var_ok = var == 0 || (kind == 0 && wn == 5 && (flags & 1) &&
                                     ((wm == 2 && var <= 3) || (wm == 1 && var == 3))) ||
                        (kind == 2 && var == 1);
    if (!tile_ok || !var_ok || ks < 1 || ks > 64) continue;
    TuneChoice c{wm, wn, ks, ms};
    c.var = var;
    t.table[std::make_tuple(kind, M, N, K, flags)] = c;
    ++n;
  }
  std::fclose(f);
  return n;
}

}  // namespace acetn
