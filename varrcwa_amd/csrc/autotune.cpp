// Autotuner of the kernel switches per harmonic count. The search times the caller's own
// batched stack problem (a sample of it) end to end under candidate switch settings,
// one switch at a time in order of impact (coordinate descent), and keeps a candidate
// only if it is at least 1 % faster than the best so far. Winners are kept per N in a
// process-wide table that tune_apply() puts in place before each call solving N x N
// systems, and can be written to / read from a plain text cache file.
#include <chrono>
#include <cstdio>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>

#include "vrc_ops.hpp"

namespace vrc {

namespace {
std::mutex g_mu;
std::map<int, TuneKnobs> g_table;  // autotuned switches per N
bool g_over = false;               // a table entry is in place of the caller's own values
TuneKnobs g_own{};                 // the caller's own values while an entry is in place
}  // namespace

TuneKnobs tune_get() {
  TuneKnobs k;
  k.aed = zgeev_get_aed();
  k.bulges = zgeev_get_bulges();
  k.qr_variant = zgeev_get_qr_variant();
  zgemm_get_knobs(&k.big_n, &k.big_waves, &k.small_waves, &k.big_min_k);
  return k;
}

void tune_put(const TuneKnobs& k) {
  zgeev_set_aed(k.aed);
  zgeev_set_bulges(k.bulges);
  zgeev_set_qr_variant(k.qr_variant);
  zgemm_set_big_n(k.big_n);
  zgemm_set_big_waves(k.big_waves);
  zgemm_set_small_waves(k.small_waves);
  zgemm_set_big_min_k(k.big_min_k);
}

void tune_apply(int N) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find(N);
  if (it != g_table.end()) {
    if (!g_over) {
      g_own = tune_get();
      g_over = true;
    }
    tune_put(it->second);
  } else if (g_over) {
    tune_put(g_own);
    g_over = false;
  }
}

void tune_manual_begin() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_over) {
    tune_put(g_own);
    g_over = false;
  }
}

static std::string knobs_text(const TuneKnobs& k) {
  char b[192];
  snprintf(b, sizeof(b),
           "qr_variant %d aed_window %d bulges %d big_n %d big_waves %d small_waves %d large_tile_min_k %d",
           k.qr_variant, k.aed, k.bulges, k.big_n, k.big_waves, k.small_waves, k.big_min_k);
  return b;
}

void tune_store(int N, const TuneKnobs& k) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[N] = k;
}

void tune_clear() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_over) tune_put(g_own);
  g_over = false;
  g_table.clear();
}

// Cache file: one line per N, "N aed bulges big_n big_waves small_waves big_min_k
// [qr_variant]" (a line without the last field leaves the QR kernel to the automatic
// choice, -1); lines starting with '#' are comments. Later lines for the same N win.
void tune_save(const std::string& path, int N, const TuneKnobs& k, const std::string& note) {
  std::ofstream f(path, std::ios::app);
  if (!f) throw ArgFailure("cannot append to the tuning cache " + path);
  f << "# " << note << "\n"
    << N << " " << k.aed << " " << k.bulges << " " << k.big_n << " " << k.big_waves << " " << k.small_waves << " "
    << k.big_min_k << " " << k.qr_variant << "\n";
}

int tune_load(const std::string& path) {
  std::ifstream f(path);
  if (!f) throw ArgFailure("cannot read the tuning cache " + path);
  std::string line;
  int n = 0;
  while (std::getline(f, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream is(line);
    int N;
    TuneKnobs k{};
    if (!(is >> N >> k.aed >> k.bulges >> k.big_n >> k.big_waves >> k.small_waves >> k.big_min_k) || N < 2)
      throw ArgFailure("malformed tuning cache line: " + line);
    if (!(is >> k.qr_variant)) k.qr_variant = -1;
    if (k.qr_variant != -1 && k.qr_variant != 0 && k.qr_variant != 2 && k.qr_variant != 4)
      throw ArgFailure("malformed tuning cache line: " + line);
    tune_store(N, k);
    ++n;
  }
  return n;
}

std::string autotune_stack(vrc_ctx& c, int nx, int ny, const std::vector<StackInstance>& sample, double budget_s,
                           TuneKnobs* best_out) {
  const int N = 2 * nx * ny;
  const long NN = (long)N * N;
  {
    std::lock_guard<std::mutex> lk(g_mu);  // search from the caller's own values
    g_table.erase(N);
    if (g_over) {
      tune_put(g_own);
      g_over = false;
    }
  }
  const TuneKnobs own = tune_get();
  TuneKnobs best = own;
  DevBuf S = c.ws.get(sizeof(cd) * 4 * NN * sample.size());
  const bool timing = c.timing;
  c.timing = false;
  using clk = std::chrono::steady_clock;
  const auto t_start = clk::now();
  auto elapsed = [&] { return std::chrono::duration<double>(clk::now() - t_start).count(); };
  int runs = 0;
  auto run = [&](const TuneKnobs& k) {
    tune_put(k);
    VRC_HIP(hipStreamSynchronize(c.stream));
    const auto t0 = clk::now();
    batch_stack_rcwa(c, nx, ny, sample, 0, S.as<cd>(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr);
    VRC_HIP(hipStreamSynchronize(c.stream));
    ++runs;
    return std::chrono::duration<double>(clk::now() - t0).count();
  };
  try {
    run(own);  // warm: code loading, pool growth
    const double t_own = run(own);
    double t_best = t_own;
    const std::vector<std::pair<int TuneKnobs::*, std::vector<int>>> space = {
        {&TuneKnobs::qr_variant, {0, 2, 4}},
        {&TuneKnobs::aed, {16, 20, 24, 32, 40, 48}},  {&TuneKnobs::bulges, {6, 8, 10, 12, 14, 16}},
        {&TuneKnobs::big_n, {64, 128}},              {&TuneKnobs::big_waves, {4, 8}},
        {&TuneKnobs::small_waves, {4, 8}},           {&TuneKnobs::big_min_k, {33, 65, 129}}};
    bool out_of_time = false;
    for (auto& [field, values] : space) {
      for (int v : values) {
        if (best.*field == v) continue;
        if (elapsed() > budget_s) {
          out_of_time = true;
          break;
        }
        TuneKnobs k = best;
        k.*field = v;
        const double t = run(k);
        if (t < 0.99 * t_best) {
          t_best = t;
          best = k;
        }
      }
      if (out_of_time) break;
    }
    tune_put(own);
    c.timing = timing;
    if (best_out) *best_out = best;
    char b[256];
    snprintf(b, sizeof(b), "N=%d: %s; sample %zu instances %.3f s (own switches %.3f s), %d runs in %.1f s%s", N,
             knobs_text(best).c_str(), sample.size(), t_best, t_own, runs, elapsed(),
             out_of_time ? " (budget reached)" : "");
    return b;
  } catch (...) {
    tune_put(own);
    c.timing = timing;
    throw;
  }
}

}  // namespace vrc
