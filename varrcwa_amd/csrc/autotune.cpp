// Autotuner of the kernel switches per harmonic count. The search times the caller's own
// batched stack problem (a sample of it) end to end under candidate switch settings,
// one switch at a time in order of impact (coordinate descent), and keeps a candidate
// only if it is at least 1 % faster than the best so far. Winners are kept per N in a
// process-wide table that tune_apply() puts in place before each call solving N x N
// systems, and can be written to / read from a plain text cache file.
#include <algorithm>
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
  k.bulges_w48 = zgeev_get_bulges_w48();
  k.bulges_w32 = zgeev_get_bulges_w32();
  zgemm_get_knobs(&k.big_n, &k.big_waves, &k.small_waves, &k.big_min_k);
  return k;
}

void tune_put(const TuneKnobs& k) {
  zgeev_set_aed(k.aed);
  zgeev_set_bulges(k.bulges);
  zgeev_set_qr_variant(k.qr_variant);
  zgeev_set_bulges_w48(k.bulges_w48);
  zgeev_set_bulges_w32(k.bulges_w32);
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
  char b[256];
  snprintf(b, sizeof(b),
           "qr_variant %d aed_window %d bulges %d bulges_w48 %d bulges_w32 %d big_n %d big_waves %d small_waves %d "
           "large_tile_min_k %d",
           k.qr_variant, k.aed, k.bulges, k.bulges_w48, k.bulges_w32, k.big_n, k.big_waves, k.small_waves,
           k.big_min_k);
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
// [qr_variant [bulges_w48 bulges_w32]]" (without qr_variant the QR kernel is left to the
// automatic choice, -1; without the last two the bulge counts of QR variants 2 and 4 are
// their defaults, 8 and 6); lines starting with '#' are comments. Later lines for the same
// N win.
void tune_save(const std::string& path, int N, const TuneKnobs& k, const std::string& note) {
  std::ofstream f(path, std::ios::app);
  if (!f) throw ArgFailure("cannot append to the tuning cache " + path);
  f << "# " << note << "\n"
    << N << " " << k.aed << " " << k.bulges << " " << k.big_n << " " << k.big_waves << " " << k.small_waves << " "
    << k.big_min_k << " " << k.qr_variant << " " << k.bulges_w48 << " " << k.bulges_w32 << "\n";
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
    k.qr_variant = -1;
    k.bulges_w48 = 8;
    k.bulges_w32 = 6;
    int qv, b48, b32;  // optional fields (a failed read leaves the defaults in place)
    if (is >> qv) {
      k.qr_variant = qv;
      if (is >> b48) {
        if (!(is >> b32)) throw ArgFailure("malformed tuning cache line: " + line);
        k.bulges_w48 = b48;
        k.bulges_w32 = b32;
      }
    }
    is.clear();
    std::string rest;
    if ((k.qr_variant != -1 && k.qr_variant != 0 && k.qr_variant != 2 && k.qr_variant != 4) || k.bulges_w48 < 1 ||
        k.bulges_w48 > 12 || k.bulges_w32 < 1 || k.bulges_w32 > 7 || (is >> rest))
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
  unsigned ran = 0;  // QR kernel variants the last run launched (bit v: variant v)
  auto run = [&](const TuneKnobs& k) {
    tune_put(k);
    VRC_HIP(hipStreamSynchronize(c.stream));
    zgeev_take_qr_kernels();
    const auto t0 = clk::now();
    batch_stack_rcwa(c, nx, ny, sample, 0, S.as<cd>(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                     nullptr);
    VRC_HIP(hipStreamSynchronize(c.stream));
    ran = zgeev_take_qr_kernels();
    ++runs;
    return std::chrono::duration<double>(clk::now() - t0).count();
  };
  // A switch is searched only where it can change what runs: the QR switches only when the
  // best setting so far launches a QR kernel, the bulge count of each QR kernel only when it
  // launches that kernel, and a deflation window only when, clamped to the window limit
  // (W - 1) of each kernel launched, it differs from the best one's.
  constexpr unsigned QR_ANY = 0u, QR_W64 = (1u << 0) | (1u << 1), QR_W48 = 1u << 2, QR_W32 = 1u << 4;
  constexpr unsigned QR_ALL = QR_W64 | QR_W48 | QR_W32;
  auto aed_effective = [](int v, unsigned kernels) {
    std::vector<int> e;
    if (kernels & QR_W64) e.push_back(std::min(v, 63));
    if (kernels & QR_W48) e.push_back(std::min(v, 47));
    if (kernels & QR_W32) e.push_back(std::min(v, 31));
    return e;
  };
  try {
    run(own);  // warm: code loading, pool growth
    const double t_own = run(own);
    double t_best = t_own;
    unsigned ran_best = ran;
    struct Switch {
      int TuneKnobs::*field;
      std::vector<int> values;
      unsigned kernels;  // QR kernels the switch acts on (QR_ANY: every call, QR_ALL: any QR kernel)
    };
    const std::vector<Switch> space = {
        {&TuneKnobs::qr_variant, {0, 2, 4}, QR_ALL},
        {&TuneKnobs::aed, {16, 20, 24, 32, 40, 48}, QR_ALL},
        {&TuneKnobs::bulges, {6, 8, 10, 12, 14, 16}, QR_W64},
        {&TuneKnobs::bulges_w48, {4, 6, 8, 10, 12}, QR_W48},
        {&TuneKnobs::bulges_w32, {3, 4, 5, 6, 7}, QR_W32},
        {&TuneKnobs::big_n, {64, 128}, QR_ANY},
        {&TuneKnobs::big_waves, {4, 8}, QR_ANY},
        {&TuneKnobs::small_waves, {4, 8}, QR_ANY},
        {&TuneKnobs::big_min_k, {33, 65, 129}, QR_ANY}};
    bool out_of_time = false;
    for (const Switch& sw : space) {
      if (sw.kernels != QR_ANY && !(sw.kernels & ran_best)) continue;
      for (int v : sw.values) {
        if (best.*sw.field == v) continue;
        if (sw.field == &TuneKnobs::aed && aed_effective(v, ran_best) == aed_effective(best.aed, ran_best)) continue;
        if (elapsed() > budget_s) {
          out_of_time = true;
          break;
        }
        TuneKnobs k = best;
        k.*sw.field = v;
        const double t = run(k);
        if (t < 0.99 * t_best) {
          t_best = t;
          best = k;
          ran_best = ran;
        }
      }
      if (out_of_time) break;
    }
    tune_put(own);
    c.timing = timing;
    if (best_out) *best_out = best;
    char b[384];
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
