// Strip-length autotuner of the interior sweeps (engine.h: tuned_strip).
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "engine.h"
#include "seis_internal.h"

namespace seis {
namespace {

constexpr int kDefaultStrip = 16;
constexpr int kCandidates[] = {8, 16, 32, 64};
// Small launches (fewer padded cells over all shots than kSmallCells): at a
// strip of 8 columns they hold too few wavefronts to fill 256 CUs, and each
// wavefront's serial sweep of its strip sets the kernel's duration, so
// shorter strips are tried as well.
constexpr long long kSmallCells = 4ll << 20;
constexpr int kSmallCandidates[] = {1, 2, 4, 8, 16};
constexpr int kReps = 3;

std::mutex g_mu;
bool g_on = true;
std::map<std::tuple<std::string, int, int, int, int>, int> g_table;  // (kind, device, NX, NZ, shots) -> bx

}  // namespace

int tuned_strip(const char* kind, int NX, int NZ, int nshots, hipStream_t stream, void (*trial)(void*, int),
                void* ctx) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  const auto key = std::make_tuple(std::string(kind), dev, NX, NZ, nshots);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_on) return kDefaultStrip;
    auto it = g_table.find(key);
    if (it != g_table.end()) return it->second;
  }
  std::vector<int> cand;
  if ((long long)NX * NZ * nshots < kSmallCells)
    cand.assign(std::begin(kSmallCandidates), std::end(kSmallCandidates));
  else
    cand.assign(std::begin(kCandidates), std::end(kCandidates));
  hipEvent_t a, b;
  if (hipEventCreate(&a) != hipSuccess) return kDefaultStrip;
  if (hipEventCreate(&b) != hipSuccess) {
    (void)hipEventDestroy(a);
    return kDefaultStrip;
  }
  int best = kDefaultStrip;
  float best_ms = 0.f;
  for (size_t ci = 0; ci < cand.size(); ci++) {
    const int bx = cand[ci];
    trial(ctx, bx);  // warm-up launch of this shape
    (void)hipEventRecord(a, stream);
    for (int r = 0; r < kReps; r++) trial(ctx, bx);
    (void)hipEventRecord(b, stream);
    float ms = 0.f;
    if (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&ms, a, b) != hipSuccess) {
      best = kDefaultStrip;
      break;
    }
    if (ci == 0 || ms < best_ms) {
      best = bx;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  std::lock_guard<std::mutex> lk(g_mu);
  g_table[key] = best;
  return best;
}

int strip_lookup(const char* kind, int NX, int NZ, int nshots) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_on) {
    auto it = g_table.find(std::make_tuple(std::string(kind), dev, NX, NZ, nshots));
    if (it != g_table.end()) return it->second;
  }
  return kDefaultStrip;
}

}  // namespace seis

using namespace seis;

extern "C" int seis_set_tuning(int on) {
  if (on != 0 && on != 1) return set_error("seis_set_tuning: on must be 0 or 1"), SEIS_ERR_ARG;
  std::lock_guard<std::mutex> lk(g_mu);
  g_on = on != 0;
  if (!g_on) g_table.clear();
  return SEIS_OK;
}

extern "C" int seis_tuned_strip(const char* kind, int nx_padded, int nz_padded, int nshots, int* bx) {
  if (!kind || !bx) return set_error("seis_tuned_strip: null argument"), SEIS_ERR_ARG;
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_table.find(std::make_tuple(std::string(kind), dev, nx_padded, nz_padded, nshots));
  *bx = it == g_table.end() ? 0 : it->second;
  return SEIS_OK;
}
