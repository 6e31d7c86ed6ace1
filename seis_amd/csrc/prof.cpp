// Kernel timing store behind seis_profile_start / seis_profile_stop.
#include "prof.h"

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "seis_internal.h"

namespace seis {
namespace {

struct Rec {
  const char* name;
  double bytes, cells;
  hipEvent_t a, b;
};

std::atomic<bool> g_on{false};
std::mutex g_mu;
std::vector<Rec> g_recs;  // records whose events have not been read yet
std::vector<hipEvent_t> g_pool;
std::map<std::string, seis_kernel_stat> g_acc;
std::vector<std::string> g_order;

// At most this many launches keep their event pair alive; older ones are
// folded into the totals (waiting for the oldest launch while the newer ones
// keep the stream busy).
constexpr size_t kMaxLive = 1024;

void fold(Rec& r) {
  float ms = 0.f;
  (void)hipEventSynchronize(r.b);
  (void)hipEventElapsedTime(&ms, r.a, r.b);
  auto it = g_acc.find(r.name);
  if (it == g_acc.end()) {
    seis_kernel_stat s{};
    std::strncpy(s.name, r.name, sizeof(s.name) - 1);
    it = g_acc.emplace(r.name, s).first;
    g_order.push_back(r.name);
  }
  it->second.launches += 1;
  it->second.total_ms += ms;
  it->second.alg_bytes += r.bytes;
  it->second.cell_updates += r.cells;
  g_pool.push_back(r.a);
  g_pool.push_back(r.b);
}

void fold_all() {
  for (Rec& r : g_recs) fold(r);
  g_recs.clear();
}

}  // namespace

bool prof_enabled() { return g_on.load(std::memory_order_relaxed); }

void prof_take_events(hipEvent_t* a, hipEvent_t* b) {
  std::lock_guard<std::mutex> lk(g_mu);
  for (hipEvent_t* e : {a, b}) {
    if (!g_pool.empty()) {
      *e = g_pool.back();
      g_pool.pop_back();
    } else {
      (void)hipEventCreate(e);
    }
  }
}

void prof_record(const char* name, double alg_bytes, double cell_updates, hipEvent_t a, hipEvent_t b) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_recs.push_back(Rec{name, alg_bytes, cell_updates, a, b});
  if (g_recs.size() >= kMaxLive) {
    const size_t n = kMaxLive / 2;
    for (size_t i = 0; i < n; i++) fold(g_recs[i]);
    g_recs.erase(g_recs.begin(), g_recs.begin() + n);
  }
}

}  // namespace seis

using namespace seis;

extern "C" {

int seis_profile_start(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  fold_all();
  g_acc.clear();
  g_order.clear();
  g_on = true;
  return SEIS_OK;
}

int seis_profile_stop(seis_kernel_stat* stats, int max_stats, int* n_stats) {
  g_on = false;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    set_error(std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
    return SEIS_ERR_HIP;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  fold_all();
  int n = 0;
  for (const std::string& k : g_order) {
    if (stats && n < max_stats) stats[n] = g_acc[k];
    n++;
  }
  if (n_stats) *n_stats = n;
  return SEIS_OK;
}

}  // extern "C"
