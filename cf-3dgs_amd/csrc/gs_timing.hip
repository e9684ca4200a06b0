// gs_timing.hip -- opt-in per-kernel timing with HIP events.
//
// When enabled, the launch helpers bracket each rasteriser stage with a pair
// of events recorded on the stream the stage is launched on; gs_timing_read
// waits for the recorded events and sums the elapsed times per stage.  Events
// come from a pool that grows only while timing is enabled; nothing is
// recorded (and nothing allocated) when it is off.
#include <mutex>
#include <string>
#include <vector>

#include "gs_common.hpp"
#include "gs_internal.hpp"

namespace gs {
namespace timing {

struct Rec {
  int stage;
  hipEvent_t a, b;
};

static std::mutex g_mu;
static bool g_on = false;
static std::vector<Rec> g_recs;
static std::vector<hipEvent_t> g_pool;

static hipEvent_t take() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

bool enabled() { return g_on; }

Scope::Scope(int stage, hipStream_t s) : stage_(stage), s_(s), a_(nullptr) {
  if (!g_on) return;
  // launches being captured into a graph are not timed (a recorded event
  // would become a graph node)
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e = take();
  if (e && hipEventRecord(e, s_) == hipSuccess) a_ = (void*)e;
}

Scope::~Scope() {
  if (!a_) return;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t b = take();
  if (!b) return;
  if (hipEventRecord(b, s_) != hipSuccess) {
    g_pool.push_back(b);
    g_pool.push_back((hipEvent_t)a_);
    return;
  }
  g_recs.push_back({stage_, (hipEvent_t)a_, b});
}

}  // namespace timing
}  // namespace gs

using namespace gs;

static const char* kStageNames[] = {"preprocess", "scan", "duplicate", "sort", "ranges", "blend",
                                    "render_bwd", "preprocess_bwd", "knn", "zero"};
static const int kNumStages = sizeof(kStageNames) / sizeof(kStageNames[0]);

extern "C" int gs_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(timing::g_mu);
  timing::g_on = on != 0;
  return GS_OK;
}

extern "C" int gs_timing_reset(void) {
  std::lock_guard<std::mutex> lk(timing::g_mu);
  for (auto& r : timing::g_recs) {
    (void)hipEventSynchronize(r.b);
    timing::g_pool.push_back(r.a);
    timing::g_pool.push_back(r.b);
  }
  timing::g_recs.clear();
  return GS_OK;
}

extern "C" int gs_timing_read(const char* stage, double* total_ms, int* launches) {
  if (!stage || !total_ms || !launches) return GS_E_ARG;
  int id = -1;
  for (int k = 0; k < kNumStages; ++k)
    if (std::string(kStageNames[k]) == stage) id = k;
  if (id < 0) return GS_E_ARG;
  std::lock_guard<std::mutex> lk(timing::g_mu);
  double t = 0.0;
  int n = 0;
  for (auto& r : timing::g_recs) {
    if (r.stage != id) continue;
    if (hipEventSynchronize(r.b) != hipSuccess) return GS_E_LAUNCH;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) return GS_E_LAUNCH;
    t += ms;
    ++n;
  }
  *total_ms = t;
  *launches = n;
  return GS_OK;
}
