// Batched NSSurfacelet transforms of many volumes of uneven sizes behind the
// C ABI (surfhip_ns_batch_*).
//
// The reference transforms one volume per call (src/transforms/ns_surfacelet.jl:
// 551-581) and has no device batching (:147).  Here a batch holds one plan per
// (lane, volume size); the volumes are spread over `lanes` HIP streams,
// longest first onto the least-loaded lane, so transforms of different
// volumes overlap where one volume's kernels leave the GPU idle.  A run is
// recorded once as a hipGraph by stream capture on the batch's own origin
// stream (every lane forked from it by one event and joined back by one event
// per lane, each recorded once per capture) and replayed with a single
// hipGraphLaunch while the data pointers, the mode and the caller's stream
// repeat.  The launch sequences hold no host synchronisation and no memset,
// and under capture each plan keeps its work on its lane (no second stream),
// so a replay computes exactly what direct launches compute.  A caller's
// stream that is itself being captured gets direct launches, which then join
// the caller's capture.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <memory>
#include <vector>

#include "hip_util.hpp"
#include "ns_plan_api.hpp"

namespace surfhip {
namespace {

void check_status(int st) {
  if (st == SURFHIP_OK) return;
  char buf[1024];
  surfhip_last_error(buf, sizeof buf);
  fail(st, buf);
}

}  // namespace
}  // namespace surfhip

using namespace surfhip;

struct surfhip_ns_batch {
  int nvol = 0, nlanes = 0, dtype = SURFHIP_F32;
  bool graphs = true;
  hipStream_t caller = nullptr, origin = nullptr;
  std::vector<std::array<int, 3>> dims;
  std::vector<int> lane_of;
  std::vector<hipStream_t> lanes;
  std::vector<surfhip_ns_plan *> plans;      // owned, one per (lane, size)
  std::vector<surfhip_ns_plan *> plan_of;    // per volume
  hipEvent_t fork = nullptr, done = nullptr, start = nullptr;
  std::vector<hipEvent_t> join;
  // The recorded run: replayed while mode and pointers repeat.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int rec_mode = -1;
  std::vector<const void *> rec_ptrs;

  ~surfhip_ns_batch() {
    if (origin) (void)hipStreamSynchronize(origin);
    for (hipStream_t s : lanes) (void)hipStreamSynchronize(s);
    drop_graph();
    for (surfhip_ns_plan *p : plans) (void)surfhip_ns_plan_destroy(p);
    for (hipStream_t s : lanes) (void)hipStreamDestroy(s);
    for (hipEvent_t e : join) (void)hipEventDestroy(e);
    if (fork) (void)hipEventDestroy(fork);
    if (done) (void)hipEventDestroy(done);
    if (start) (void)hipEventDestroy(start);
    if (origin) (void)hipStreamDestroy(origin);
  }
  // An executable graph is destroyed only once its last replay (followed by
  // `done`) has finished, never while it may still run.
  void drop_graph() {
    if (exec && done) (void)hipEventSynchronize(done);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    rec_mode = -1;
    rec_ptrs.clear();
  }

  // Every volume's launches, each on its lane, forked from and joined back to
  // `s` (the origin stream under capture, else the caller's stream).
  void launch_all(hipStream_t s, int mode, const void *const *xs, void *const *cs, void *const *ys) {
    SH_HIP(hipEventRecord(fork, s));
    for (hipStream_t l : lanes) SH_HIP(hipStreamWaitEvent(l, fork, 0));
    for (int k = 0; k < nvol; ++k) {
      surfhip_ns_plan *p = plan_of[k];
      if (mode == 0 || mode == 2) ns_analyze_raw(p, xs[k], cs[k]);
      if (mode == 1 || mode == 2) ns_synthesize_raw(p, cs[k], ys[k]);
    }
    for (int i = 0; i < nlanes; ++i) {
      SH_HIP(hipEventRecord(join[i], lanes[i]));
      SH_HIP(hipStreamWaitEvent(s, join[i], 0));
    }
  }
};

namespace {

bool capturing(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return st != hipStreamCaptureStatusNone;
}

void create_batch(surfhip_ns_batch **out, const surfhip_ns_params *p, int32_t nvol, const int32_t *dims,
                  int32_t dtype, int32_t lanes, int32_t flags, void *stream) {
  if (!out || !p || !dims) fail(SURFHIP_ERR_ARG, "null argument");
  *out = nullptr;
  if (nvol < 1) fail(SURFHIP_ERR_ARG, "a batch needs at least one volume");
  if (lanes < 1) fail(SURFHIP_ERR_ARG, fmt("lanes must be ≥ 1, got %d", lanes));
  if (flags & ~(SURFHIP_BATCH_TUNE | SURFHIP_BATCH_NO_GRAPH)) fail(SURFHIP_ERR_ARG, fmt("unknown flags %d", flags));
  auto B = std::make_unique<surfhip_ns_batch>();
  B->nvol = nvol;
  B->nlanes = std::min(lanes, nvol);
  B->dtype = dtype;
  B->graphs = !(flags & SURFHIP_BATCH_NO_GRAPH);
  B->caller = (hipStream_t)stream;
  for (int k = 0; k < nvol; ++k) {
    std::array<int, 3> d{dims[3 * k], dims[3 * k + 1], dims[3 * k + 2]};
    if (d[0] < 1 || d[1] < 1 || d[2] < 1)
      fail(SURFHIP_ERR_ARG, fmt("volume %d: size must be positive, got (%d, %d, %d)", k, d[0], d[1], d[2]));
    B->dims.push_back(d);
  }
  // Longest-processing-time assignment: biggest volumes first, each onto the
  // lane with the fewest voxels so far (ties: the lowest lane).
  std::vector<int> order(nvol);
  for (int k = 0; k < nvol; ++k) order[k] = k;
  auto vox = [&](int k) { return (int64_t)B->dims[k][0] * B->dims[k][1] * B->dims[k][2]; };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return vox(a) > vox(b); });
  std::vector<int64_t> load(B->nlanes, 0);
  B->lane_of.assign(nvol, 0);
  for (int k : order) {
    const int l = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    B->lane_of[k] = l;
    load[l] += vox(k);
  }
  // Streams at the default priority.
  SH_HIP(hipStreamCreateWithFlags(&B->origin, hipStreamNonBlocking));
  B->lanes.assign(B->nlanes, nullptr);
  for (int i = 0; i < B->nlanes; ++i) SH_HIP(hipStreamCreateWithFlags(&B->lanes[i], hipStreamNonBlocking));
  SH_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
  SH_HIP(hipEventCreateWithFlags(&B->done, hipEventDisableTiming));
  SH_HIP(hipEventCreateWithFlags(&B->start, hipEventDisableTiming));
  B->join.assign(B->nlanes, nullptr);
  for (int i = 0; i < B->nlanes; ++i) SH_HIP(hipEventCreateWithFlags(&B->join[i], hipEventDisableTiming));
  std::map<std::pair<int, std::array<int, 3>>, surfhip_ns_plan *> by_key;
  B->plan_of.assign(nvol, nullptr);
  for (int k = 0; k < nvol; ++k) {
    const auto key = std::make_pair(B->lane_of[k], B->dims[k]);
    auto it = by_key.find(key);
    if (it == by_key.end()) {
      surfhip_ns_plan *pl = nullptr;
      const auto &d = B->dims[k];
      check_status(surfhip_ns_plan_create(&pl, p, d[0], d[1], d[2], dtype, B->lanes[key.first]));
      B->plans.push_back(pl);
      if (flags & SURFHIP_BATCH_TUNE) check_status(surfhip_ns_plan_tune(pl, nullptr));
      it = by_key.emplace(key, pl).first;
    }
    B->plan_of[k] = it->second;
  }
  for (hipStream_t s : B->lanes) SH_HIP(hipStreamSynchronize(s));
  *out = B.release();
}

void run_batch(surfhip_ns_batch *b, int32_t mode, const void *const *xs, void *const *cs, void *const *ys) {
  if (!b) fail(SURFHIP_ERR_ARG, "null batch");
  if (mode < 0 || mode > 2) fail(SURFHIP_ERR_ARG, fmt("mode must be 0 (analyze), 1 (synthesize) or 2 (both), got %d", mode));
  if (!cs || ((mode == 0 || mode == 2) && !xs) || ((mode == 1 || mode == 2) && !ys))
    fail(SURFHIP_ERR_ARG, "null argument");
  // Every caller buffer is checked before anything is launched or recorded.
  std::vector<const void *> ptrs;
  for (int k = 0; k < b->nvol; ++k) {
    surfhip_ns_plan *p = b->plan_of[k];
    const size_t vb = ns_volume_bytes(p), cb = ns_coefficient_bytes(p);
    if (mode != 1) {
      if (!xs[k]) fail(SURFHIP_ERR_ARG, fmt("volume %d: null pointer", k));
      check_device_buffer(xs[k], vb, fmt("volume %d", k).c_str());
    }
    if (!cs[k]) fail(SURFHIP_ERR_ARG, fmt("coefficients %d: null pointer", k));
    check_device_buffer(cs[k], cb, fmt("coefficient array %d", k).c_str());
    if (mode != 0) {
      if (!ys[k]) fail(SURFHIP_ERR_ARG, fmt("output %d: null pointer", k));
      check_device_buffer(ys[k], vb, fmt("output %d", k).c_str());
    }
    ptrs.push_back(mode != 1 ? xs[k] : nullptr);
    ptrs.push_back(cs[k]);
    ptrs.push_back(mode != 0 ? ys[k] : nullptr);
  }
  if (!b->graphs || capturing(b->caller)) {
    b->launch_all(b->caller, mode, xs, cs, ys);
    return;
  }
  if (!(b->exec && b->rec_mode == mode && b->rec_ptrs == ptrs)) {
    b->drop_graph();
    SH_HIP(hipStreamBeginCapture(b->origin, hipStreamCaptureModeThreadLocal));
    try {
      b->launch_all(b->origin, mode, xs, cs, ys);
    } catch (...) {
      hipGraph_t g = nullptr;
      (void)hipStreamEndCapture(b->origin, &g);
      if (g) (void)hipGraphDestroy(g);
      throw;
    }
    SH_HIP(hipStreamEndCapture(b->origin, &b->graph));
    SH_HIP(hipGraphInstantiateWithFlags(&b->exec, b->graph, 0));
    b->rec_mode = mode;
    b->rec_ptrs = ptrs;
  }
  // Ordered after the caller's earlier work and before its later work.
  SH_HIP(hipEventRecord(b->start, b->caller));
  SH_HIP(hipStreamWaitEvent(b->origin, b->start, 0));
  SH_HIP(hipGraphLaunch(b->exec, b->origin));
  SH_HIP(hipEventRecord(b->done, b->origin));
  SH_HIP(hipStreamWaitEvent(b->caller, b->done, 0));
}

}  // namespace

extern "C" {

int surfhip_ns_batch_create(surfhip_ns_batch **out, const surfhip_ns_params *p, int32_t nvol, const int32_t *dims,
                            int32_t dtype, int32_t lanes, int32_t flags, void *stream) {
  return guarded([&] { create_batch(out, p, nvol, dims, dtype, lanes, flags, stream); });
}

int surfhip_ns_batch_run(surfhip_ns_batch *b, int32_t mode, const void *const *xs, void *const *cs,
                         void *const *ys) {
  return guarded([&] { run_batch(b, mode, xs, cs, ys); });
}

int surfhip_ns_batch_set_stream(surfhip_ns_batch *b, void *stream) {
  return guarded([&] {
    if (!b) fail(SURFHIP_ERR_ARG, "null batch");
    b->caller = (hipStream_t)stream;
  });
}

int surfhip_ns_batch_info(surfhip_ns_batch *b, int32_t *lane_of, int32_t *graph_recorded) {
  return guarded([&] {
    if (!b) fail(SURFHIP_ERR_ARG, "null batch");
    if (lane_of) std::copy(b->lane_of.begin(), b->lane_of.end(), lane_of);
    if (graph_recorded) *graph_recorded = b->exec ? 1 : 0;
  });
}

int surfhip_ns_batch_destroy(surfhip_ns_batch *b) {
  return guarded([&] { delete b; });
}

}  // extern "C"
