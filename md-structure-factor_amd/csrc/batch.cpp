// Batches of independent systems (mixed grid sizes) run as one HIP graph.
//
// Each system keeps its own plan (grid, types, buffers, FFT choices picked by
// the autotuner).  The first run after the frames change captures every
// system's frame loop -- S(q) reset, then per frame prepare / deposit / three
// FFT passes with the |F|^2 accumulation -- into one graph, systems spread
// over forked capture streams so that the graph has one independent branch
// per stream; later runs replay the instantiated graph with a single launch.
// Coordinates are read, never written (no in-place rescale / wrap), so every
// replay computes the same S(q) from the same inputs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mdsf_internal.h"

namespace mdsf {
namespace {

// Capture streams (graph branches).  Branches run concurrently: on the
// 32-system workload 1 / 4 / 8 / 16 / 32 branches gave 895 / 1352 / 1608 /
// 1683 / 1718 system-frames/s.
constexpr int kMaxBatchStreams = 16;

struct SysFrames {
  const void* coords = nullptr;
  int dtype = MDSF_F32;
  int64_t nframes = 0;
  const int32_t* types = nullptr;
  const double* scale = nullptr;
  double box[3] = {0, 0, 0};
  int arith = MDSF_F32;
  int64_t wrap_lo = 0, wrap_hi = 0;
};

struct Batch {
  std::vector<mdsf_plan*> plans;
  std::vector<double> cost;       // estimated work per frame (voxel-update units)
  std::vector<int> stream_of;     // capture stream of each system
  std::vector<SysFrames> frames;
  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;  // [0] fork, [1..] joins
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool dirty = true;
};

void drop_graph(Batch& b) {
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
}

void free_batch(Batch* b) {
  if (!b) return;
  drop_graph(*b);
  for (hipEvent_t e : b->events) (void)hipEventDestroy(e);
  for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
  for (mdsf_plan* p : b->plans) mdsf_plan_destroy(p);
  delete b;
}

// Enqueue system `i`'s whole frame loop on stream s (captured, not run).
hipError_t enqueue_system(Batch& b, int i, hipStream_t s) {
  Plan& p = *reinterpret_cast<Plan*>(b.plans[i]);
  const SysFrames& f = b.frames[i];
  const long long nxy = (long long)p.sys.n[0] * p.sys.n[1];
  hipError_t e = launch_fill_zero(p.sf_main, nxy * (p.sys.n[2] / 2), s);
  if (e == hipSuccess) e = launch_fill_zero(p.sf_nyq, nxy, s);
  const size_t esz = f.dtype == MDSF_F32 ? 4 : 8;
  const size_t fstride = (size_t)p.sys.natoms * 3 * esz;
  for (int64_t t = 0; t < f.nframes && e == hipSuccess; ++t) {
    void* c = const_cast<char*>(static_cast<const char*>(f.coords)) + (size_t)t * fstride;
    e = launch_prep(p, c, f.dtype, f.arith, f.types, f.scale + 3 * t, f.box, f.wrap_lo, f.wrap_hi,
                    nullptr, nullptr, s, false, /*write_back=*/false);
    if (e == hipSuccess) e = launch_binning(p, s);
    if (e == hipSuccess) e = launch_deposit(p, p.rho, s);
    if (e == hipSuccess) e = launch_parked(p, p.rho, s);
    if (e == hipSuccess) e = fft_forward(p, p.rho, true, s);
  }
  return e;
}

// Longest-first assignment of systems to the capture streams (each stream is
// one graph branch): the heaviest remaining system goes to the least loaded
// stream, so that the branches finish close together.
void assign_streams(Batch& b) {
  const int ns = (int)b.streams.size();
  std::vector<int> order(b.plans.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::vector<double> work(b.plans.size());
  for (size_t i = 0; i < order.size(); ++i) work[i] = b.cost[i] * (double)b.frames[i].nframes;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return work[x] > work[y]; });
  std::vector<double> load(ns, 0.0);
  b.stream_of.assign(b.plans.size(), 0);
  for (int i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    b.stream_of[i] = k;
    load[k] += work[i];
  }
}

int capture(Batch& b) {
  drop_graph(b);
  assign_streams(b);
  const int ns = (int)b.streams.size();
  hipStream_t s0 = b.streams[0];
  MDSF_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
  hipError_t e = hipEventRecord(b.events[0], s0);
  for (int k = 1; k < ns && e == hipSuccess; ++k) e = hipStreamWaitEvent(b.streams[k], b.events[0], 0);
  for (size_t i = 0; i < b.plans.size() && e == hipSuccess; ++i)
    e = enqueue_system(b, (int)i, b.streams[b.stream_of[i]]);
  for (int k = 1; k < ns && e == hipSuccess; ++k) {
    e = hipEventRecord(b.events[k], b.streams[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s0, b.events[k], 0);
  }
  hipGraph_t g = nullptr;
  const hipError_t ee = hipStreamEndCapture(s0, &g);
  if (e == hipSuccess) e = ee;
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return hip_fail(e, "batch graph capture");
  }
  b.graph = g;
  e = hipGraphInstantiate(&b.exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    drop_graph(b);
    return hip_fail(e, "batch graph instantiation");
  }
  b.dirty = false;
  return MDSF_OK;
}

}  // namespace
}  // namespace mdsf

using namespace mdsf;

extern "C" {

int mdsf_batch_create(const mdsf_grid_desc* descs, int32_t nsys, mdsf_batch** out) {
  if (!descs || !out || nsys <= 0) return fail(MDSF_ERR_ARG, "need at least one system");
  *out = nullptr;
  Batch* b = new Batch();
  for (int i = 0; i < nsys; ++i) {
    mdsf_grid_desc d = descs[i];
    d.flags |= MDSF_AUTOTUNE;
    mdsf_plan* p = nullptr;
    const int rc = mdsf_plan_create(&d, &p);
    if (rc != MDSF_OK) {
      const std::string msg = mdsf_last_error();
      free_batch(b);
      return fail(rc, "system " + std::to_string(i) + ": " + msg);
    }
    b->plans.push_back(p);
    // per-frame work estimate: voxel updates of the deposition plus the
    // FFT / prepare passes as ~64 update-equivalents per grid point
    double c = 64.0 * (double)d.n[0] * d.n[1] * d.n[2];
    for (int t = 0; t < d.ntypes; ++t)
      c += (double)d.types[t].count * 8.0 * d.types[t].half[0] * d.types[t].half[1] *
           d.types[t].half[2];
    b->cost.push_back(c);
  }
  b->frames.resize(nsys);
  const int ns = std::min<int>(nsys, kMaxBatchStreams);
  for (int k = 0; k < ns; ++k) {
    hipStream_t s = nullptr;
    hipEvent_t e = nullptr;
    hipError_t err = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (err == hipSuccess) b->streams.push_back(s);
    if (err == hipSuccess) err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (err == hipSuccess) b->events.push_back(e);
    if (err != hipSuccess) {
      free_batch(b);
      return hip_fail(err, "batch streams");
    }
  }
  *out = reinterpret_cast<mdsf_batch*>(b);
  return MDSF_OK;
}

int mdsf_batch_destroy(mdsf_batch* batch) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b) return MDSF_OK;
  (void)hipDeviceSynchronize();
  free_batch(b);
  return MDSF_OK;
}

mdsf_plan* mdsf_batch_plan(mdsf_batch* batch, int32_t sys) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b || sys < 0 || sys >= (int)b->plans.size()) {
    set_error("batch system index out of range");
    return nullptr;
  }
  return b->plans[sys];
}

int mdsf_batch_set_frames(mdsf_batch* batch, int32_t sys, const void* d_coords, int dtype,
                          int64_t nframes, const int32_t* d_types, const double* d_scale,
                          const double box[3], int arith, int64_t wrap_lo, int64_t wrap_hi) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b || sys < 0 || sys >= (int)b->plans.size())
    return fail(MDSF_ERR_ARG, "batch system index out of range");
  const Plan& p = *reinterpret_cast<Plan*>(b->plans[sys]);
  if (nframes < 0) return fail(MDSF_ERR_ARG, "negative frame count");
  if (nframes > 0 && p.sys.natoms > 0 && (!d_coords || !d_types || !d_scale))
    return fail(MDSF_ERR_ARG, "null coordinate / type / scale pointer");
  if (!box) return fail(MDSF_ERR_ARG, "null box");
  if (dtype != MDSF_F32 && dtype != MDSF_F64) return fail(MDSF_ERR_ARG, "dtype must be F32 or F64");
  if (arith != MDSF_F32 && arith != MDSF_F64) return fail(MDSF_ERR_ARG, "arith must be F32 or F64");
  if (dtype == MDSF_F64 && arith == MDSF_F32)
    return fail(MDSF_ERR_ARG, "float64 coordinates are rescaled in float64");
  SysFrames& f = b->frames[sys];
  f.coords = d_coords;
  f.dtype = dtype;
  f.nframes = nframes;
  f.types = d_types;
  f.scale = d_scale;
  for (int d = 0; d < 3; ++d) f.box[d] = box[d];
  f.arith = arith;
  f.wrap_lo = wrap_lo;
  f.wrap_hi = wrap_hi;
  b->dirty = true;
  return MDSF_OK;
}

int mdsf_batch_run(mdsf_batch* batch, mdsf_stream stream) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b) return fail(MDSF_ERR_ARG, "null batch");
  if (b->dirty) {
    // capture streams start after everything already queued on `stream`
    MDSF_HIP(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)));
    const int rc = capture(*b);
    if (rc != MDSF_OK) return rc;
  }
  MDSF_HIP(hipGraphLaunch(b->exec, reinterpret_cast<hipStream_t>(stream)));
  return MDSF_OK;
}

int mdsf_batch_check_errors(mdsf_batch* batch, mdsf_stream stream) {
  Batch* b = reinterpret_cast<Batch*>(batch);
  if (!b) return fail(MDSF_ERR_ARG, "null batch");
  for (size_t i = 0; i < b->plans.size(); ++i) {
    const int rc = mdsf_check_errors(b->plans[i], stream);
    if (rc != MDSF_OK) return fail(rc, "system " + std::to_string(i) + ": " + mdsf_last_error());
  }
  return MDSF_OK;
}

}  // extern "C"
