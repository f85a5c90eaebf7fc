// hk_engine.h -- the wavefront scheduler behind hk_render, hk_render_batch,
// hk_render_pixel_samples and the hk_server request loop.
//
// Work model.  A request is cut into chunks of `chunk` samples per pixel
// (chunk = min(spp, max_samples / pixels), a function of the request and the
// context's sample capacity only).  A chunk is admitted when an item slot and
// a contiguous range of the sample buffer are free; every bounce, k_regen
// starts as many not-yet-started samples of the admitted chunks (in
// admission order) as there are free places in the extension queue.  Paths
// count themselves finished per chunk; a chunk whose samples are all done is
// resolved into its request's film, and a request's chunks are resolved in
// chunk order.  So a request's film does not depend on what else runs: alone,
// batched or admitted mid-flight, it accumulates the same values in the same
// order.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "hk_kernels.h"
#include "hk_runtime.h"

namespace hk {

// One render request inside an engine.
struct RequestRec {
  uint64_t id = 0;
  DevRequest dev;
  uint32_t npix = 0, spp = 0, spp_offset = 0, chunk = 1, n_chunks = 0;
  uint32_t slot = 0;               // request slot of the device table
  DBuf accum;                      // npix x float4 (rgb * w, w)
  float *out_rgb = nullptr;        // host film (height x width x 3) or null
  bool keep_samples = false;       // pixel-sample mode: keep per-sample radiance
  std::vector<float> samples;      // keep_samples: n x 4 radiance (rgb, unused)
  uint32_t next_admit = 0, next_resolve = 0;
  uint64_t started = 0;            // samples handed to k_regen so far
  uint64_t left() const { return cancelled ? 0 : uint64_t(npix) * spp - started; }
  std::vector<int32_t> chunk_item;  // item slot per admitted, unresolved chunk (-1 otherwise)
  double t_submit = 0, t_admit = -1, t_done = -1;  // ms on the engine clock
  bool done = false;
  bool cancelled = false;          // Engine::cancel: no film is produced
  uint64_t invalid_samples = 0;    // samples the resolve dropped (set at finalize)
  uint64_t film_bytes() const { return uint64_t(npix) * 16; }
};

// An admitted chunk (one item slot).
struct ChunkRec {
  RequestRec *req = nullptr;
  uint32_t index = 0;     // chunk index inside the request
  uint32_t size = 0;      // samples = npix * k (a cancelled chunk: the samples it started)
  uint32_t started = 0;   // samples handed to k_regen so far
  uint32_t buf_base = 0;  // first sample-buffer entry
  uint32_t buf_len = 0;   // sample-buffer entries held (npix * k, whatever size becomes)
  bool complete = false;
};

// Validates and sets up a request on scene s (film, camera, filter, chunk
// size from the sample capacity chunk_cap); does not start it.
std::unique_ptr<RequestRec> prepare_request(hk_scene *s, const hk_request &req, float *out_rgb, uint32_t chunk_cap);

class Engine {
 public:
  // ws: device/pinned buffers (owned by the caller, sized here); st: the
  // stream every launch of this engine goes to.
  Engine(hk_scene *s, Workspace &ws, hipStream_t st, uint32_t pool_paths, uint32_t buffer_samples,
         uint32_t chunk_cap, bool profile, hipEvent_t ev0, hipEvent_t ev1);
  ~Engine();

  // Hands a prepared request to the scheduler (FIFO admission).  Throws
  // without changing the engine when the request cannot be taken (no free
  // request slot, a chunk larger than the sample buffer, no memory for its
  // film).
  void submit(RequestRec *r);
  // Request slots still free (submit needs one per request in flight).
  uint32_t free_request_slots() const { return uint32_t(free_reqs_.size()); }
  uint32_t requests_in_flight() const { return kMaxRequests - uint32_t(free_reqs_.size()); }
  // Stops a submitted request: its unadmitted chunks are dropped and its
  // admitted ones shrink to the samples already started, so they finish with
  // the paths in flight and free their slots, buffer ranges and film without
  // a resolve.  Returns the requests finalized by the call (r itself when
  // nothing of it is in flight); the rest come out of later collect() calls,
  // marked cancelled.  No effect on a finished request.
  std::vector<RequestRec *> cancel(RequestRec *r);
  // Non-blocking progress: a launched bounce whose work has finished on the
  // device (collect() then returns without waiting).
  bool launched() const { return launched_; }
  bool ready() const { return launched_ && hipStreamQuery(st_) == hipSuccess; }
  // Samples of submitted requests not yet handed to k_regen, and whether
  // this engine's next bounces keep at least half its wavefront busy.
  uint64_t unstarted() const;
  bool saturating() const { return busy() && 2 * (uint64_t(in_flight_) + unstarted()) >= pool_; }
  // Submit time (engine clock) of the oldest request in flight; +inf if none.
  double oldest_submit() const;
  // Admission / start order: true the fewest samples left first, false
  // first come first served.
  void set_shortest_first(bool v) { shortest_first_ = v; }
  bool busy() const { return !pending_.empty() || !active_.empty() || in_flight_ > 0; }
  // One bounce: admit, start new samples, extend, shade, shadow, collect
  // finished chunks, resolve them, finalize finished requests.  Returns the
  // requests that finished during this bounce.  launch() issues the bounce's
  // work without waiting (false: nothing to run); collect() waits for it and
  // does the rest, so several engines on their own streams overlap.
  std::vector<RequestRec *> bounce();
  bool launch();
  std::vector<RequestRec *> collect();
  // Bounces until every submitted request has finished.
  void run();
  // Totals since the last reset_stats (device counters are read by take_stats).
  void reset_stats();
  void take_stats(hk_render_stats *st);
  double now_ms() const;

 private:
  void admit();
  bool admit_chunks(RequestRec *r);
  bool alloc_range(uint32_t n, uint32_t &base);
  void free_range(uint32_t base, uint32_t n);
  void resolve_ready(std::vector<RequestRec *> &finished);
  void finalize(RequestRec *r);
  RenderArgs args() const;
  void enqueue_bounce(int qi, bool graph, uint32_t n_plan, uint32_t paths);
  void capture(int qi);
  static constexpr uint64_t kGraphAfter = 4;   // eager bounces before capturing

  hk_scene *s_;
  Workspace &ws_;
  hipStream_t st_;
  uint32_t pool_, cap_samples_, chunk_cap_;
  bool profile_;
  hipEvent_t ev0_, ev1_;
  std::deque<RequestRec *> pending_;            // submitted, chunks left to admit (submission order)
  bool shortest_first_ = false;                 // schedule: fewest samples left first (else FIFO)
  std::vector<ChunkRec> slots_;                 // by item slot
  std::vector<uint32_t> free_items_, free_reqs_;
  std::vector<uint32_t> active_;                // admitted item slots, admission order
  std::map<uint32_t, uint32_t> free_ranges_;    // sample buffer: start -> length
  std::vector<RequestRec *> reqs_by_slot_;
  uint32_t in_flight_ = 0;                      // paths in the current extension queue
  int qi_ = 0;
  bool launched_ = false;
  // statistics
  uint64_t paths_ = 0, bounces_ = 0, trace_launches_ = 0, graph_bounces_ = 0;
  double trace_ms_ = 0;
  hipEvent_t ev_start_ = nullptr, ev_end_ = nullptr;
  bool timing_ = false;
  bool use_graph_ = true;
  hipGraphExec_t gexec_[2] = {nullptr, nullptr};  // the bounce, captured per queue parity
};

}  // namespace hk
