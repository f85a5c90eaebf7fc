// hk_engine.cpp -- the wavefront scheduler (see hk_engine.h for the model).
// Per bounce the host launches k_regen (new samples from the plan), k_extend,
// k_shade and k_shadow on the engine's stream, then reads back the queue
// counters and the per-chunk finished counts (one 128 KiB copy: eight partial
// counts per chunk) to plan the next bounce and to resolve finished chunks.
#include "hk_engine.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

namespace hk {

// ------------------------------------------------------------------ pinned host memory
void PinnedBuf::alloc(size_t n) {
  release();
  if (n == 0) return;
  hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
  bytes = n;
}
void PinnedBuf::release() {
  if (p) (void)hipHostFree(p);
  p = nullptr;
  bytes = 0;
}

// ------------------------------------------------------------------ workspace
void Workspace::ensure(uint32_t paths, uint32_t emitters, uint32_t samples) {
  emitters = std::max(1u, emitters);
  if (!tables) {
    reqs.alloc(sizeof(DevRequest) * kMaxRequests);
    items.alloc(sizeof(WorkItem) * kMaxItems);
    plan.alloc(sizeof(RegenEntry) * kMaxItems);
    item_done.alloc(sizeof(uint32_t) * kMaxItems * kDoneWays);
    hdr.alloc(sizeof(BounceHdr));
    counters.alloc(sizeof(uint32_t) * kCounterWords);
    stats.alloc(64);
    req_invalid.alloc(sizeof(unsigned long long) * kMaxRequests);
    h_reqs.alloc(sizeof(DevRequest) * kMaxRequests);
    h_req_invalid.alloc(sizeof(unsigned long long) * kMaxRequests);
    h_items.alloc(sizeof(WorkItem) * kMaxItems);
    h_plan.alloc(sizeof(RegenEntry) * kMaxItems);
    h_done.alloc(sizeof(uint32_t) * kMaxItems * kDoneWays);
    h_counters.alloc(sizeof(uint32_t) * (ctr(1) + 1));
    h_hdr.alloc(sizeof(BounceHdr));
    tables = true;
  }
  if (samples > cap_samples) {
    sample_L.alloc(size_t(samples) * 16);
    cap_samples = samples;
  }
  if (paths > cap_paths) {
    size_t P = paths;
    for (int i = 0; i < 2; i++) {
      q_o[i].alloc(P * 16); q_d[i].alloc(P * 16); q_beta[i].alloc(P * 16); q_L[i].alloc(P * 16);
      q_aux[i].alloc(P * 16); q_meta[i].alloc(P * 4);
    }
    hit.alloc(P * 16);
    s_ent.alloc(P * 16);
    cap_paths = paths;
  }
  // shadow slot (e, i) at e x paths + i; mask word w >= 1 of path i at
  // (w - 1) x paths + i
  size_t slots = size_t(paths) * emitters;
  if (slots > cap_slots) {
    sh_o.alloc(slots * 16); sh_d.alloc(slots * 16); sh_c.alloc(slots * 16);
    cap_slots = slots;
  }
  size_t words = size_t(paths) * ((emitters + 31) / 32 - 1);
  if (words > cap_words) {
    sh_mask.alloc(words * 4);
    cap_words = words;
  }
}

uint32_t paths_for_emitters(uint32_t paths, uint32_t emitters) {
  uint64_t cap = kShadowSlotBytes / (uint64_t(48) * std::max(1u, emitters));
  return uint32_t(std::max<uint64_t>(1, std::min<uint64_t>(paths, cap)));
}

uint64_t Workspace::device_bytes() const {
  const DBuf *all[] = {&sample_L, &q_o[0], &q_o[1], &q_d[0], &q_d[1], &q_beta[0], &q_beta[1], &q_L[0], &q_L[1],
                       &q_aux[0], &q_aux[1], &q_meta[0], &q_meta[1], &hit, &s_ent, &sh_o, &sh_d, &sh_c,
                       &sh_mask, &counters, &stats, &reqs, &items, &plan, &item_done, &hdr, &req_invalid};
  uint64_t b = 0;
  for (const DBuf *x : all) b += x->bytes;
  return b;
}

void Workspace::release() {
  DBuf *all[] = {&sample_L, &q_o[0], &q_o[1], &q_d[0], &q_d[1], &q_beta[0], &q_beta[1], &q_L[0], &q_L[1],
                 &q_aux[0], &q_aux[1], &q_meta[0], &q_meta[1], &hit, &s_ent, &sh_o, &sh_d, &sh_c, &sh_mask, &counters,
                 &stats, &reqs, &items, &plan, &item_done, &hdr, &req_invalid};
  for (DBuf *b : all) b->release();
  PinnedBuf *pin[] = {&h_reqs, &h_items, &h_plan, &h_done, &h_counters, &h_hdr, &h_req_invalid};
  for (PinnedBuf *b : pin) b->release();
  cap_paths = cap_samples = 0;
  cap_slots = cap_words = 0;
  tables = false;
}

// Queue kernels get one lane per item (optionally capped by the "grid" option).
static uint32_t kernel_grid(uint32_t count, uint32_t cap) {
  uint32_t g = (count + 255) / 256;
  if (cap) g = std::min(g, cap);
  return std::max(1u, g);
}

// ------------------------------------------------------------------ engine
Engine::Engine(hk_scene *s, Workspace &ws, hipStream_t st, uint32_t pool_paths, uint32_t buffer_samples,
               uint32_t chunk_cap, bool profile, hipEvent_t ev0, hipEvent_t ev1)
    : s_(s), ws_(ws), st_(st), pool_(paths_for_emitters(std::max(1u, pool_paths), uint32_t(std::max(1, s->dev.n_emitters)))),
      cap_samples_(std::max(1u, buffer_samples)),
      chunk_cap_(std::max(1u, chunk_cap)), profile_(profile), ev0_(ev0), ev1_(ev1) {
  ws_.ensure(pool_, uint32_t(s->dev.n_emitters), cap_samples_);
  shortest_first_ = s->ctx->schedule == 1;
  slots_.assign(kMaxItems, ChunkRec());
  reqs_by_slot_.assign(kMaxRequests, nullptr);
  for (uint32_t i = kMaxItems; i-- > 0;) free_items_.push_back(i);
  for (uint32_t i = kMaxRequests; i-- > 0;) free_reqs_.push_back(i);
  free_ranges_[0] = cap_samples_;
  hip_check(hipEventCreate(&ev_start_), "hipEventCreate");
  hip_check(hipEventCreate(&ev_end_), "hipEventCreate");
  // k_extend's profiling events need eager launches (an event recorded by a
  // graph node carries no usable timestamp for hipEventElapsedTime)
  use_graph_ = s->ctx->graph && !profile_;
  reset_stats();
}

Engine::~Engine() {
  for (hipGraphExec_t &x : gexec_)
    if (x) (void)hipGraphExecDestroy(x);
  if (ev_start_) (void)hipEventDestroy(ev_start_);
  if (ev_end_) (void)hipEventDestroy(ev_end_);
}

double Engine::now_ms() const {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

RenderArgs Engine::args() const {
  RenderArgs A;
  std::memset(&A, 0, sizeof A);
  A.sc = s_->dev;
  A.reqs = ws_.reqs.as<DevRequest>();
  A.items = ws_.items.as<WorkItem>();
  A.plan = ws_.plan.as<RegenEntry>();
  A.hdr = ws_.hdr.as<BounceHdr>();
  A.n_paths = pool_;
  A.refill_extend = s_->ctx->refill_extend;
  A.refill_shadow = s_->ctx->refill_shadow;
  A.fetch_batch = s_->ctx->fetch_batch;
  A.item_done = ws_.item_done.as<uint32_t>();
  A.sample_L = ws_.sample_L.as<float4>();
  for (int i = 0; i < 2; i++) {
    A.q_o[i] = ws_.q_o[i].as<float4>();
    A.q_d[i] = ws_.q_d[i].as<float4>();
    A.q_beta[i] = ws_.q_beta[i].as<float4>();
    A.q_L[i] = ws_.q_L[i].as<float4>();
    A.q_aux[i] = ws_.q_aux[i].as<uint4>();
    A.q_meta[i] = ws_.q_meta[i].as<uint32_t>();
  }
  A.hit = ws_.hit.as<float4>();
  A.s_ent = ws_.s_ent.as<uint4>();
  A.sh_o = ws_.sh_o.as<float4>();
  A.sh_d = ws_.sh_d.as<float4>();
  A.sh_c = ws_.sh_c.as<float4>();
  A.sh_mask = ws_.sh_mask.as<uint32_t>();
  A.counters = ws_.counters.as<uint32_t>();
  A.stats = ws_.stats.as<unsigned long long>();
  A.req_invalid = ws_.req_invalid.as<unsigned long long>();
  return A;
}

std::unique_ptr<RequestRec> prepare_request(hk_scene *s, const hk_request &req, float *out_rgb, uint32_t chunk_cap) {
  const hk_scene_desc &d = s->desc;
  const float *c2w = req.use_scene_camera ? d.cam_to_world : req.cam_to_world;
  float fov = req.use_scene_camera ? d.fov : req.fov;
  int w = req.use_scene_camera ? d.width : req.width;
  int h = req.use_scene_camera ? d.height : req.height;
  if (w <= 0 || h <= 0 || uint64_t(w) * uint64_t(h) > (1u << 26))
    throw HkError(HK_ERR_INVALID, "bad film size %dx%d", w, h);
  if (!(fov > 0.0f && fov < 180.0f)) throw HkError(HK_ERR_INVALID, "bad field of view %f", fov);
  std::unique_ptr<RequestRec> r(new RequestRec());
  std::memset(&r->dev, 0, sizeof r->dev);
  camera_setup(c2w, fov, w, h, d.near_clip, d.far_clip, r->dev);
  filter_setup(d, r->dev);
  r->dev.seed = req.seed;
  r->dev.tile = s->ctx->tile;  // whole images start their camera rays in tiles
  r->npix = uint32_t(w) * uint32_t(h);
  if (r->npix > chunk_cap)
    throw HkError(HK_ERR_INVALID, "film of %u pixels exceeds the sample capacity (%u)", r->npix, chunk_cap);
  r->spp = req.spp;
  r->spp_offset = req.spp_offset;
  r->chunk = std::max(1u, std::min(r->spp == 0 ? 1u : r->spp, chunk_cap / r->npix));
  r->n_chunks = (r->spp + r->chunk - 1) / r->chunk;
  r->chunk_item.assign(r->n_chunks, -1);
  r->out_rgb = out_rgb;
  return r;
}

void Engine::submit(RequestRec *r) {
  if (free_reqs_.empty()) throw HkError(HK_ERR_BUDGET, "too many requests in flight (%u)", kMaxRequests);
  if (r->n_chunks > 0 && uint64_t(r->npix) * r->chunk > cap_samples_)
    throw HkError(HK_ERR_INVALID, "a chunk of %u x %u samples exceeds the sample buffer (%u)", r->npix, r->chunk,
                  cap_samples_);
  // allocations first: a request that cannot get its memory leaves the
  // engine as it was
  if (!r->keep_samples) {
    r->accum.alloc(size_t(r->npix) * 16);
    r->dev.accum = r->accum.as<float>();
  } else {
    r->samples.assign(size_t(r->spp) * r->npix * 4, 0.0f);
  }
  r->slot = free_reqs_.back();
  free_reqs_.pop_back();
  reqs_by_slot_[r->slot] = r;
  if (!r->keep_samples) hip_check(hipMemsetAsync(r->accum.p, 0, size_t(r->npix) * 16, st_), "accum clear");
  hip_check(launch_fill_u32(reinterpret_cast<uint32_t *>(ws_.req_invalid.as<unsigned long long>() + r->slot), 0u, 2u,
                            st_),
            "invalid-sample counter clear");
  DevRequest *hr = ws_.h_reqs.as<DevRequest>() + r->slot;
  *hr = r->dev;
  hip_check(hipMemcpyAsync(ws_.reqs.as<DevRequest>() + r->slot, hr, sizeof(DevRequest), hipMemcpyHostToDevice, st_),
            "request upload");
  r->t_submit = r->t_submit > 0 ? r->t_submit : now_ms();
  if (r->n_chunks == 0) {  // nothing to render: an empty (zero) film
    finalize(r);
    return;
  }
  pending_.push_back(r);
}

bool Engine::alloc_range(uint32_t n, uint32_t &base) {
  for (auto it = free_ranges_.begin(); it != free_ranges_.end(); ++it) {
    if (it->second >= n) {
      base = it->first;
      uint32_t len = it->second;
      free_ranges_.erase(it);
      if (len > n) free_ranges_[base + n] = len - n;
      return true;
    }
  }
  return false;
}

void Engine::free_range(uint32_t base, uint32_t n) {
  auto it = free_ranges_.emplace(base, n).first;
  auto nx = std::next(it);
  if (nx != free_ranges_.end() && it->first + it->second == nx->first) {
    it->second += nx->second;
    free_ranges_.erase(nx);
  }
  if (it != free_ranges_.begin()) {
    auto pv = std::prev(it);
    if (pv->first + pv->second == it->first) {
      pv->second += it->second;
      free_ranges_.erase(it);
    }
  }
}

// Admits chunks while an item slot and a sample-buffer range are free.  FIFO:
// in submission order, stopping at the first chunk that does not fit.
// Shortest first: requests with the fewest samples left go first, and a
// request whose next chunk does not fit is passed over for smaller ones.
void Engine::admit() {
  if (shortest_first_ && pending_.size() > 1) {
    std::vector<RequestRec *> order(pending_.begin(), pending_.end());
    std::stable_sort(order.begin(), order.end(),
                     [](const RequestRec *a, const RequestRec *b) { return a->left() < b->left(); });
    for (RequestRec *r : order) {
      if (free_items_.empty()) break;
      admit_chunks(r);
    }
    pending_.erase(std::remove_if(pending_.begin(), pending_.end(),
                                  [](const RequestRec *r) { return r->next_admit == r->n_chunks; }),
                   pending_.end());
    return;
  }
  while (!pending_.empty()) {
    RequestRec *r = pending_.front();
    if (!admit_chunks(r)) return;
    pending_.pop_front();
  }
}

// Admits r's next chunks while they fit; true when all are admitted.
bool Engine::admit_chunks(RequestRec *r) {
  while (r->next_admit < r->n_chunks) {
    uint32_t k = std::min(r->chunk, r->spp - r->next_admit * r->chunk);
    uint32_t size = r->npix * k;
    uint32_t base = 0;
    if (free_items_.empty() || !alloc_range(size, base)) return false;
    uint32_t slot = free_items_.back();
    free_items_.pop_back();
    ChunkRec &c = slots_[slot];
    c = ChunkRec();
    c.req = r;
    c.index = r->next_admit;
    c.size = size;
    c.buf_base = base;
    c.buf_len = size;
    WorkItem *hw = ws_.h_items.as<WorkItem>() + slot;
    std::memset(hw, 0, sizeof *hw);
    hw->req = r->slot;
    hw->npix = r->npix;
    hw->s0 = r->spp_offset + r->next_admit * r->chunk;
    hw->k = k;
    hw->buf_base = base;
    hip_check(hipMemcpyAsync(ws_.items.as<WorkItem>() + slot, hw, sizeof(WorkItem), hipMemcpyHostToDevice, st_),
              "item upload");
    hip_check(launch_clear_done(ws_.item_done.as<uint32_t>(), slot, st_), "item counter clear");
    r->chunk_item[r->next_admit] = int32_t(slot);
    if (r->t_admit < 0) r->t_admit = now_ms();
    r->next_admit++;
    active_.push_back(slot);
  }
  return true;
}

std::vector<RequestRec *> Engine::bounce() {
  if (!launch()) return {};
  return collect();
}

bool Engine::launch() {
  admit();
  // plan: new samples for the free places of the queue, oldest chunks first
  // (shortest first: chunks of the requests with the fewest samples left)
  uint32_t room = pool_ - in_flight_, n_new = 0, n_plan = 0;
  RegenEntry *plan = ws_.h_plan.as<RegenEntry>();
  const std::vector<uint32_t> *order = &active_;
  std::vector<uint32_t> by_left;
  if (shortest_first_ && active_.size() > 1) {
    std::vector<std::pair<uint64_t, uint32_t>> keyed;
    keyed.reserve(active_.size());
    for (uint32_t slot : active_) keyed.push_back({slots_[slot].req->left(), slot});
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<uint64_t, uint32_t> &a, const std::pair<uint64_t, uint32_t> &b) {
                       return a.first < b.first;
                     });
    for (auto &k : keyed) by_left.push_back(k.second);
    order = &by_left;
  }
  for (uint32_t slot : *order) {
    if (room == 0) break;
    ChunkRec &c = slots_[slot];
    if (c.started == c.size) continue;
    uint32_t take = std::min(room, c.size - c.started);
    plan[n_plan++] = RegenEntry{slot, c.started, take, n_new};
    c.started += take;
    c.req->started += take;
    n_new += take;
    room -= take;
  }
  if (in_flight_ + n_new == 0) {
    // nothing runs: every admitted chunk must have been resolved by now
    if (!active_.empty() || !pending_.empty())
      throw HkError(HK_ERR_INVALID, "scheduler stalled with %zu chunks active, %zu requests pending",
                    active_.size(), pending_.size());
    return false;
  }
  if (!timing_) {
    hip_check(hipEventRecord(ev_start_, st_), "event");
    timing_ = true;
  }
  *ws_.h_hdr.as<BounceHdr>() = BounceHdr{in_flight_, n_new, n_plan, 0u};
  // Long renders replay the bounce as a captured graph (one per queue
  // parity) once a few bounces have run eagerly; short ones stay eager, and
  // so do bounces with less than half a wavefront (the graph's grids are
  // sized for a full one).
  bool full = uint64_t(in_flight_ + n_new) * 2 >= pool_;
  if (use_graph_ && full && bounces_ >= kGraphAfter && !gexec_[qi_]) capture(qi_);
  if (full && gexec_[qi_]) {
    hip_check(hipGraphLaunch(gexec_[qi_], st_), "bounce graph launch");
    graph_bounces_++;
  } else {
    enqueue_bounce(qi_, false, n_plan, in_flight_ + n_new);
  }
  in_flight_ += n_new;
  paths_ += n_new;
  launched_ = true;
  return true;
}

// The work of one bounce on st_: bounce header (and plan) upload, k_regen,
// k_extend, k_shade, k_shadow, and the read-back of the queue counters and
// per-chunk done counts.  graph: the form captured into a graph, whose
// launches and copies must serve every bounce (k_regen always runs, grids and
// the plan copy are sized for the full wavefront); otherwise sized for this
// bounce (n_plan plan entries, `paths` paths in the queue after k_regen).
void Engine::enqueue_bounce(int qi, bool graph, uint32_t n_plan, uint32_t paths) {
  RenderArgs A = args();
  const BounceHdr *hh = ws_.h_hdr.as<BounceHdr>();
  hip_check(hipMemcpyAsync(ws_.hdr.p, hh, sizeof(BounceHdr), hipMemcpyHostToDevice, st_), "header upload");
  bool regen = graph || hh->n_new > 0;
  if (regen) {
    size_t plan_bytes = sizeof(RegenEntry) * (graph ? kMaxItems : n_plan);
    if (plan_bytes)
      hip_check(hipMemcpyAsync(ws_.plan.p, ws_.h_plan.p, plan_bytes, hipMemcpyHostToDevice, st_), "plan upload");
    hip_check(launch_regen(A, qi, graph ? pool_ : hh->n_new, st_), "k_regen");
  }
  uint32_t g = kernel_grid(graph ? pool_ : paths, s_->ctx->grid);
  if (profile_) hip_check(hipEventRecord(ev0_, st_), "event");
  hip_check(launch_extend(A, qi, g, st_), "k_extend");
  if (profile_) hip_check(hipEventRecord(ev1_, st_), "event");
  hip_check(launch_shade(A, qi, g, st_), "k_shade");
  // the shadow queue holds at most one entry per shaded path: size the grid
  // for that bound, the kernel reads the real count
  hip_check(launch_shadow(A, qi ^ 1, g, st_), "k_shadow");
  hip_check(hipMemcpyAsync(ws_.h_counters.p, A.counters, sizeof(uint32_t) * (ctr(1) + 1), hipMemcpyDeviceToHost, st_),
            "counters D2H");
  hip_check(hipMemcpyAsync(ws_.h_done.p, A.item_done, sizeof(uint32_t) * kMaxItems * kDoneWays, hipMemcpyDeviceToHost,
                           st_),
            "item counts D2H");
}

// Captures enqueue_bounce(qi, graph form) on st_ and instantiates it.  Nothing
// captured runs during the capture; a capture or instantiation the runtime
// refuses leaves the engine eager (the bounce is then enqueued directly).
void Engine::capture(int qi) {
  hipGraph_t g = nullptr;
  bool ok = hipStreamBeginCapture(st_, hipStreamCaptureModeThreadLocal) == hipSuccess;
  if (ok) {
    try {
      enqueue_bounce(qi, true, kMaxItems, pool_);
    } catch (const HkError &) {
      ok = false;
    }
    hipGraph_t got = nullptr;
    if (hipStreamEndCapture(st_, &got) != hipSuccess) ok = false;
    g = got;
  }
  if (ok && hipGraphInstantiate(&gexec_[qi], g, nullptr, nullptr, 0) != hipSuccess) ok = false;
  if (g) (void)hipGraphDestroy(g);
  if (!ok) {
    (void)hipGetLastError();
    if (gexec_[qi]) (void)hipGraphExecDestroy(gexec_[qi]);
    gexec_[qi] = nullptr;
    use_graph_ = false;
  }
}

std::vector<RequestRec *> Engine::collect() {
  std::vector<RequestRec *> finished;
  if (!launched_) return finished;
  launched_ = false;
  hip_check(hipStreamSynchronize(st_), "bounce sync");
  if (profile_) {
    float ms = 0;
    hip_check(hipEventElapsedTime(&ms, ev0_, ev1_), "event time");
    trace_ms_ += ms;
    trace_launches_++;
  }
  bounces_++;
  qi_ ^= 1;
  in_flight_ = ws_.h_counters.as<uint32_t>()[ctr(uint32_t(qi_))];
  const uint32_t *done = ws_.h_done.as<uint32_t>();
  for (uint32_t slot : active_) {
    ChunkRec &c = slots_[slot];
    if (c.complete) continue;
    uint32_t n = 0;
    for (uint32_t w = 0; w < kDoneWays; w++) n += done[w * kMaxItems + slot];
    if (n == c.size) c.complete = true;
  }
  resolve_ready(finished);
  return finished;
}

// Resolves, per request and in chunk order, every chunk whose samples have
// all finished; frees their slots and buffer ranges; finalizes requests.
void Engine::resolve_ready(std::vector<RequestRec *> &finished) {
  std::vector<uint32_t> keep;
  keep.reserve(active_.size());
  std::vector<RequestRec *> touched;
  for (uint32_t slot : active_) {
    ChunkRec &c = slots_[slot];
    if (c.complete && std::find(touched.begin(), touched.end(), c.req) == touched.end()) touched.push_back(c.req);
  }
  RenderArgs A = args();
  for (RequestRec *r : touched) {
    while (r->next_resolve < r->next_admit) {
      int32_t slot = r->chunk_item[r->next_resolve];
      ChunkRec &c = slots_[uint32_t(slot)];
      if (!c.complete) break;
      if (r->cancelled) {
        // the chunk's paths have finished writing its buffer range: nothing to resolve
      } else if (r->keep_samples) {
        float *dst = r->samples.data() + size_t(r->next_resolve) * r->chunk * r->npix * 4;
        hip_check(hipMemcpyAsync(dst, ws_.sample_L.as<float4>() + c.buf_base, size_t(c.size) * 16,
                                 hipMemcpyDeviceToHost, st_),
                  "samples D2H");
        hip_check(hipStreamSynchronize(st_), "samples sync");
      } else {
        hip_check(launch_resolve(A, uint32_t(slot), r->dev, r->npix, s_->ctx->resolve_tiles, st_), "k_resolve");
      }
      free_range(c.buf_base, c.buf_len);
      free_items_.push_back(uint32_t(slot));
      c.req = nullptr;
      r->chunk_item[r->next_resolve] = -1;
      r->next_resolve++;
    }
    if (r->next_resolve == r->n_chunks) {
      finalize(r);
      finished.push_back(r);
    }
  }
  for (uint32_t slot : active_)
    if (slots_[slot].req != nullptr) keep.push_back(slot);
  active_.swap(keep);
}

void Engine::finalize(RequestRec *r) {
  unsigned long long *inv = ws_.h_req_invalid.as<unsigned long long>() + r->slot;
  hip_check(hipMemcpyAsync(inv, ws_.req_invalid.as<unsigned long long>() + r->slot, sizeof *inv, hipMemcpyDeviceToHost,
                           st_),
            "invalid-sample count D2H");
  DBuf rgb;
  if (r->out_rgb && !r->keep_samples && !r->cancelled) {
    rgb.alloc(size_t(r->npix) * 12);
    hip_check(launch_finalize(r->accum.p, rgb.as<float>(), r->npix, st_), "k_finalize");
    hip_check(hipMemcpyAsync(r->out_rgb, rgb.p, size_t(r->npix) * 12, hipMemcpyDeviceToHost, st_), "film D2H");
  }
  hip_check(hipStreamSynchronize(st_), "film sync");
  r->invalid_samples = *inv;
  r->accum.release();
  reqs_by_slot_[r->slot] = nullptr;
  free_reqs_.push_back(r->slot);
  r->t_done = now_ms();
  r->done = true;
}

void Engine::run() {
  while (busy()) bounce();
}

std::vector<RequestRec *> Engine::cancel(RequestRec *r) {
  std::vector<RequestRec *> finished;
  if (r->done || r->cancelled) return finished;
  r->cancelled = true;
  r->n_chunks = r->next_admit;
  pending_.erase(std::remove(pending_.begin(), pending_.end(), r), pending_.end());
  // Host-side started counts include the samples of a bounce launched but
  // not yet collected, so a shrunk chunk completes when every path it
  // started has counted itself done: at once if it started none or the last
  // collected counts already show them all (nothing may be in flight then),
  // otherwise in a later collect().
  const uint32_t *done = ws_.h_done.as<uint32_t>();
  for (uint32_t k = r->next_resolve; k < r->next_admit; k++) {
    uint32_t slot = uint32_t(r->chunk_item[k]);
    ChunkRec &c = slots_[slot];
    c.size = c.started;
    uint32_t n = 0;
    if (!launched_)
      for (uint32_t w = 0; w < kDoneWays; w++) n += done[w * kMaxItems + slot];
    if (c.size == 0 || n == c.size) c.complete = true;
  }
  if (r->next_resolve == r->n_chunks) {  // nothing admitted and unresolved
    finalize(r);
    finished.push_back(r);
    return finished;
  }
  resolve_ready(finished);
  return finished;
}

uint64_t Engine::unstarted() const {
  uint64_t n = 0;
  if (free_reqs_.size() == kMaxRequests) return 0;
  for (const RequestRec *r : reqs_by_slot_)
    if (r) n += r->left();
  return n;
}

double Engine::oldest_submit() const {
  double t = HUGE_VAL;
  if (free_reqs_.size() == kMaxRequests) return t;
  for (const RequestRec *r : reqs_by_slot_)
    if (r) t = std::min(t, r->t_submit);
  return t;
}

void Engine::reset_stats() {
  hip_check(hipMemsetAsync(ws_.stats.p, 0, 64, st_), "stats clear");
  paths_ = bounces_ = trace_launches_ = graph_bounces_ = 0;
  trace_ms_ = 0;
  timing_ = false;
}

void Engine::take_stats(hk_render_stats *st) {
  hip_check(hipEventRecord(ev_end_, st_), "event");
  unsigned long long hs[6] = {0, 0, 0, 0, 0, 0};
  hip_check(hipMemcpyAsync(hs, ws_.stats.p, sizeof hs, hipMemcpyDeviceToHost, st_), "stats D2H");
  hip_check(hipStreamSynchronize(st_), "stats sync");
  if (!st) return;
  float ms = 0;
  if (timing_) hip_check(hipEventElapsedTime(&ms, ev_start_, ev_end_), "event time");
  st->paths += paths_;
  st->extension_rays += hs[0];
  st->shadow_rays += hs[1];
  st->bounces_launched += bounces_;
  st->gpu_ms += ms;
  st->trace_ms += trace_ms_;
  st->trace_launches += trace_launches_;
  st->node_visits += hs[2] + hs[4];
  st->triangle_tests += hs[3] + hs[5];
  st->graph_bounces += graph_bounces_;
}

}  // namespace hk
