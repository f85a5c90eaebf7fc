// hk_server.cpp -- request loop with continuous batching over resident scenes.
//
// Every resident scene has two lanes, each an Engine with its own HIP stream
// (default priority):
//   * the bulk lane: a wavefront of max_paths (64 M when the option is not
//     set) and a sample buffer of 4 x max_samples holding up
//     to four chunks of max_samples.  Bulk lanes never run at the same time
//     (below), so the scenes' bulk engines share one workspace, sized for the
//     largest of them when the scenes are added: a resident scene adds its
//     express lane, not a second wavefront;
//   * the express lane (its own workspace): a wavefront of max_paths / 16 (4 M
//     when the option is not set) and a buffer of
//     max_samples / 8, for requests of at most half that many samples,
//     fewest samples left first (context option "express" = 0 turns it off:
//     every request then waits its turn on the bulk lane).
// Bulk requests run one at a time, in arrival order over all scenes, as a
// first-come-first-served queue of plain renders would run them.  A request
// finishes only when its longest Russian-roulette path does; sharing a
// wavefront (or the device) with the next request's full bounces would stretch
// that tail from a few hundred tiny bounces into as many full ones, so large
// requests never overlap each other.  Small requests pass them on the
// express lanes, whose bounces run beside the bulk lane's.
// One worker thread drives all lanes without a common barrier: it collects a
// lane's bounce as soon as the device has finished it (hipStreamQuery) and
// launches the lane's next one, so each lane advances at its own pace and the
// lanes' kernels overlap on the device.  A request that arrives while others
// run is admitted (HBM budget permitting); an express request's paths join
// its lane's next bounce.
// Results equal the same requests rendered alone with hk_render on a context
// with the same options, bit for bit: a film depends neither on the wavefront
// size nor on the chunking (see hk_engine.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/hikari_amd.h"
#include "hk_engine.h"
#include "hk_runtime.h"

using namespace hk;

namespace hk {
extern thread_local std::string g_create_error;
}

namespace {

enum { kBulk = 0, kExpress = 1 };

struct Lane {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Workspace ws;                // the express lane's (bulk lanes use the server's)
  std::unique_ptr<Engine> eng;
  uint32_t chunk_cap = 0;      // samples per chunk at most (prepare_request)
  uint64_t max_request = 0;    // express: largest request (samples) it takes
  ~Lane() {
    eng.reset();
    ws.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct ServerScene {
  hk_scene *scene = nullptr;
  std::unique_ptr<Lane> lane[2];
};

struct Ticket {
  std::unique_ptr<RequestRec> rec;
  int32_t scene = 0;
  int lane = kBulk;
  int status = HK_OK;
  std::string err;
  bool done = false;
  bool admitted = false;  // its film counts against the budget
  bool submitted = false; // handed to its lane's engine (worker only)
};

}  // namespace

struct hk_server {
  hk_ctx *ctx = nullptr;
  uint64_t budget = 0, fixed_bytes = 0, film_bytes = 0;
  Workspace bulk_ws;  // shared by the bulk lanes (declared before `scenes`: freed after their engines)
  std::vector<std::unique_ptr<ServerScene>> scenes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> arrivals;          // tickets waiting for admission
  std::vector<uint64_t> cancels;          // admitted tickets to cancel (the worker's next round)
  std::map<uint64_t, Ticket> tickets;
  uint64_t next_ticket = 1;
  std::thread worker;
  bool started = false, stop = false;
  std::string fatal;                      // error that stopped the worker
  hk_server_stats st;
  double t0 = 0;                          // server clock origin (steady clock, ms)
};

static double clock_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

static void finish_ticket(hk_server *srv, Ticket &t, int status, const std::string &err) {
  if (status == HK_ERR_CANCELLED) srv->st.cancelled++;
  t.status = status;
  t.err = err;
  t.done = true;
  if (t.admitted) srv->film_bytes -= t.rec->film_bytes();
  t.admitted = false;
}

// A failure of a bounce (a HIP error, a lost device) ends the worker: every
// unfinished ticket fails with it and later submits are refused.
static void stop_worker(hk_server *srv, int code, const std::string &msg) {
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->fatal = msg;
  for (auto &kv : srv->tickets)
    if (!kv.second.done) finish_ticket(srv, kv.second, code, msg);
  srv->arrivals.clear();
  srv->cv_done.notify_all();
}

static bool any_lane_active(const hk_server *srv) {
  for (auto &sc : srv->scenes)
    for (auto &l : sc->lane)
      if (l->eng->busy() || l->eng->launched()) return true;
  return false;
}

static bool bulk_idle(const hk_server *srv) {
  for (auto &sc : srv->scenes) {
    const Engine &e = *sc->lane[kBulk]->eng;
    if (e.busy() || e.launched()) return false;
  }
  return true;
}

static void worker_main(hk_server *srv) {
  (void)hipSetDevice(srv->ctx->device);
  // bulk requests admitted (HBM budget) and waiting for their turn, in
  // arrival order over all scenes; only the worker touches it
  std::deque<Ticket *> bulk_wait;
  for (;;) {
    std::vector<Ticket *> admitted;
    std::vector<uint64_t> cancels;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      // HBM-budget admission, first come first served.  A request whose
      // lane has no free request slot stays queued (later requests for
      // other lanes may pass it); one whose film does not fit the budget
      // waits, with everything behind it, until running requests free memory.
      std::vector<uint32_t> slots(srv->scenes.size() * 2);
      for (size_t k = 0; k < srv->scenes.size(); k++)
        for (int l = 0; l < 2; l++) slots[2 * k + l] = srv->scenes[k]->lane[l]->eng->free_request_slots();
      for (auto it = srv->arrivals.begin(); it != srv->arrivals.end();) {
        Ticket &t = srv->tickets[*it];
        uint32_t &free = slots[2 * size_t(t.scene) + t.lane];
        if (free == 0) {
          ++it;
          continue;
        }
        uint64_t need = t.rec->film_bytes();
        bool fits = srv->budget == 0 || srv->fixed_bytes + srv->film_bytes + need <= srv->budget;
        if (!fits && srv->film_bytes > 0) break;  // wait for running requests to free memory
        it = srv->arrivals.erase(it);
        if (!fits) {
          finish_ticket(srv, t, HK_ERR_BUDGET, "request film exceeds the HBM budget");
          t.rec.reset();
          srv->cv_done.notify_all();
          continue;
        }
        srv->film_bytes += need;
        t.admitted = true;
        free--;
        admitted.push_back(&t);
      }
      cancels.swap(srv->cancels);
      if (admitted.empty() && bulk_wait.empty() && cancels.empty() && !any_lane_active(srv)) {
        // (nothing active means no request is in flight, so every queued
        // request would have been admitted or refused above)
        if (srv->stop) return;
        srv->cv_work.wait(lk);
        continue;
      }
    }
    // Express requests go to their lane's engine at once; bulk requests one
    // at a time, the next when no bulk lane has one in flight.  A request the
    // engine refuses (e.g. no device memory for its film) fails alone; the
    // engine is left as it was.  Only the worker finishes tickets, so the
    // accepted ones stay valid until this pass of the loop ends.
    std::vector<Ticket *> to_submit;
    for (Ticket *tp : admitted) {
      if (tp->lane == kBulk) bulk_wait.push_back(tp);
      else to_submit.push_back(tp);
    }
    // Cancellations of admitted requests: one waiting for the bulk lane is
    // dropped here; one in an engine stops there, and finishes (cancelled)
    // now or when its started paths have finished.
    std::vector<RequestRec *> cancelled_now;
    for (uint64_t id : cancels) {
      Ticket *tp = nullptr;
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->tickets.find(id);
        if (it != srv->tickets.end() && !it->second.done) tp = &it->second;
      }
      if (!tp) continue;
      auto wit = std::find(bulk_wait.begin(), bulk_wait.end(), tp);
      auto sit = std::find(to_submit.begin(), to_submit.end(), tp);
      if (wit != bulk_wait.end() || sit != to_submit.end()) {
        if (wit != bulk_wait.end()) bulk_wait.erase(wit);
        if (sit != to_submit.end()) to_submit.erase(sit);
        std::lock_guard<std::mutex> lk(srv->mu);
        finish_ticket(srv, *tp, HK_ERR_CANCELLED, "request cancelled");
        srv->cv_done.notify_all();
      } else if (tp->submitted) {
        try {
          std::vector<RequestRec *> f = srv->scenes[size_t(tp->scene)]->lane[tp->lane]->eng->cancel(tp->rec.get());
          cancelled_now.insert(cancelled_now.end(), f.begin(), f.end());
        } catch (const HkError &e) {
          stop_worker(srv, e.code, e.msg);
          return;
        } catch (const std::exception &e) {
          stop_worker(srv, HK_ERR_HIP, e.what());
          return;
        }
      }
    }
    if (!bulk_wait.empty() && bulk_idle(srv)) {
      to_submit.push_back(bulk_wait.front());
      bulk_wait.pop_front();
    }
    std::vector<Ticket *> accepted;
    for (Ticket *tp : to_submit) {
      Ticket &t = *tp;
      int code = HK_OK;
      std::string msg;
      try {
        srv->scenes[size_t(t.scene)]->lane[t.lane]->eng->submit(t.rec.get());
      } catch (const HkError &e) {
        code = e.code;
        msg = e.msg;
      } catch (const std::exception &e) {
        code = HK_ERR_OOM;
        msg = e.what();
      }
      if (code == HK_OK) {
        t.submitted = true;
        accepted.push_back(&t);
        continue;
      }
      std::lock_guard<std::mutex> lk(srv->mu);
      finish_ticket(srv, t, code, msg);
      srv->cv_done.notify_all();
    }
    try {
      bool progressed = !to_submit.empty() || !cancels.empty();
      std::vector<RequestRec *> finished = cancelled_now;
      // bounces the device has finished
      for (auto &sc : srv->scenes)
        for (auto &l : sc->lane)
          if (l->eng->ready()) {
            std::vector<RequestRec *> f = l->eng->collect();
            finished.insert(finished.end(), f.begin(), f.end());
            progressed = true;
          }
      // next bounces of every lane with work
      uint64_t in_flight = 0;  // requests holding an engine's request slot
      for (auto &sc : srv->scenes)
        for (auto &l : sc->lane) in_flight += l->eng->requests_in_flight();
      uint64_t launches = 0;
      for (auto &sc : srv->scenes)
        for (auto &l : sc->lane) {
          Engine &e = *l->eng;
          if (!e.launched() && e.busy() && e.launch()) launches++;
        }
      progressed = progressed || launches > 0;
      // requests with no samples finish at submit
      for (Ticket *t : accepted)
        if (t->rec->done && std::find(finished.begin(), finished.end(), t->rec.get()) == finished.end())
          finished.push_back(t->rec.get());
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->st.bounces += launches;
        if (in_flight >= 2) srv->st.shared_bounces += launches;
        srv->st.max_in_flight = std::max<uint64_t>(srv->st.max_in_flight, in_flight);
        srv->st.requests_in_flight = in_flight;
        for (RequestRec *r : finished) {
          Ticket &t = srv->tickets[r->id];
          if (r->cancelled) {
            finish_ticket(srv, t, HK_ERR_CANCELLED, "request cancelled");
            continue;
          }
          finish_ticket(srv, t, HK_OK, "");
          srv->st.completed++;
          srv->st.paths += uint64_t(r->npix) * r->spp;
        }
        if (!finished.empty()) srv->cv_done.notify_all();
      }
      // every launched bounce is still running: wait a little for one
      if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    } catch (const HkError &e) {
      stop_worker(srv, e.code, e.msg);
      return;
    } catch (const std::exception &e) {
      stop_worker(srv, HK_ERR_HIP, e.what());
      return;
    }
  }
}

#define HK_API_BEGIN try {
#define HK_API_END(ctxp)                 \
  }                                      \
  catch (const HkError &e) {             \
    if (ctxp) (ctxp)->err = e.msg;       \
    else g_create_error = e.msg;         \
    return e.code;                       \
  }                                      \
  catch (const std::exception &e) {      \
    if (ctxp) (ctxp)->err = e.what();    \
    else g_create_error = e.what();      \
    return HK_ERR_INVALID;               \
  }

static std::unique_ptr<Lane> make_lane(hk_ctx *ctx, hk_scene *scene, uint32_t paths, uint32_t samples,
                                       uint32_t chunk_cap, bool shortest_first, Workspace *shared = nullptr) {
  std::unique_ptr<Lane> l(new Lane());
  hip_check(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreate(&l->ev0), "hipEventCreate");
  hip_check(hipEventCreate(&l->ev1), "hipEventCreate");
  l->chunk_cap = chunk_cap;
  l->eng.reset(new Engine(scene, shared ? *shared : l->ws, l->stream, paths, samples, chunk_cap, false, l->ev0, l->ev1));
  l->eng->set_shortest_first(shortest_first);
  hip_check(hipStreamSynchronize(l->stream), "workspace");
  return l;
}

extern "C" {

int hk_server_create(hk_ctx *ctx, uint64_t hbm_budget, hk_server **out) {
  HK_API_BEGIN
  if (!ctx || !out) throw HkError(HK_ERR_INVALID, "null argument");
  std::unique_ptr<hk_server> srv(new hk_server());
  srv->ctx = ctx;
  srv->budget = hbm_budget;
  std::memset(&srv->st, 0, sizeof srv->st);
  srv->t0 = clock_ms();
  *out = srv.release();
  return HK_OK;
  HK_API_END(ctx)
}

int hk_server_add_scene(hk_server *srv, hk_scene *scene, int32_t *scene_id) {
  hk_ctx *ctx = srv ? srv->ctx : nullptr;
  HK_API_BEGIN
  if (!srv || !scene) throw HkError(HK_ERR_INVALID, "null argument");
  if (scene->ctx != srv->ctx) throw HkError(HK_ERR_INVALID, "scene belongs to another context");
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->started) throw HkError(HK_ERR_INVALID, "scenes must be added before the first submit");
  hip_check(hipSetDevice(ctx->device), "hipSetDevice");
  std::unique_ptr<ServerScene> sc(new ServerScene());
  sc->scene = scene;
  uint64_t bulk_samples = std::min<uint64_t>(uint64_t(ctx->max_samples) * 4, 0x7fffffffu);
  // the bulk lane joins the shared workspace (which grows to this scene's
  // needs: its emitters' shadow slots, the wavefront); the growth counts
  // against the budget even when the scene is refused, since it stays held
  uint64_t shared_before = srv->bulk_ws.device_bytes();
  uint32_t ex_paths = 1u, ex_samples = 1u;
  try {
    uint32_t bulk_paths = ctx->max_paths_set ? ctx->max_paths : kDefaultServerBulkPaths;
    sc->lane[kBulk] = make_lane(ctx, scene, bulk_paths, uint32_t(bulk_samples), ctx->max_samples, false,
                                &srv->bulk_ws);
    // express lane (option "express" = 0: a placeholder that takes nothing,
    // so every request runs on the bulk lane in arrival order)
    if (ctx->express) {
      ex_samples = std::max<uint32_t>(1, ctx->max_samples / 8);
      ex_paths = ctx->max_paths_set ? std::max<uint32_t>(1, ctx->max_paths / 16) : kDefaultExpressPaths;
    }
    sc->lane[kExpress] = make_lane(ctx, scene, ex_paths, ex_samples, ex_samples, true);
  } catch (...) {
    srv->fixed_bytes += srv->bulk_ws.device_bytes() - shared_before;
    throw;
  }
  sc->lane[kExpress]->max_request = ctx->express ? ex_samples / 2 : 0;
  uint64_t grown = srv->bulk_ws.device_bytes() - shared_before;
  uint64_t ws_bytes = grown + sc->lane[kExpress]->ws.device_bytes();
  if (srv->budget && srv->fixed_bytes + ws_bytes > srv->budget) {
    srv->fixed_bytes += grown;
    throw HkError(HK_ERR_BUDGET, "scene workspaces (%llu B) exceed the HBM budget (%llu B)",
                  (unsigned long long)(srv->fixed_bytes + ws_bytes - grown), (unsigned long long)srv->budget);
  }
  srv->fixed_bytes += ws_bytes;
  if (scene_id) *scene_id = int32_t(srv->scenes.size());
  srv->scenes.push_back(std::move(sc));
  return HK_OK;
  HK_API_END(ctx)
}

int hk_server_submit(hk_server *srv, int32_t scene_id, const hk_request *req, float *out_rgb, uint64_t *ticket) {
  hk_ctx *ctx = srv ? srv->ctx : nullptr;
  HK_API_BEGIN
  if (!srv || !req || !ticket) throw HkError(HK_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(srv->mu);
  if (!srv->fatal.empty()) throw HkError(HK_ERR_HIP, "server stopped: %s", srv->fatal.c_str());
  if (scene_id < 0 || size_t(scene_id) >= srv->scenes.size()) throw HkError(HK_ERR_INVALID, "bad scene id %d", scene_id);
  ServerScene &sc = *srv->scenes[size_t(scene_id)];
  // the lane: express for requests of at most its limit whose film fits its
  // chunks, bulk otherwise
  const hk_scene_desc &d = sc.scene->desc;
  uint64_t w = uint64_t(req->use_scene_camera ? d.width : req->width);
  uint64_t h = uint64_t(req->use_scene_camera ? d.height : req->height);
  const Lane &ex = *sc.lane[kExpress];
  int lane = (ex.max_request > 0 && w * h * req->spp <= ex.max_request && w * h <= ex.chunk_cap) ? kExpress : kBulk;
  std::unique_ptr<RequestRec> r = prepare_request(sc.scene, *req, out_rgb, sc.lane[lane]->chunk_cap);
  uint64_t id = srv->next_ticket++;
  r->id = id;
  r->t_submit = clock_ms();
  Ticket &t = srv->tickets[id];
  t.rec = std::move(r);
  t.scene = scene_id;
  t.lane = lane;
  srv->arrivals.push_back(id);
  srv->st.submitted++;
  if (!srv->started) {
    srv->started = true;
    srv->worker = std::thread(worker_main, srv);
  }
  lk.unlock();
  srv->cv_work.notify_one();
  *ticket = id;
  return HK_OK;
  HK_API_END(ctx)
}

int hk_server_wait(hk_server *srv, uint64_t ticket, hk_server_result *res) {
  hk_ctx *ctx = srv ? srv->ctx : nullptr;
  HK_API_BEGIN
  if (!srv) throw HkError(HK_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->tickets.find(ticket);
  if (it == srv->tickets.end()) throw HkError(HK_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
  srv->cv_done.wait(lk, [&] { return it->second.done; });
  Ticket &t = it->second;
  int status = t.status;
  std::string err = t.err;
  if (res) {
    std::memset(res, 0, sizeof *res);
    res->status = status;
    if (t.rec) {
      res->queue_ms = t.rec->t_admit >= 0 ? t.rec->t_admit - t.rec->t_submit : 0.0;
      res->latency_ms = t.rec->t_done >= 0 ? t.rec->t_done - t.rec->t_submit : 0.0;
      res->submit_ms = t.rec->t_submit - srv->t0;
      res->paths = uint64_t(t.rec->npix) * t.rec->spp;
      res->lane = t.lane;
      res->invalid_samples = t.rec->invalid_samples;
    }
  }
  srv->tickets.erase(it);
  if (status != HK_OK) throw HkError(status, "%s", err.c_str());
  return HK_OK;
  HK_API_END(ctx)
}

int hk_server_cancel(hk_server *srv, uint64_t ticket) {
  hk_ctx *ctx = srv ? srv->ctx : nullptr;
  HK_API_BEGIN
  if (!srv) throw HkError(HK_ERR_INVALID, "null argument");
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end()) throw HkError(HK_ERR_INVALID, "unknown ticket %llu", (unsigned long long)ticket);
    Ticket &t = it->second;
    if (t.done) return HK_OK;  // finished (or failed) before the call: keeps its result
    auto a = std::find(srv->arrivals.begin(), srv->arrivals.end(), ticket);
    if (a != srv->arrivals.end()) {  // not admitted yet: the worker never saw it
      srv->arrivals.erase(a);
      finish_ticket(srv, t, HK_ERR_CANCELLED, "request cancelled");
      srv->cv_done.notify_all();
      return HK_OK;
    }
    if (std::find(srv->cancels.begin(), srv->cancels.end(), ticket) == srv->cancels.end())
      srv->cancels.push_back(ticket);
  }
  srv->cv_work.notify_one();
  return HK_OK;
  HK_API_END(ctx)
}

int hk_server_get_stats(hk_server *srv, hk_server_stats *out) {
  hk_ctx *ctx = srv ? srv->ctx : nullptr;
  HK_API_BEGIN
  if (!srv || !out) throw HkError(HK_ERR_INVALID, "null argument");
  std::lock_guard<std::mutex> lk(srv->mu);
  *out = srv->st;
  out->fixed_bytes = srv->fixed_bytes;
  out->film_bytes = srv->film_bytes;
  return HK_OK;
  HK_API_END(ctx)
}

int hk_server_destroy(hk_server *srv) {
  if (!srv) return HK_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  delete srv;
  return HK_OK;
}

}  // extern "C"
