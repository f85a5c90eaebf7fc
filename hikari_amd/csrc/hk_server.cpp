// hk_server.cpp -- request loop with continuous batching over resident scenes.
//
// One worker thread owns one Engine per scene, each on its own HIP stream
// (default priority).  Every round it admits arrived requests while the HBM
// budget allows (FIFO), issues one bounce on every busy engine, then collects
// them, so the scenes' kernels overlap on the device and a request that
// arrives while others run joins their very next bounce.  Results equal the
// same requests rendered alone with hk_render on a context with the same
// max_paths / max_samples, bit for bit (see hk_engine.h).
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

struct ServerScene {
  hk_scene *scene = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  Workspace ws;
  std::unique_ptr<Engine> eng;
  ~ServerScene() {
    eng.reset();
    ws.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

struct Ticket {
  std::unique_ptr<RequestRec> rec;
  int32_t scene = 0;
  int status = HK_OK;
  std::string err;
  bool done = false;
  bool admitted = false;  // its film counts against the budget
  double t_arrive = 0;
};

}  // namespace

struct hk_server {
  hk_ctx *ctx = nullptr;
  uint64_t budget = 0, fixed_bytes = 0, film_bytes = 0;
  std::vector<std::unique_ptr<ServerScene>> scenes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<uint64_t> arrivals;          // tickets waiting for admission
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

static void worker_main(hk_server *srv) {
  (void)hipSetDevice(srv->ctx->device);
  for (;;) {
    std::vector<std::pair<uint64_t, Ticket *>> admitted;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      // HBM-budget admission, first come first served.  A request whose
      // scene has no free request slot stays queued (later requests for
      // other scenes may pass it); one whose film does not fit the budget
      // waits, with everything behind it, until running requests free memory.
      std::vector<uint32_t> slots(srv->scenes.size());
      for (size_t k = 0; k < slots.size(); k++) slots[k] = srv->scenes[k]->eng->free_request_slots();
      for (auto it = srv->arrivals.begin(); it != srv->arrivals.end();) {
        uint64_t id = *it;
        Ticket &t = srv->tickets[id];
        if (slots[size_t(t.scene)] == 0) {
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
        slots[size_t(t.scene)]--;
        admitted.push_back({id, &t});
      }
      bool any = !admitted.empty();
      for (auto &sc : srv->scenes) any = any || sc->eng->busy();
      if (!any) {
        // (nothing busy means no request is in flight, so every queued
        // request would have been admitted or refused above)
        if (srv->stop) return;
        srv->cv_work.wait(lk);
        continue;
      }
    }
    // A request the engine refuses (e.g. no device memory for its film)
    // fails alone; the engine is left as it was.  Only the worker finishes
    // tickets, so the accepted ones stay valid until this round ends.
    std::vector<Ticket *> accepted;
    for (auto &a : admitted) {
      Ticket &t = *a.second;
      ServerScene &sc = *srv->scenes[size_t(t.scene)];
      int code = HK_OK;
      std::string msg;
      try {
        sc.eng->submit(t.rec.get());
      } catch (const HkError &e) {
        code = e.code;
        msg = e.msg;
      } catch (const std::exception &e) {
        code = HK_ERR_OOM;
        msg = e.what();
      }
      if (code == HK_OK) {
        accepted.push_back(&t);
        continue;
      }
      std::lock_guard<std::mutex> lk(srv->mu);
      finish_ticket(srv, t, code, msg);
      srv->cv_done.notify_all();
    }
    try {
      uint64_t in_flight = 0;
      for (auto &sc : srv->scenes) in_flight += sc->eng->requests_in_flight();
      std::vector<ServerScene *> launched;
      for (auto &sc : srv->scenes)
        if (sc->eng->busy() && sc->eng->launch()) launched.push_back(sc.get());
      std::vector<RequestRec *> finished;
      for (ServerScene *sc : launched) {
        std::vector<RequestRec *> f = sc->eng->collect();
        finished.insert(finished.end(), f.begin(), f.end());
      }
      // requests with no samples finish at submit
      for (Ticket *t : accepted)
        if (t->rec->done && std::find(finished.begin(), finished.end(), t->rec.get()) == finished.end())
          finished.push_back(t->rec.get());
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->st.bounces += launched.size();
      if (in_flight >= 2) srv->st.shared_bounces += launched.size();
      srv->st.max_in_flight = std::max<uint64_t>(srv->st.max_in_flight, in_flight);
      for (RequestRec *r : finished) {
        Ticket &t = srv->tickets[r->id];
        finish_ticket(srv, t, HK_OK, "");
        srv->st.completed++;
        srv->st.paths += uint64_t(r->npix) * r->spp;
      }
      if (!finished.empty()) srv->cv_done.notify_all();
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
  std::unique_ptr<ServerScene> sc(new ServerScene());
  sc->scene = scene;
  hip_check(hipSetDevice(ctx->device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking), "hipStreamCreate");
  hip_check(hipEventCreate(&sc->ev0), "hipEventCreate");
  hip_check(hipEventCreate(&sc->ev1), "hipEventCreate");
  sc->eng.reset(new Engine(scene, sc->ws, sc->stream, ctx->max_paths, ctx->max_samples, ctx->max_samples, false,
                           sc->ev0, sc->ev1));
  hip_check(hipStreamSynchronize(sc->stream), "workspace");
  uint64_t ws_bytes = sc->ws.device_bytes();
  if (srv->budget && srv->fixed_bytes + ws_bytes > srv->budget)
    throw HkError(HK_ERR_BUDGET, "scene workspaces (%llu B) exceed the HBM budget (%llu B)",
                  (unsigned long long)(srv->fixed_bytes + ws_bytes), (unsigned long long)srv->budget);
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
  std::unique_ptr<RequestRec> r = prepare_request(srv->scenes[size_t(scene_id)]->scene, *req, out_rgb,
                                                  ctx->max_samples);
  uint64_t id = srv->next_ticket++;
  r->id = id;
  r->t_submit = clock_ms();
  Ticket &t = srv->tickets[id];
  t.rec = std::move(r);
  t.scene = scene_id;
  t.t_arrive = t.rec->t_submit;
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
      res->latency_ms = t.rec->t_done - t.rec->t_submit;
      res->submit_ms = t.rec->t_submit - srv->t0;
      res->paths = uint64_t(t.rec->npix) * t.rec->spp;
    }
  }
  srv->tickets.erase(it);
  if (status != HK_OK) throw HkError(status, "%s", err.c_str());
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
