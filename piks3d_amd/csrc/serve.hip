// serve.hip -- continuous batching of independent solves on one GPU.
//
// The reference runs one solve per program (src/main.F90); a service that
// answers many permeability requests needs them to share the device.  This
// server keeps the reference's per-solve loop (iterate, chkConverge every
// chkConvergeStep iterations, stop at error <= eps or maxStep) and runs many
// such loops at once:
//   * submit copies the request and queues it for host preprocessing
//     (velocity grid, image labelling, sweep schedule) on a pool of threads;
//   * the scheduler thread admits preprocessed requests in arrival order
//     whenever a slot is free and their exact device footprint fits the
//     memory budget, allocates them stream-ordered on the slot's stream,
//     captures one iteration as a HIP graph, and enqueues iterations up to
//     the next convergence check followed by an asynchronous copy of the
//     check plane into pinned memory;
//   * it polls the slots' events, evaluates each check on the host in the
//     reference's order, and either enqueues the next chunk or retires the
//     request, releasing its slot and memory to the next one in line.
// Requests that arrive while others run are admitted as soon as resources
// free up; their kernels overlap on the device through the slots' streams.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/piks3d.h"
#include "context.h"

namespace piks {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
  int64_t ticket = 0;
  piks_params prm{};
  std::vector<uint8_t> image;
  int32_t max_step = 0;
  double eps = 0.0;
  int status = PIKS_REQ_QUEUED;
  int error_code = 0;
  std::string error_msg;
  Ctx* ctx = nullptr;
  size_t bytes = 0;
  int slot = -1;
  // solve state (src/main.F90:95-121)
  int step = 0, steps32 = 0, chunk = 0;
  double mass = 1.0, perm = 0.0, error = 1.0;
  bool converged = false;
  Clock::time_point t_submit, t_admit, t_done;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  double* plane = nullptr;  // pinned
  size_t plane_cap = 0;
  Request* req = nullptr;
};

struct Server {
  piks_serve_params prm{};
  size_t budget = 0, in_use = 0;
  std::mutex mu;
  std::condition_variable cv_prep, cv_sched, cv_done;
  std::deque<Request*> prep_q, ready_q;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  int64_t next_ticket = 1;
  int pending = 0;  // submitted, not finished
  int waiters = 0;  // threads inside piks_serve_wait / piks_serve_drain
  bool stop = false;
  std::vector<Slot> slots;
  std::vector<std::thread> workers;
  std::thread sched;
  std::string sched_error;
};

// Publish a request's end (under the lock).  A request ends once: the server's
// teardown may fail a request whose last chunk the scheduler is retiring.
void finish(Server& s, Request& r, int status) {
  if (r.status == PIKS_REQ_DONE || r.status == PIKS_REQ_FAILED) return;
  r.status = status;
  r.t_done = Clock::now();
  --s.pending;
  s.cv_done.notify_all();
}

void prep_worker(Server* s) {
  // the workers preprocess different requests at once: share the host threads
  set_host_threads(std::max(1, host_threads() / std::max(1, s->prm.prep_threads)));
  for (;;) {
    Request* r = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_prep.wait(lk, [&] { return s->stop || !s->prep_q.empty(); });
      if (s->stop) return;
      r = s->prep_q.front();
      s->prep_q.pop_front();
    }
    Ctx* c = nullptr;
    const int rc = ctx_create_host(&r->prm, r->image.data(), (int64_t)r->image.size(), &c);
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc != 0) {
      r->error_code = rc;
      r->error_msg = piks_last_error();
      finish(*s, *r, PIKS_REQ_FAILED);
      continue;
    }
    c->use_graph = s->prm.use_graphs != 0;
    r->ctx = c;
    r->bytes = ctx_device_bytes(*c);
    std::vector<uint8_t>().swap(r->image);
    if (r->bytes > s->budget) {
      r->error_code = PIKS_E_MEM;
      r->error_msg = "request needs " + std::to_string(r->bytes >> 20) + " MiB, the server's budget is " +
                     std::to_string(s->budget >> 20) + " MiB";
      ctx_destroy(c);
      r->ctx = nullptr;
      finish(*s, *r, PIKS_REQ_FAILED);
      continue;
    }
    s->ready_q.push_back(r);
    s->cv_sched.notify_all();
  }
}

// Iterations up to the next check (or maxStep), then the check plane.
int enqueue_chunk(Slot& sl, Request& r) {
  Ctx& c = *r.ctx;
  const int chk = c.prm.chk_converge_step;
  r.chunk = std::min(chk - r.step % chk, r.max_step - r.step);
  if (int rc = ctx_enqueue(c, r.chunk)) return rc;
  if (int rc = ctx_enqueue_plane(c, sl.plane)) return rc;
  if (hipEventRecord(sl.done, sl.stream) != hipSuccess) return set_error(PIKS_E_HIP, "hipEventRecord failed");
  return 0;
}

// The scheduler thread alone touches the slots, the contexts of running
// requests and their solve state; s.mu guards the queues, the published
// result fields, the memory accounting and the request status.  Device work
// (allocation, upload, enqueue, teardown) runs without the lock, so that
// submit, poll, wait and the preparation workers never wait behind it.

// Release a failed running request's slot and memory (called without the lock).
void fail_request(Server& s, Slot& sl, Request& r, int rc) {
  const std::string msg = piks_last_error();
  ctx_destroy(r.ctx);
  std::lock_guard<std::mutex> lk(s.mu);
  r.ctx = nullptr;
  r.error_code = rc;
  r.error_msg = msg;
  s.in_use -= r.bytes;
  sl.req = nullptr;
  finish(s, r, PIKS_REQ_FAILED);
}

// A chunk finished: evaluate the check like the reference's loop (called
// without the lock).
void on_chunk_done(Server& s, Slot& sl, Request& r) {
  Ctx& c = *r.ctx;
  const int step = r.step + r.chunk;
  const int chk = c.prm.chk_converge_step;
  bool done = false, converged = false;
  double mass = r.mass, perm = r.perm, error = r.error;
  if (step % chk == 0) {
    double out[3];
    ctx_check(c, 0, ctx_mass(c, sl.plane), out);
    ctx_dc_policy(c, out[2]);  // FP32_REFINED: the next defect-correction blocks
    mass = out[0];
    perm = out[1];
    error = out[2];
    if (error <= r.eps) {
      converged = true;
      done = true;
    }
  }
  if (step >= r.max_step) done = true;
  {
    std::lock_guard<std::mutex> lk(s.mu);
    r.step = step;
    r.steps32 = (int)c.dc_inner;
    r.mass = mass;
    r.perm = perm;
    r.error = error;
    r.converged = converged;
  }
  if (!done) {
    if (int rc = enqueue_chunk(sl, r)) fail_request(s, sl, r, rc);
    return;
  }
  const double final_mass = c.mass;
  ctx_destroy(r.ctx);  // stream-ordered frees: no device-wide wait
  std::lock_guard<std::mutex> lk(s.mu);
  r.ctx = nullptr;
  r.mass = final_mass;
  s.in_use -= r.bytes;
  sl.req = nullptr;
  finish(s, r, PIKS_REQ_DONE);
}

// Device set-up of an admitted request on its slot (called without the lock).
void start_request(Server& s, Slot& sl, Request& r) {
  const size_t need = (size_t)r.ctx->g.Ny * r.ctx->g.Nz;
  if (need > sl.plane_cap) {
    if (sl.plane) (void)hipHostFree(sl.plane);
    sl.plane = nullptr;
    sl.plane_cap = 0;
    if (hipHostMalloc((void**)&sl.plane, need * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      set_error(PIKS_E_MEM, "pinned host allocation failed");
      return fail_request(s, sl, r, PIKS_E_MEM);
    }
    sl.plane_cap = need;
  }
  int rc = ctx_create_device(*r.ctx, sl.stream);
  if (rc == 0) rc = enqueue_chunk(sl, r);
  if (rc) fail_request(s, sl, r, rc);
}

void scheduler(Server* s) {
  if (hipSetDevice(s->prm.device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->sched_error = "hipSetDevice failed";
    return;
  }
  std::vector<std::pair<Slot*, Request*>> admit;
  for (;;) {
    admit.clear();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->stop) return;
      // admit in arrival order while a slot and the memory are free
      while (!s->ready_q.empty()) {
        Request* r = s->ready_q.front();
        Slot* free_slot = nullptr;
        for (auto& sl : s->slots)
          if (!sl.req) {
            free_slot = &sl;
            break;
          }
        if (!free_slot || s->in_use + r->bytes > s->budget) break;
        s->ready_q.pop_front();
        r->slot = (int)(free_slot - s->slots.data());
        r->t_admit = Clock::now();
        r->status = PIKS_REQ_RUNNING;
        s->in_use += r->bytes;
        free_slot->req = r;
        admit.emplace_back(free_slot, r);
      }
    }
    bool progressed = !admit.empty();
    for (auto& a : admit) start_request(*s, *a.first, *a.second);
    // retire or continue the running requests whose chunk has finished
    for (auto& sl : s->slots) {
      if (!sl.req) continue;
      const hipError_t q = hipEventQuery(sl.done);
      if (q == hipErrorNotReady) continue;
      progressed = true;
      if (q != hipSuccess) {
        set_error(PIKS_E_HIP, std::string("solve failed on the device: ") + hipGetErrorString(q));
        fail_request(*s, sl, *sl.req, PIKS_E_HIP);
        continue;
      }
      on_chunk_done(*s, sl, *sl.req);
    }
    if (!progressed) {
      bool running = false;
      for (auto& sl : s->slots) running |= sl.req != nullptr;
      std::unique_lock<std::mutex> lk(s->mu);
      if (running)
        s->cv_sched.wait_for(lk, std::chrono::microseconds(50));
      else
        s->cv_sched.wait(lk, [&] { return s->stop || !s->ready_q.empty(); });
    }
  }
}

void fill_result(const Request& r, piks_serve_result* out) {
  std::memset(out, 0, sizeof(*out));
  out->status = r.status;
  out->error_code = r.error_code;
  out->steps = r.step;
  out->steps32 = r.steps32;
  out->converged = r.converged ? 1 : 0;
  out->permeability = r.perm;
  out->mass = r.mass;
  out->error = r.error;
  out->device_bytes = (int64_t)r.bytes;
  if (r.status == PIKS_REQ_DONE || r.status == PIKS_REQ_FAILED) {
    const bool admitted = r.t_admit.time_since_epoch().count() != 0;
    out->queue_ms = ms_between(r.t_submit, admitted ? r.t_admit : r.t_done);
    out->run_ms = admitted ? ms_between(r.t_admit, r.t_done) : 0.0;
    out->total_ms = ms_between(r.t_submit, r.t_done);
  }
}

}  // namespace
}  // namespace piks

using namespace piks;

extern "C" {

int piks_serve_create(const piks_serve_params* sp, void** server_out) {
  if (!server_out) return set_error(PIKS_E_ARG, "null argument");
  *server_out = nullptr;
  std::unique_ptr<Server> s(new Server());
  if (sp) s->prm = *sp;
  if (s->prm.max_slots <= 0) s->prm.max_slots = 8;
  if (s->prm.prep_threads <= 0) s->prm.prep_threads = 4;
  if (!sp) s->prm.use_graphs = 1;
  if (s->prm.max_slots > 64) return set_error(PIKS_E_ARG, "at most 64 slots");
  if (hipSetDevice(s->prm.device) != hipSuccess) return set_error(PIKS_E_HIP, "hipSetDevice failed");
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return set_error(PIKS_E_HIP, "hipMemGetInfo failed");
  s->budget = s->prm.memory_budget > 0 ? (size_t)s->prm.memory_budget : free_b / 10 * 9;
  if (s->budget > free_b) return set_error(PIKS_E_MEM, "memory budget exceeds the free device memory");
  s->slots.resize((size_t)s->prm.max_slots);
  for (auto& sl : s->slots) {
    if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess)
      return set_error(PIKS_E_HIP, "stream or event creation failed");
  }
  for (int q = 0; q < s->prm.prep_threads; ++q) s->workers.emplace_back(prep_worker, s.get());
  s->sched = std::thread(scheduler, s.get());
  *server_out = s.release();
  return 0;
}

int piks_serve_submit(void* server, const piks_params* p, const uint8_t* raw_image, int64_t image_bytes,
                      int32_t max_step, double eps, double refine_switch, int64_t* ticket_out) {
  if (!server || !p || !raw_image || image_bytes < 0 || !ticket_out)
    return set_error(PIKS_E_ARG, "null argument");
  if (max_step < 1) return set_error(PIKS_E_ARG, "max_step must be >= 1");
  if (p->chk_converge_step < 1) return set_error(PIKS_E_ARG, "chkConvergeStep must be >= 1");
  Server& s = *static_cast<Server*>(server);
  std::unique_ptr<Request> r(new Request());
  r->prm = *p;
  r->prm.device = s.prm.device;
  r->image.assign(raw_image, raw_image + image_bytes);
  r->max_step = max_step;
  r->eps = eps;
  (void)refine_switch;  // FP32_REFINED requests run the defect correction
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(s.mu);
  r->ticket = s.next_ticket++;
  *ticket_out = r->ticket;
  s.prep_q.push_back(r.get());
  ++s.pending;
  s.reqs[r->ticket] = std::move(r);
  s.cv_prep.notify_one();
  return 0;
}

int piks_serve_poll(void* server, int64_t ticket, piks_serve_result* out) {
  if (!server || !out) return set_error(PIKS_E_ARG, "null argument");
  Server& s = *static_cast<Server*>(server);
  std::lock_guard<std::mutex> lk(s.mu);
  auto it = s.reqs.find(ticket);
  if (it == s.reqs.end()) return set_error(PIKS_E_ARG, "unknown ticket " + std::to_string(ticket));
  fill_result(*it->second, out);
  return 0;
}

int piks_serve_wait(void* server, int64_t ticket, piks_serve_result* out) {
  if (!server || !out) return set_error(PIKS_E_ARG, "null argument");
  Server& s = *static_cast<Server*>(server);
  std::unique_lock<std::mutex> lk(s.mu);
  auto it = s.reqs.find(ticket);
  if (it == s.reqs.end()) return set_error(PIKS_E_ARG, "unknown ticket " + std::to_string(ticket));
  Request& r = *it->second;
  ++s.waiters;
  s.cv_done.wait(lk, [&] { return r.status == PIKS_REQ_DONE || r.status == PIKS_REQ_FAILED; });
  fill_result(r, out);
  --s.waiters;
  s.cv_done.notify_all();
  return 0;
}

int piks_serve_drain(void* server) {
  if (!server) return set_error(PIKS_E_ARG, "null argument");
  Server& s = *static_cast<Server*>(server);
  std::unique_lock<std::mutex> lk(s.mu);
  ++s.waiters;
  s.cv_done.wait(lk, [&] { return s.pending == 0; });
  --s.waiters;
  s.cv_done.notify_all();
  return 0;
}

int piks_serve_destroy(void* server) {
  if (!server) return 0;
  Server* s = static_cast<Server*>(server);
  {
    // every request that has not finished fails with PIKS_E_STATE, which
    // releases the threads blocked in wait and drain; they leave before the
    // server is freed
    std::unique_lock<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& kv : s->reqs) {
      Request& r = *kv.second;
      if (r.status == PIKS_REQ_DONE || r.status == PIKS_REQ_FAILED) continue;
      r.error_code = PIKS_E_STATE;
      r.error_msg = "the server was destroyed";
      finish(*s, r, PIKS_REQ_FAILED);
    }
    s->cv_done.notify_all();
    s->cv_done.wait(lk, [&] { return s->waiters == 0; });
  }
  s->cv_prep.notify_all();
  s->cv_sched.notify_all();
  for (auto& t : s->workers) t.join();
  if (s->sched.joinable()) s->sched.join();
  // teardown is best effort: every release is attempted whatever the others
  // return
  (void)hipSetDevice(s->prm.device);
  for (auto& sl : s->slots) {
    if (sl.req && sl.req->ctx) {
      (void)hipStreamSynchronize(sl.stream);
      ctx_destroy(sl.req->ctx);
      sl.req->ctx = nullptr;
    }
  }
  for (auto& kv : s->reqs)
    if (kv.second->ctx) ctx_destroy(kv.second->ctx);
  for (auto& sl : s->slots) {
    if (sl.stream) {
      (void)hipStreamSynchronize(sl.stream);
      (void)hipStreamDestroy(sl.stream);
    }
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.plane) (void)hipHostFree(sl.plane);
  }
  delete s;
  return 0;
}

}  // extern "C"
