// Continuous-batching server over bh_step_batch (no reference counterpart:
// the reference steps one system per call, main.cc:199-246; this is that
// leapfrog for many independent systems sharing the GPU).
//
// State. Active systems sit back to back in one device buffer (the step's
// input, `in`); the step writes each system's Morton-ordered result to the
// same range of `out`. Every system advances by one force evaluation per
// server step, at its own point of the leapfrog schedule: (dt/2, dt) first,
// then (dt, dt), and (dt/2, 0) as its last of steps + 1 evaluations.
//
// Admission. Requests queue in submission order; at each step boundary the
// head of the queue is admitted while the active systems' footprints
// (bh_server_footprint) stay within the memory cap and the system count
// within its limit. A request admitted while others are active is copied
// (from pinned staging, on a copy stream) into `out` behind the active
// range during the running step, so its upload overlaps their compute, and
// it joins the batch from the next step on.
//
// Retirement. A system whose last evaluation ran is copied to its result
// buffer; the survivors and the newly staged systems are then packed into
// `in` for the next step (or, with nothing retired, `out` simply becomes
// `in`).
//
// Each system's arithmetic is independent of the others in the batch, so a
// request's result is bit-identical to stepping it alone.
//
// Failures. A request with a non-finite value is refused at submission. If
// the batched step fails anyway, every active system is stepped alone from
// the same input (which the step never writes): the ones that succeed have
// exactly the result the batch would have given them and carry on, the ones
// that fail complete with the error as their status. If the buffers cannot
// grow for newly admitted requests, those go back to the head of the queue.
// Either way no request is lost and no footprint leaks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "bh_internal.h"

namespace {

struct Req {
  uint64_t ticket = 0;
  bh_request r{};
  bh_particle* staged = nullptr;  // pinned copy of the initial particles
  bh_particle* result = nullptr;  // caller's destination
  double submit_s = 0.0, admit_s = 0.0;
  uint64_t footprint = 0;
};

struct Active {
  Req q;
  uint32_t phase = 0;  // index of its next force evaluation, 0..steps
  uint64_t off = 0;    // particle offset in the current input buffer
};

struct Server {
  bh_ctx* ctx = nullptr;
  int device = 0;
  uint64_t cap = 0, used = 0;
  uint32_t max_systems = 0;
  uint64_t next_ticket = 1;
  hipStream_t copy = nullptr;
  hipEvent_t staged_ev = nullptr;
  bh_particle* buf[2] = {nullptr, nullptr};
  uint64_t buf_cap = 0;  // particles per buffer
  int in = 0;            // buf[in] is the next step's input
  std::deque<Req> pending;
  std::vector<Active> active;
  std::vector<bh_completion> done;
  std::chrono::steady_clock::time_point t0;
  std::string err;
};

double now_s(const Server* s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

int fail(Server* s, int code, const std::string& msg) {
  s->err = msg;
  return code;
}

#define SRV_HIP(s, expr)                                                                          \
  do {                                                                                            \
    hipError_t e_ = (expr);                                                                       \
    if (e_ != hipSuccess) return fail((s), BH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

uint64_t active_particles(const Server* s) {
  uint64_t n = 0;
  for (const auto& a : s->active) n += a.q.r.n;
  return n;
}

// Grows both buffers to hold `need` particles, keeping the first `keep`
// particles of the input buffer.
int grow(Server* s, uint64_t need, uint64_t keep) {
  if (need <= s->buf_cap) return BH_OK;
  const uint64_t want = std::min<uint64_t>(std::max<uint64_t>(need + need / 4, 1u << 16), bhk::kMaxN);
  SRV_HIP(s, hipStreamSynchronize(s->copy));
  SRV_HIP(s, hipDeviceSynchronize());
  bh_particle* nb[2] = {nullptr, nullptr};
  for (int k = 0; k < 2; ++k) {
    if (hipMalloc(reinterpret_cast<void**>(&nb[k]), want * sizeof(bh_particle)) != hipSuccess) {
      (void)hipGetLastError();
      if (nb[0]) (void)hipFree(nb[0]);
      return fail(s, BH_ERR_OOM, "server buffers: device allocation of " + std::to_string(want) + " particles failed");
    }
  }
  if (keep) SRV_HIP(s, hipMemcpy(nb[0], s->buf[s->in], keep * sizeof(bh_particle), hipMemcpyDeviceToDevice));
  for (int k = 0; k < 2; ++k)
    if (s->buf[k]) (void)hipFree(s->buf[k]);
  s->buf[0] = nb[0];
  s->buf[1] = nb[1];
  s->in = 0;
  s->buf_cap = want;
  return BH_OK;
}

void kick_drift(const Active& a, double* kick, double* drift) {
  const double dt = a.q.r.dt;
  if (a.phase == 0) {
    *kick = dt / 2;
    *drift = dt;
  } else if (a.phase < a.q.r.steps) {
    *kick = dt;
    *drift = dt;
  } else {
    *kick = dt / 2;
    *drift = 0.0;
  }
}

bh_system system_of(const Active& a) {
  bh_system y{};
  y.n = a.q.r.n;
  y.fp = a.q.r.fp;
  y.max_leaf = a.q.r.max_leaf;
  y.group_size = a.q.r.group_size;
  y.target_size = a.q.r.target_size;
  y.walk_tile = 0;  // the context's
  kick_drift(a, &y.kick_dt, &y.drift_dt);
  return y;
}

}  // namespace

extern "C" {

// Device bytes per request: per particle the server's two state copies
// (128 B), sort keys and indices (24), the walk copy (32), build and target
// scratch (~20), the group sums (24) and marks (8), and per walk group (at
// least one per 128 particles, ~40 particles on average) its deferred list
// (2 KB) and deep stack frames (2.5 KB): ~350 B; per tree node ~330 B, with
// nodes ~ n min(1.6, 4 / max_leaf) (measured on Plummer, Hernquist and
// uniform trees).
uint64_t bh_server_footprint(const bh_request* r) {
  if (!r) return 0;
  const double per_node = 4.0 / std::max(1u, r->max_leaf);
  const double nodes = static_cast<double>(r->n) * std::min(1.6, per_node) + 64.0;
  return static_cast<uint64_t>(static_cast<double>(r->n) * 350.0 + nodes * 330.0);
}

int bh_server_create(bh_server** out, bh_ctx* ctx, uint64_t mem_cap_bytes, uint32_t max_systems) {
  if (!out || !ctx || max_systems == 0 || max_systems > bhk::kMaxSystems) return BH_ERR_INVALID;
  *out = nullptr;
  Server* s = new Server();
  s->ctx = ctx;
  s->device = bhk::ctx_device(ctx);
  s->cap = mem_cap_bytes;
  s->max_systems = max_systems;
  s->t0 = std::chrono::steady_clock::now();
  if (hipSetDevice(s->device) != hipSuccess || hipStreamCreateWithFlags(&s->copy, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&s->staged_ev, hipEventDisableTiming) != hipSuccess) {
    (void)hipGetLastError();
    if (s->copy) (void)hipStreamDestroy(s->copy);
    delete s;
    return BH_ERR_HIP;
  }
  *out = reinterpret_cast<bh_server*>(s);
  return BH_OK;
}

void bh_server_destroy(bh_server* h) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return;
  (void)hipSetDevice(s->device);
  (void)hipStreamSynchronize(s->copy);
  for (auto& q : s->pending)
    if (q.staged) (void)hipHostFree(q.staged);
  for (auto& a : s->active)
    if (a.q.staged) (void)hipHostFree(a.q.staged);
  for (int k = 0; k < 2; ++k)
    if (s->buf[k]) (void)hipFree(s->buf[k]);
  (void)hipEventDestroy(s->staged_ev);
  (void)hipStreamDestroy(s->copy);
  delete s;
}

const char* bh_server_last_error(const bh_server* h) {
  const Server* s = reinterpret_cast<const Server*>(h);
  return s ? s->err.c_str() : "null server";
}

double bh_server_clock(const bh_server* h) { return h ? now_s(reinterpret_cast<const Server*>(h)) : 0.0; }

int bh_server_submit(bh_server* h, const bh_particle* particles, bh_particle* result, const bh_request* r,
                     uint64_t* ticket) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return BH_ERR_INVALID;
  if (!r || !result || (r->n && !particles) || !ticket) return fail(s, BH_ERR_INVALID, "null argument");
  if (r->n == 0) return fail(s, BH_ERR_INVALID, "a request needs at least one particle");
  if (r->steps == 0) return fail(s, BH_ERR_INVALID, "a request needs at least one leapfrog step");
  if (!std::isfinite(r->dt)) return fail(s, BH_ERR_INVALID, "dt must be finite");
  if (!(r->fp.theta >= 0.0) || !std::isfinite(r->fp.theta) || !(r->fp.eps2 >= 0.0) || !std::isfinite(r->fp.eps2) ||
      !std::isfinite(r->fp.G))
    return fail(s, BH_ERR_INVALID, "invalid force parameters");
  if (r->max_leaf > bhk::kMaxLeafCap || r->group_size == 0)
    return fail(s, BH_ERR_INVALID, "max_leaf must be <= 1024 and group_size >= 1");
  if (r->n > bhk::kMaxN) return fail(s, BH_ERR_TOO_LARGE, "request exceeds the per-batch particle limit");
  for (uint64_t i = 0; i < r->n; ++i) {
    const bh_particle& q = particles[i];
    if (!(std::isfinite(q.mass) && std::isfinite(q.pos[0]) && std::isfinite(q.pos[1]) && std::isfinite(q.pos[2]) &&
          std::isfinite(q.vel[0]) && std::isfinite(q.vel[1]) && std::isfinite(q.vel[2])))
      return fail(s, BH_ERR_NONFINITE, "particle " + std::to_string(i) + " has a non-finite mass, position or velocity");
  }
  const uint64_t fp = bh_server_footprint(r);
  if (fp > s->cap) return fail(s, BH_ERR_TOO_LARGE, "request footprint exceeds the server's memory cap");
  Req q;
  q.ticket = s->next_ticket++;
  q.r = *r;
  q.result = result;
  q.footprint = fp;
  q.submit_s = now_s(s);
  if (hipSetDevice(s->device) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&q.staged), r->n * sizeof(bh_particle), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    return fail(s, BH_ERR_OOM, "pinned staging allocation failed");
  }
  std::memcpy(q.staged, particles, r->n * sizeof(bh_particle));
  s->pending.push_back(q);
  *ticket = q.ticket;
  return BH_OK;
}

int bh_server_step(bh_server* h, bh_server_stats* stats) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s) return BH_ERR_INVALID;
  if (stats) *stats = bh_server_stats{};
  SRV_HIP(s, hipSetDevice(s->device));
  const uint64_t n_act = active_particles(s);

  // admission: the head of the queue while it fits
  std::vector<Active> admitted;
  uint64_t n_new = 0;
  while (!s->pending.empty() && s->active.size() + admitted.size() < s->max_systems) {
    const Req& q = s->pending.front();
    if (s->used + q.footprint > s->cap || n_act + n_new + q.r.n > bhk::kMaxN) break;
    Active a;
    a.q = q;
    a.q.admit_s = now_s(s);
    a.off = n_act + n_new;
    n_new += q.r.n;
    s->used += q.footprint;
    admitted.push_back(a);
    s->pending.pop_front();
  }
  const bool idle = s->active.empty();
  if (n_new) {
    const int rc = grow(s, n_act + n_new, n_act);
    if (rc != BH_OK) {
      // back to the head of the queue, in their order, with their footprint released
      for (auto it = admitted.rbegin(); it != admitted.rend(); ++it) {
        s->used -= it->q.footprint;
        it->q.admit_s = 0.0;
        s->pending.push_front(it->q);
      }
      return rc;
    }
    // idle server: stage straight into the input and step now; otherwise
    // behind the active range of the output, overlapping this step
    bh_particle* dst = s->buf[idle ? s->in : 1 - s->in];
    for (const auto& a : admitted)
      SRV_HIP(s, hipMemcpyAsync(dst + a.off, a.q.staged, a.q.r.n * sizeof(bh_particle), hipMemcpyHostToDevice, s->copy));
    SRV_HIP(s, hipEventRecord(s->staged_ev, s->copy));
  }
  if (idle) {
    if (n_new) SRV_HIP(s, hipEventSynchronize(s->staged_ev));
    for (auto& a : admitted) {
      (void)hipHostFree(a.q.staged);
      a.q.staged = nullptr;
      s->active.push_back(a);
    }
    admitted.clear();
  }
  if (stats) {
    stats->admitted = static_cast<uint32_t>(admitted.size() + (idle ? s->active.size() : 0));
    stats->pending = static_cast<uint32_t>(s->pending.size());
  }
  if (s->active.empty()) return BH_OK;

  // one force evaluation of every active system
  std::vector<bh_system> table(s->active.size());
  for (size_t k = 0; k < s->active.size(); ++k) table[k] = system_of(s->active[k]);
  bh_particle* in = s->buf[s->in];
  bh_particle* out = s->buf[1 - s->in];
  bh_step_timing t{};
  const double t_step = now_s(s);
  std::vector<int> failed(s->active.size(), BH_OK);
  const int rc = bh_step_batch(s->ctx, in, out, table.data(), static_cast<uint32_t>(table.size()), &t);
  if (rc != BH_OK) {
    // each system alone, from the same (unwritten) input
    s->err = std::string("batched step: ") + bh_last_error(s->ctx);
    for (size_t k = 0; k < s->active.size(); ++k) {
      const Active& a = s->active[k];
      const bh_system& y = table[k];
      failed[k] = bh_step_device(s->ctx, in + a.off, out + a.off, y.n, &y.fp, y.max_leaf, y.group_size, y.target_size,
                                 y.kick_dt, y.drift_dt, nullptr);
    }
  }
  SRV_HIP(s, hipEventSynchronize(s->staged_ev));
  const double t_end = now_s(s);
  for (auto& a : admitted) {
    (void)hipHostFree(a.q.staged);
    a.q.staged = nullptr;
  }
  uint64_t stepped = 0;
  for (const auto& a : s->active) stepped += a.q.r.n;
  const uint32_t stepped_systems = static_cast<uint32_t>(s->active.size());

  // retire finished systems, pack the rest (then the new ones) into `in`
  std::vector<Active> keep;
  keep.reserve(s->active.size() + admitted.size());
  uint32_t completed = 0;
  for (size_t k = 0; k < s->active.size(); ++k) {
    Active& a = s->active[k];
    if (failed[k] == BH_OK) a.phase += 1;
    if (a.phase > a.q.r.steps || failed[k] != BH_OK) {
      if (failed[k] == BH_OK)
        SRV_HIP(s, hipMemcpy(a.q.result, out + a.off, a.q.r.n * sizeof(bh_particle), hipMemcpyDeviceToHost));
      bh_completion c{};
      c.ticket = a.q.ticket;
      c.submit_s = a.q.submit_s;
      c.admit_s = a.q.admit_s;
      c.done_s = now_s(s);
      c.evaluations = a.phase;
      c.status = failed[k];
      s->done.push_back(c);
      s->used -= a.q.footprint;
      ++completed;
    } else {
      keep.push_back(a);
    }
  }
  for (auto& a : admitted) keep.push_back(a);
  if (completed == 0) {
    s->in = 1 - s->in;  // survivors keep their offsets, the new ones follow
  } else {
    uint64_t pos = 0;
    for (auto& a : keep) {
      SRV_HIP(s, hipMemcpyAsync(in + pos, out + a.off, a.q.r.n * sizeof(bh_particle), hipMemcpyDeviceToDevice, s->copy));
      a.off = pos;
      pos += a.q.r.n;
    }
    SRV_HIP(s, hipStreamSynchronize(s->copy));
  }
  s->active.swap(keep);
  if (stats) {
    stats->active_systems = static_cast<uint32_t>(s->active.size());
    stats->active_particles = active_particles(s);
    stats->stepped_particles = stepped;
    stats->stepped_systems = stepped_systems;
    stats->completed = completed;
    stats->step_ms = 1e3 * (t_end - t_step);
    stats->used_bytes = s->used;
  }
  return BH_OK;
}

int bh_server_poll(bh_server* h, bh_completion* out, uint32_t max, uint32_t* count) {
  Server* s = reinterpret_cast<Server*>(h);
  if (!s || !count || (max && !out)) return BH_ERR_INVALID;
  const uint32_t k = static_cast<uint32_t>(std::min<size_t>(max, s->done.size()));
  std::copy(s->done.begin(), s->done.begin() + k, out);
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *count = k;
  return BH_OK;
}

int bh_server_pending(const bh_server* h, uint32_t* pending, uint32_t* active) {
  const Server* s = reinterpret_cast<const Server*>(h);
  if (!s) return BH_ERR_INVALID;
  if (pending) *pending = static_cast<uint32_t>(s->pending.size());
  if (active) *active = static_cast<uint32_t>(s->active.size());
  return BH_OK;
}

}  // extern "C"
