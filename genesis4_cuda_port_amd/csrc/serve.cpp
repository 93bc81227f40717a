// serve.cpp - a request loop that keeps many runs resident on one GPU.
//
// Requests arrive at any time (g4h_server_submit).  One scheduler thread
// owns every HIP call: it admits queued requests while their estimated HBM
// footprint fits the budget and a stream lane is free, builds their state,
// enqueues all their steps on the lane's stream without waiting, and polls
// the lanes for completion; finished runs give their diagnostics back to the
// host and release their memory, which admits the next requests.  Work of
// different lanes overlaps on the device.  Steady-state requests that share
// a lattice, precision and solver and are queued together run as one batch
// (g4h_batch: one launch per step for all of them); a time-dependent request
// is one fused run (g4h_run) whose beam is generated and whose seed field
// is filled on the device (g4h_beam_generate, g4h_field_fill_planes): the
// scheduler thread never builds host arrays of a request's size, which
// took ~1 s of host time for the largest requests and held every other lane
// back meanwhile.  Latency is measured from submission to the moment its
// records are on the host.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "kernels.hpp"

namespace g4 {
void set_last_error(const std::string& m);
}

namespace {

using Clock = std::chrono::steady_clock;

double now_s() {
  return std::chrono::duration<double>(Clock::now().time_since_epoch()).count();
}

struct Req {
  int64_t id = 0;
  g4h_request q{};
  std::vector<g4h_step> steps;
  std::vector<double> aw;
  uint64_t bytes = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  int status = -1;  // -1 queued/running, else g4h_status
  std::string error;
  int nrec = 0;
  std::vector<double> braw, fraw;
};

// HBM estimate of a request's resident state (beam, field ring, source
// planes, FP64 ADI intermediate and the x sweep's output plane set,
// diagnostics history).
uint64_t estimate_bytes(const g4h_request& q, int nsteps) {
  const uint64_t np = static_cast<uint64_t>(q.nslice) * q.npart;
  const uint64_t cells = static_cast<uint64_t>(q.nslice) * q.ngrid * q.ngrid;
  const uint64_t sp = q.precision == G4H_FP64 ? 48 : 28;
  const uint64_t sc = q.precision == G4H_FP64 ? 16 : 8;
  int nout = 0;
  for (int i = 0; i < nsteps; ++i) nout += q.steps[i].outstep ? 1 : 0;
  const uint64_t hist =
      static_cast<uint64_t>(1 + nout) * q.nslice * (G4H_BEAM_RAW(q.nharm) + G4H_FIELD_RAW) * 8;
  // a steady-state request in a batch also holds two parameter blocks per
  // step (its group's table and the batch's per-step table of all runs)
  const uint64_t tables =
      q.kind == 0 ? 2 * static_cast<uint64_t>(nsteps) * sizeof(g4::ParticleParams) : 0;
  return np * sp + cells * (2 * sc + (q.solver == G4H_SOLVER_ADI ? 16 + sc : 0)) + hist + tables +
         (1u << 20);
}

struct Lane {
  g4h_context* ctx = nullptr;
  std::vector<std::shared_ptr<Req>> reqs;  // running on this lane
  g4h_batch* batch = nullptr;
  g4h_beam* beam = nullptr;
  g4h_field* field = nullptr;
  g4h_run* run = nullptr;
  hipEvent_t done_ev = nullptr;
  bool busy = false;
  uint64_t bytes = 0;
};

}  // namespace

struct g4h_server {
  int device = 0;
  uint64_t budget = 0;
  std::vector<Lane> lanes;
  std::thread worker;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Req>> queue;
  std::map<int64_t, std::shared_ptr<Req>> all;
  int64_t next_id = 1;
  bool stop = false;
  uint64_t in_use = 0;
  // statistics over completed requests
  std::vector<double> latencies;
  double first_submit = -1, last_end = 0, work = 0;
  std::string fatal;

  void loop();
  bool admit(Lane& lane);
  void finish(Lane& lane);
  void fail_lane(Lane& lane, const std::string& msg, int status);
  void clear_lane(Lane& lane);
};

namespace {

void ok(g4h_status s) {
  if (s != G4H_OK) throw g4::Error(s, g4h_last_error());
}

bool same_batch(const Req& a, const Req& b) {
  const g4h_request &x = a.q, &y = b.q;
  return x.kind == 0 && y.kind == 0 && x.precision == y.precision && x.solver == y.solver &&
         x.lambda0 == y.lambda0 && x.nharm == y.nharm && a.steps.size() == b.steps.size() &&
         std::memcmp(a.steps.data(), b.steps.data(), a.steps.size() * sizeof(g4h_step)) == 0;
}

}  // namespace

void g4h_server::clear_lane(Lane& lane) {
  if (lane.batch) g4h_batch_destroy(lane.batch);
  if (lane.run) g4h_run_destroy(lane.run);
  if (lane.field) g4h_field_destroy(lane.field);
  if (lane.beam) g4h_beam_destroy(lane.beam);
  lane.batch = nullptr;
  lane.run = nullptr;
  lane.field = nullptr;
  lane.beam = nullptr;
  lane.reqs.clear();
  lane.busy = false;
}

void g4h_server::fail_lane(Lane& lane, const std::string& msg, int status) {
  {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& r : lane.reqs) {
      r->status = status;
      r->error = msg;
      r->t_end = now_s();
    }
    in_use -= lane.bytes;
    lane.bytes = 0;
  }
  clear_lane(lane);
  cv_done.notify_all();
}

// Take the head of the queue (and, for a steady-state head, the queued
// requests that can share its batch) if it fits the budget; build and
// enqueue it on `lane`.  Returns false when nothing was admitted.
bool g4h_server::admit(Lane& lane) {
  std::vector<std::shared_ptr<Req>> take;
  uint64_t bytes = 0;
  {
    std::lock_guard<std::mutex> lk(mu);
    if (queue.empty()) return false;
    auto head = queue.front();
    if (head->bytes > budget) {  // can never fit
      queue.pop_front();
      head->status = G4H_ERR_OUT_OF_MEMORY;
      head->error = "request exceeds the server's HBM budget";
      head->t_end = now_s();
      cv_done.notify_all();
      return true;
    }
    if (in_use + head->bytes > budget) return false;
    take.push_back(head);
    bytes = head->bytes;
    queue.pop_front();
    if (head->q.kind == 0) {
      for (auto it = queue.begin(); it != queue.end();) {
        if (same_batch(*head, **it) && in_use + bytes + (*it)->bytes <= budget) {
          bytes += (*it)->bytes;
          take.push_back(*it);
          it = queue.erase(it);
        } else {
          ++it;
        }
      }
    }
    in_use += bytes;
  }
  lane.reqs = take;
  lane.bytes = bytes;
  lane.busy = true;
  const double t0 = now_s();
  for (auto& r : take) r->t_start = t0;
  try {
    const Req& h = *take.front();
    const g4h_request& q = h.q;
    const int nsteps = static_cast<int>(h.steps.size());
    if (q.kind == 0) {
      std::vector<g4h_batch_run> runs(take.size());
      for (size_t i = 0; i < take.size(); ++i) {
        const g4h_request& r = take[i]->q;
        g4h_batch_run& b = runs[i];
        std::memset(&b, 0, sizeof(b));
        b.npart = r.npart;
        b.nbins = r.nbins;
        b.ngrid = r.ngrid;
        b.shotnoise = r.shotnoise;
        b.seed = r.seed;
        b.gridmax = r.gridmax;
        b.current = r.current;
        b.gamma0 = r.gamma0;
        b.espread_rel = r.espread_rel;
        b.emitn_x = b.emitn_y = r.emitn;
        b.betax = b.betay = r.beta;
        b.seed_power = r.seed_power;
        b.seed_w0 = r.seed_w0;
        b.aw = take[i]->aw.empty() ? nullptr : take[i]->aw.data();
      }
      ok(g4h_batch_create(lane.ctx, q.precision, q.solver, static_cast<int>(runs.size()),
                          runs.data(), h.steps.data(), nsteps, q.lambda0, q.nharm, &lane.batch));
      ok(g4h_batch_advance(lane.batch, nsteps, nullptr));
    } else {
      // beam generated and field seeded on the device: no host arrays of
      // the request's size
      g4h_beam_spec s{};
      s.gamma0 = q.gamma0;
      s.espread_rel = q.espread_rel;
      s.chirp_rel = q.chirp_rel;
      s.emitn_x = s.emitn_y = q.emitn;
      s.betax = s.betay = q.beta;
      s.slicelength = q.lambda0 * q.sample;
      s.nslice = q.nslice;
      s.npart = q.npart;
      s.nbins = q.nbins;
      s.shotnoise = q.shotnoise;
      s.seed = q.seed;
      std::vector<double> cur(static_cast<size_t>(q.nslice), q.current);
      std::vector<int32_t> off(static_cast<size_t>(q.nslice) + 1);
      for (int i = 0; i <= q.nslice; ++i) off[static_cast<size_t>(i)] = i * q.npart;
      ok(g4h_beam_create(lane.ctx, q.precision, q.nslice, off.data(), cur.data(), q.gamma0,
                         &lane.beam));
      ok(g4h_beam_generate(lane.beam, &s, cur.data()));
      std::vector<double> re(static_cast<size_t>(q.ngrid) * q.ngrid, 0.0), im(re.size(), 0.0);
      if (q.seed_power > 0)
        ok(g4h_seed_gaussian(1, q.ngrid, q.gridmax, q.lambda0, q.seed_power, q.seed_w0,
                             re.data(), im.data()));
      ok(g4h_field_create(lane.ctx, q.precision, q.nslice, q.ngrid, q.gridmax, q.lambda0,
                          q.lambda0 * q.sample, 1, &lane.field));
      ok(g4h_field_fill_planes(lane.field, re.data(), im.data(), 0));
      ok(g4h_field_init_solver(lane.field, q.solver, 0, 1.0, 1.0, 1.0, 0));
      g4h_run_config cfg{q.lambda0, q.sample, 1, q.periodic, q.nharm, 1};
      ok(g4h_run_create(lane.ctx, lane.beam, lane.field, h.steps.data(), nsteps, &cfg, &lane.run));
      ok(g4h_run_advance(lane.run, nsteps, nullptr));
    }
    G4_HIP(hipEventRecord(lane.done_ev, lane.ctx->stream));
  } catch (const g4::Error& e) {
    fail_lane(lane, e.what(), e.code);
  } catch (const std::exception& e) {
    fail_lane(lane, e.what(), G4H_ERR_DEVICE);
  }
  return true;
}

// A lane's work has completed: copy every request's records to the host.
void g4h_server::finish(Lane& lane) {
  try {
    if (lane.batch) {
      int nrec = 0;
      ok(g4h_batch_records(lane.batch, &nrec));
      for (size_t i = 0; i < lane.reqs.size(); ++i) {
        Req& r = *lane.reqs[i];
        r.nrec = nrec;
        r.braw.resize(static_cast<size_t>(nrec) * G4H_BEAM_RAW(r.q.nharm));
        r.fraw.resize(static_cast<size_t>(nrec) * G4H_FIELD_RAW);
        ok(g4h_batch_download(lane.batch, static_cast<int>(i), r.braw.data(), r.fraw.data()));
      }
    } else {
      Req& r = *lane.reqs.front();
      int nrec = 0;
      ok(g4h_run_records(lane.run, &nrec));
      r.nrec = nrec;
      r.braw.resize(static_cast<size_t>(nrec) * r.q.nslice * G4H_BEAM_RAW(r.q.nharm));
      r.fraw.resize(static_cast<size_t>(nrec) * r.q.nslice * G4H_FIELD_RAW);
      ok(g4h_run_download_diag(lane.run, r.braw.data(), r.fraw.data()));
    }
  } catch (const g4::Error& e) {
    fail_lane(lane, e.what(), e.code);
    return;
  }
  const double t = now_s();
  {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& r : lane.reqs) {
      r->t_end = t;
      r->status = G4H_OK;
      latencies.push_back(t - r->t_submit);
      work += static_cast<double>(r->q.nslice) * r->q.npart * static_cast<double>(r->steps.size());
      last_end = std::max(last_end, t);
    }
    in_use -= lane.bytes;
    lane.bytes = 0;
  }
  clear_lane(lane);
  cv_done.notify_all();
}

void g4h_server::loop() {
  (void)hipSetDevice(device);
  for (;;) {
    bool any_busy = false, progressed = false;
    for (Lane& lane : lanes) {
      if (lane.busy) {
        const hipError_t q = hipEventQuery(lane.done_ev);
        if (q == hipSuccess) {
          finish(lane);
          progressed = true;
        } else if (q != hipErrorNotReady) {
          fail_lane(lane, hipGetErrorString(q), G4H_ERR_DEVICE);
          progressed = true;
        } else {
          any_busy = true;
        }
      }
      if (!lane.busy && admit(lane)) progressed = true;
      any_busy = any_busy || lane.busy;
    }
    std::unique_lock<std::mutex> lk(mu);
    if (stop && queue.empty() && !any_busy) return;
    if (!progressed) {
      if (any_busy)
        cv_work.wait_for(lk, std::chrono::microseconds(200));
      else
        cv_work.wait(lk, [&] { return stop || !queue.empty(); });
    }
  }
}

extern "C" {

g4h_status g4h_server_create(int device, uint64_t hbm_budget, int nstreams, g4h_server** out) {
  G4_API_BEGIN
  g4::require(out && nstreams >= 1 && nstreams <= 16 && hbm_budget > 0, "bad server arguments");
  *out = nullptr;
  auto s = std::make_unique<g4h_server>();
  s->device = device;
  s->budget = hbm_budget;
  s->lanes.resize(static_cast<size_t>(nstreams));
  for (Lane& l : s->lanes) {
    ok(g4h_context_create(device, &l.ctx));
    G4_HIP(hipEventCreateWithFlags(&l.done_ev, hipEventDisableTiming));
  }
  g4h_server* raw = s.get();
  s->worker = std::thread([raw] { raw->loop(); });
  *out = s.release();
  G4_API_END
}

void g4h_server_destroy(g4h_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Lane& l : s->lanes) {
    s->clear_lane(l);
    if (l.done_ev) (void)hipEventDestroy(l.done_ev);
    g4h_context_destroy(l.ctx);
  }
  delete s;
}

g4h_status g4h_server_submit(g4h_server* s, const g4h_request* q, int64_t* id) {
  G4_API_BEGIN
  g4::require(s && q && id && q->steps && q->nsteps >= 1, "null argument or empty lattice");
  g4::require(q->kind == 0 || q->kind == 1, "kind is 0 (steady state) or 1 (time-dependent)");
  g4::require(q->kind == 1 || q->nslice == 1, "a steady-state request has one slice");
  g4::require(q->nslice >= 1 && q->npart >= 1 && q->nbins >= 1 && q->npart % q->nbins == 0,
              "bad beam size");
  g4::require(q->ngrid >= 3 && q->ngrid <= 1024 && q->gridmax > 0, "bad grid");
  g4::require(q->nharm >= 1 && q->nharm <= 8 && q->lambda0 > 0 && q->gamma0 > 1,
              "bad nharm, lambda0 or gamma0");
  auto r = std::make_shared<Req>();
  r->q = *q;
  r->steps.assign(q->steps, q->steps + q->nsteps);
  if (q->aw) r->aw.assign(q->aw, q->aw + q->nsteps);
  r->q.steps = nullptr;
  r->q.aw = nullptr;
  r->bytes = estimate_bytes(*q, q->nsteps);
  r->t_submit = now_s();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    if (s->first_submit < 0) s->first_submit = r->t_submit;
    s->all[r->id] = r;
    s->queue.push_back(r);
    *id = r->id;
  }
  s->cv_work.notify_all();
  G4_API_END
}

g4h_status g4h_server_wait(g4h_server* s, int64_t id, double timeout_s, g4h_result_info* info) {
  G4_API_BEGIN
  g4::require(s, "server is null");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  g4::require(it != s->all.end(), "unknown request id");
  auto r = it->second;
  const auto ready = [&] { return r->status >= 0; };
  if (timeout_s < 0)
    s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    throw g4::Error(G4H_ERR_NOT_INITIALIZED, "request not finished within the timeout");
  if (info) {
    info->status = r->status;
    info->nrecords = r->nrec;
    info->latency_s = r->t_end - r->t_submit;
    info->queued_s = r->t_start > 0 ? r->t_start - r->t_submit : r->t_end - r->t_submit;
  }
  if (r->status != G4H_OK) throw g4::Error(static_cast<g4h_status>(r->status), r->error);
  G4_API_END
}

g4h_status g4h_server_result(g4h_server* s, int64_t id, double* beam_raw, double* field_raw) {
  G4_API_BEGIN
  g4::require(s, "server is null");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  g4::require(it != s->all.end(), "unknown request id");
  const Req& r = *it->second;
  g4::require(r.status == G4H_OK, "request has not completed successfully");
  if (beam_raw) std::memcpy(beam_raw, r.braw.data(), r.braw.size() * sizeof(double));
  if (field_raw) std::memcpy(field_raw, r.fraw.data(), r.fraw.size() * sizeof(double));
  G4_API_END
}

g4h_status g4h_server_release(g4h_server* s, int64_t id) {
  G4_API_BEGIN
  g4::require(s, "server is null");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  g4::require(it != s->all.end(), "unknown request id");
  g4::require(it->second->status >= 0, "request still queued or running");
  s->all.erase(it);
  G4_API_END
}

g4h_status g4h_server_stats(g4h_server* s, g4h_server_stats_t* st) {
  G4_API_BEGIN
  g4::require(s && st, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  std::vector<double> l = s->latencies;
  std::sort(l.begin(), l.end());
  auto pct = [&](double p) {
    if (l.empty()) return 0.0;
    const double k = p * static_cast<double>(l.size() - 1);
    const size_t i = static_cast<size_t>(k);
    const double f = k - static_cast<double>(i);
    return i + 1 < l.size() ? l[i] * (1 - f) + l[i + 1] * f : l[i];
  };
  st->completed = static_cast<int64_t>(l.size());
  st->p50_s = pct(0.5);
  st->p99_s = pct(0.99);
  st->mean_s = 0;
  for (double v : l) st->mean_s += v / static_cast<double>(l.size());
  const double span = s->last_end - s->first_submit;
  st->particle_slice_steps_per_s = span > 0 ? s->work / span : 0.0;
  st->hbm_in_use = s->in_use;
  G4_API_END
}

}  // extern "C"
