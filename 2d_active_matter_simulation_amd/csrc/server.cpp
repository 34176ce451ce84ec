// Serving loop: continuous batching of simulation requests on one GPU.
//
// Requests (grid size, parameters, horizon, snapshot cadence, IC seed) arrive
// at given times.  The loop keeps every size class of the engine moving on
// its own: whenever a class's chunk of steps completes, its runs' step
// counts, snapshots, completions and NaN flags are handled and the class's
// next chunk is launched at once (k = the steps until the next snapshot or
// horizon of any of its runs, capped by max_tick_steps), while other classes
// are still computing.  Arrived requests are admitted in arrival order while
// the engine's HBM budget allows (a request that does not fit waits, and the
// ones behind it wait too) and join their class's next chunk.  Snapshots are
// converted to the reference layout on the class's stream and copied to
// pinned buffers from a reused pool on a copy stream, then handed to the
// callback; retired runs' device blocks are reused by later requests.
//
// Times: chunk completions come from timing events on the class streams
// (seconds since the start event), admissions from the host clock.  Per-step
// latency is the interval between a request's consecutive chunk completions
// divided by the chunk's steps (weight = steps); time to completion is the
// final chunk's completion minus arrival.  No reference counterpart: the
// reference runs one simulation per process.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "engine_impl.hpp"

using namespace af;

namespace {

double pct(std::vector<std::pair<double, long long>>& v, double q) {
  // weighted percentile of (value, weight) pairs
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  long long tot = 0;
  for (auto& x : v) tot += x.second;
  const double target = q * (double)tot;
  long long acc = 0;
  for (auto& x : v) {
    acc += x.second;
    if ((double)acc >= target) return x.first;
  }
  return v.back().first;
}

double pct_plain(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  size_t i = (size_t)(q * (double)(v.size() - 1) + 0.5);
  return v[std::min(i, v.size() - 1)];
}

struct Resident {
  int slot = -1;
  long long steps = 0;
  double last = 0;   // completion time of its last chunk (or its admission)
};

// A class's chunk in flight.
struct ClassWork {
  std::vector<int> reqs;   // resident requests of the class
  bool busy = false;
  int k = 0;
  std::vector<int> chunk;  // requests the chunk in flight covers
  hipEvent_t done = nullptr;
};

// A device staging buffer (reference layout) and its pinned host twin.
struct SnapBuf {
  size_t n = 0;            // doubles
  double* dev = nullptr;
  double* host = nullptr;
};

struct PendingSnap {
  int req;
  long long step;
  int nx, ny;
  double M[2];
  bool final_state;
  SnapBuf buf;
  hipEvent_t done = nullptr;
};

}  // namespace

extern "C" {

af_status af_serve(const af_request* reqs, int n, size_t hbm_budget, int max_tick_steps, af_snapshot_fn fn,
                   void* user, double** final_states, af_request_result* results, af_serve_stats* stats) {
  if (n < 0 || (n > 0 && (!reqs || !results)) || max_tick_steps < 1) return fail(AF_EINVAL, "bad arguments");
  for (int i = 0; i < n; ++i) {
    std::memset(&results[i], 0, sizeof(results[i]));
    results[i].admit_s = -1;
    results[i].finish_s = -1;
  }
  af_engine* eng = nullptr;
  af_status st = af_engine_create(hbm_budget, &eng);
  if (st != AF_OK) return st;
  hipStream_t copy_stream = nullptr;
  hipEvent_t t0ev = nullptr;
  if (hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&t0ev) != hipSuccess) {
    if (copy_stream) (void)hipStreamDestroy(copy_stream);
    af_engine_destroy(eng);
    return fail(AF_EDEVICE, "serving streams");
  }

  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  std::vector<std::vector<std::pair<double, long long>>> lat(n);
  std::vector<std::pair<double, long long>> lat_all;
  std::vector<Resident> res(n);
  std::map<ClassKey, ClassWork> work;
  std::deque<PendingSnap> snaps;
  std::multimap<size_t, SnapBuf> spare;   // snapshot buffers for reuse, by doubles
  std::vector<hipEvent_t> spare_events;
  std::vector<double> host_state;
  int next = 0, max_resident = 0, ticks = 0, resident = 0;
  size_t peak = 0;
  double gps = 0;

  // the start of the clock, on the host and on the device
  hipError_t he = hipEventRecord(t0ev, eng->stream);
  if (he == hipSuccess) he = hipEventSynchronize(t0ev);
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
  if (he != hipSuccess) st = hip_fail(he, "start event");

  auto new_event = [&](hipEvent_t* ev, bool timing) -> hipError_t {
    if (!timing && !spare_events.empty()) {
      *ev = spare_events.back();
      spare_events.pop_back();
      return hipSuccess;
    }
    return timing ? hipEventCreate(ev) : hipEventCreateWithFlags(ev, hipEventDisableTiming);
  };
  auto take_buf = [&](size_t ndoubles, SnapBuf* b) -> af_status {
    auto it = spare.find(ndoubles);
    if (it != spare.end()) {
      *b = it->second;
      spare.erase(it);
      return AF_OK;
    }
    b->n = ndoubles;
    if (hipMalloc(&b->dev, ndoubles * sizeof(double)) != hipSuccess) return fail(AF_ENOMEM, "snapshot staging");
    if (hipHostMalloc((void**)&b->host, ndoubles * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      (void)hipFree(b->dev);
      return fail(AF_ENOMEM, "snapshot pinned buffer");
    }
    return AF_OK;
  };

  // Hands completed copies to the callback / final states, in enqueue order.
  auto deliver = [&](bool wait_all) -> af_status {
    while (!snaps.empty()) {
      PendingSnap& s = snaps.front();
      hipError_t q = wait_all ? hipEventSynchronize(s.done) : hipEventQuery(s.done);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) return hip_fail(q, "snapshot copy");
      if (s.final_state) {
        if (final_states && final_states[s.req])
          std::memcpy(final_states[s.req], s.buf.host, s.buf.n * sizeof(double));
      } else {
        if (fn) fn(user, s.req, s.step, s.buf.host, s.M, s.nx, s.ny);
        results[s.req].snapshots += 1;
      }
      spare.emplace(s.buf.n, s.buf);
      spare_events.push_back(s.done);
      snaps.pop_front();
    }
    return AF_OK;
  };

  // The state of request q (resident in slot) at this point of its class's
  // stream: reference layout on the class stream, copy to pinned memory on
  // the copy stream.
  auto start_copy = [&](int q, const MeanState& ms, bool final_state) -> af_status {
    const Slot& sl = eng->slots[res[q].slot];
    Class* c = engine_class(eng, res[q].slot);
    PendingSnap s;
    s.req = q;
    s.step = res[q].steps;
    s.nx = sl.p.nx;
    s.ny = sl.p.ny;
    s.M[0] = ms.M[0];
    s.M[1] = ms.M[1];
    s.final_state = final_state;
    af_status a = take_buf((size_t)sl.g.ny * (sl.g.nx2 + 1) * 2, &s.buf);
    if (a != AF_OK) return a;
    hipEvent_t conv = nullptr;
    hipError_t e = sl.g.from_bricks(sl.base + sl.g.off_field[0], s.buf.dev, sl.p.nx, sl.p.ny, sl.g.by, sl.g.bx,
                                    c->stream);
    if (e == hipSuccess) e = new_event(&conv, false);
    if (e == hipSuccess) e = hipEventRecord(conv, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(copy_stream, conv, 0);
    if (e == hipSuccess) e = new_event(&s.done, false);
    if (e == hipSuccess)
      e = hipMemcpyAsync(s.buf.host, s.buf.dev, s.buf.n * sizeof(double), hipMemcpyDeviceToHost, copy_stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, copy_stream);
    if (conv) spare_events.push_back(conv);   // reusable: the copy stream already waits on it
    if (e != hipSuccess) {
      spare.emplace(s.buf.n, s.buf);
      return hip_fail(e, "snapshot");
    }
    snaps.push_back(s);
    return AF_OK;
  };

  auto leave_class = [&](int q) {
    const Slot& sl = eng->slots[res[q].slot];
    auto& v = work[sl.key].reqs;
    v.erase(std::remove(v.begin(), v.end(), q), v.end());
  };

  // Processes a completed chunk of class key.
  auto finish_chunk = [&](const ClassKey& key, ClassWork& w) -> af_status {
    float ms = 0;
    hipError_t e = hipEventElapsedTime(&ms, t0ev, w.done);
    if (e != hipSuccess) return hip_fail(e, "chunk time");
    const double tend = ms * 1e-3;
    w.busy = false;
    Class& c = eng->classes.at(key);
    for (int q : w.chunk) {
      Resident& r = res[q];
      const af_request& rq = reqs[q];
      const int k = w.k;
      r.steps += k;
      eng->slots[r.slot].steps += k;
      const double per = (tend - r.last) / (double)k;
      lat[q].push_back({per, k});
      lat_all.push_back({per, k});
      r.last = tend;
      gps += (double)rq.params.nx * rq.params.ny * (double)k;
      const MeanState m = c.mirror[eng->slots[r.slot].entry];
      if (m.nan_step >= 0) {   // the reference stops such a run (.cu:850-861): it fails alone
        results[q].status = AF_ENAN;
        leave_class(q);
        af_status a = engine_retire(eng, r.slot);
        if (a != AF_OK) return a;
        --resident;
        continue;
      }
      if (rq.snapshot_every > 0 && r.steps % rq.snapshot_every == 0) {
        af_status a = start_copy(q, m, false);
        if (a != AF_OK) return a;
      }
      if (r.steps >= rq.horizon) {
        af_status a = start_copy(q, m, true);
        if (a != AF_OK) return a;
        results[q].M[0] = m.M[0];
        results[q].M[1] = m.M[1];
        results[q].finish_s = tend;
        results[q].status = AF_OK;
        leave_class(q);
        a = engine_retire(eng, r.slot);   // after the copy on the class's stream
        if (a != AF_OK) return a;
        --resident;
      }
    }
    return AF_OK;
  };

  while (st == AF_OK && (next < n || resident > 0)) {
    bool progressed = false;
    // 1. completed chunks
    for (auto& kv : work) {
      ClassWork& w = kv.second;
      if (!w.busy) continue;
      hipError_t q = hipEventQuery(w.done);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess) { st = hip_fail(q, "chunk"); break; }
      st = finish_chunk(kv.first, w);
      if (st != AF_OK) break;
      progressed = true;
    }
    if (st != AF_OK) break;
    if ((st = deliver(false)) != AF_OK) break;
    engine_reclaim(eng);
    // 2. admission in arrival order
    const double tnow = now();
    while (next < n && reqs[order[next]].arrival_s <= tnow) {
      const int q = order[next];
      const af_request& rq = reqs[q];
      const size_t nc = (size_t)rq.params.ny * (rq.params.nx / 2 + 1);
      host_state.assign(2 * nc, 0.0);
      af_status s2 = af_ic_bandlimited(rq.params.nx, rq.params.ny, rq.seed, rq.ic_amp, rq.ic_kmax,
                                       host_state.data());
      const double M0[2] = {0.0, 0.0};
      int slot = -1;
      if (s2 == AF_OK) s2 = af_engine_add(eng, &rq.params, host_state.data(), M0, &slot);
      if (s2 == AF_ENOMEM && (resident > 0 || !eng->pending.empty())) break;   // wait for memory
      ++next;
      progressed = true;
      if (s2 != AF_OK) {
        results[q].status = s2;
        continue;
      }
      results[q].admit_s = now();
      res[q].slot = slot;
      res[q].steps = 0;
      res[q].last = results[q].admit_s;
      ++resident;
      work[eng->slots[slot].key].reqs.push_back(q);
      if (rq.horizon <= 0) {   // nothing to step: the initial state is the result
        MeanState m;
        std::memset(&m, 0, sizeof(m));
        if ((st = start_copy(q, m, true)) != AF_OK) break;
        results[q].finish_s = results[q].admit_s;
        results[q].status = AF_OK;
        leave_class(q);
        if ((st = engine_retire(eng, slot)) != AF_OK) break;
        --resident;
      }
    }
    if (st != AF_OK) break;
    max_resident = std::max(max_resident, resident);
    peak = std::max(peak, af_engine_bytes_in_use(eng));
    // 3. every idle class with runs starts its next chunk
    for (auto& kv : work) {
      ClassWork& w = kv.second;
      if (w.busy || w.reqs.empty()) continue;
      long long k = max_tick_steps;
      for (int q : w.reqs) {
        const af_request& rq = reqs[q];
        long long left = rq.horizon - res[q].steps;
        if (rq.snapshot_every > 0) left = std::min(left, rq.snapshot_every - res[q].steps % rq.snapshot_every);
        k = std::min(k, std::max(1LL, left));
      }
      Class& c = eng->classes.at(kv.first);
      if ((st = engine_enqueue(eng, c, (int)k)) != AF_OK) break;
      if (!w.done && (he = hipEventCreate(&w.done)) != hipSuccess) { st = hip_fail(he, "chunk event"); break; }
      if ((he = hipEventRecord(w.done, c.stream)) != hipSuccess) { st = hip_fail(he, "chunk event"); break; }
      w.busy = true;
      w.k = (int)k;
      w.chunk = w.reqs;
      ++ticks;
      progressed = true;
    }
    if (st != AF_OK) break;
    // 4. nothing to do right now: wait for the device or the next arrival
    if (!progressed) {
      bool busy = false;
      for (auto& kv : work) busy = busy || kv.second.busy;
      if (!busy && resident == 0 && next < n) {
        const double wait = reqs[order[next]].arrival_s - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.002)));
      } else {
        std::this_thread::yield();
      }
    }
  }
  if (st == AF_OK) st = deliver(true);
  const double wall = now();
  for (int i = 0; i < n; ++i) {
    results[i].step_p50_s = pct(lat[i], 0.5);
    results[i].step_p99_s = pct(lat[i], 0.99);
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    std::vector<double> ttc, qw;
    for (int i = 0; i < n; ++i) {
      if (results[i].finish_s >= 0) ttc.push_back(results[i].finish_s - reqs[i].arrival_s);
      if (results[i].admit_s >= 0) qw.push_back(results[i].admit_s - reqs[i].arrival_s);
    }
    stats->wall_s = wall;
    stats->grid_point_steps = gps;
    stats->throughput = wall > 0 ? gps / wall : 0;
    stats->step_p50_s = pct(lat_all, 0.5);
    stats->step_p99_s = pct(lat_all, 0.99);
    stats->ttc_p50_s = pct_plain(ttc, 0.5);
    stats->ttc_p99_s = pct_plain(ttc, 0.99);
    stats->queue_p50_s = pct_plain(qw, 0.5);
    stats->queue_p99_s = pct_plain(qw, 0.99);
    stats->max_resident = max_resident;
    stats->peak_bytes = peak;
    stats->ticks = ticks;
    stats->completed = (int)ttc.size();
  }
  // release everything (on an error path some copies may still be queued)
  (void)hipStreamSynchronize(copy_stream);
  for (auto& s : snaps) {
    (void)hipFree(s.buf.dev);
    (void)hipHostFree(s.buf.host);
    (void)hipEventDestroy(s.done);
  }
  for (auto& kv : spare) {
    (void)hipFree(kv.second.dev);
    (void)hipHostFree(kv.second.host);
  }
  for (auto ev : spare_events) (void)hipEventDestroy(ev);
  for (auto& kv : work)
    if (kv.second.done) (void)hipEventDestroy(kv.second.done);
  (void)hipEventDestroy(t0ev);
  (void)hipStreamDestroy(copy_stream);
  af_engine_destroy(eng);
  return st;
}

}  // extern "C"
