// Serving loop: continuous batching of simulation requests on one GPU.
//
// Requests (grid size, parameters, horizon, snapshot cadence, IC seed) arrive
// at given times.  At every tick boundary the loop
//   1. admits arrived requests in arrival order while the engine's HBM budget
//      allows (a request that does not fit waits; later ones wait behind it),
//   2. steps every resident run by k steps in shared launches (k = the
//      largest number of steps before some run reaches a snapshot or its
//      horizon, capped by max_tick_steps),
//   3. starts the device->host copy of due snapshots on a copy stream into
//      pinned buffers (delivered to the callback once complete, while the next
//      tick computes),
//   4. retires finished runs, which frees their memory for waiting requests.
// Per-step latency is the interval between a request's consecutive step
// completions (a tick of k steps after an interval dt counts k steps of dt/k);
// time to completion is finish - arrival.  No reference counterpart: the
// reference runs one simulation per process.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
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
  int req;
  int slot;
  long long steps = 0;
  double last = 0;
};

struct PendingSnap {
  int req;
  long long step;
  int nx, ny;
  double M[2];
  double* dev = nullptr;    // staging (reference layout) on the device
  double* host = nullptr;   // pinned
  size_t n = 0;             // doubles
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
  if (hipStreamCreateWithFlags(&copy_stream, hipStreamNonBlocking) != hipSuccess) {
    af_engine_destroy(eng);
    return fail(AF_EDEVICE, "copy stream");
  }

  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return reqs[a].arrival_s < reqs[b].arrival_s; });
  std::vector<std::vector<std::pair<double, long long>>> lat(n);
  std::vector<std::pair<double, long long>> lat_all;
  std::vector<Resident> res;
  std::deque<PendingSnap> snaps;
  std::vector<double> host_state;
  int next = 0, max_resident = 0, ticks = 0;
  size_t peak = 0;
  double gps = 0;
  const auto t0 = std::chrono::steady_clock::now();
  auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };

  auto deliver = [&](bool wait_all) -> af_status {
    while (!snaps.empty()) {
      PendingSnap& s = snaps.front();
      hipError_t q = wait_all ? hipEventSynchronize(s.done) : hipEventQuery(s.done);
      if (q == hipErrorNotReady) break;
      if (q != hipSuccess) return hip_fail(q, "snapshot copy");
      if (fn) fn(user, s.req, s.step, s.host, s.M, s.nx, s.ny);
      results[s.req].snapshots += 1;
      (void)hipFree(s.dev);
      (void)hipHostFree(s.host);
      (void)hipEventDestroy(s.done);
      snaps.pop_front();
    }
    return AF_OK;
  };

  auto start_snapshot = [&](int req, const Resident& r) -> af_status {
    const Slot& sl = eng->slots[r.slot];
    PendingSnap s;
    s.req = req;
    s.step = r.steps;
    s.nx = sl.p.nx;
    s.ny = sl.p.ny;
    s.n = (size_t)sl.g.ny * (sl.g.nx2 + 1) * 2;
    MeanState ms;
    if (hipMalloc(&s.dev, s.n * sizeof(double)) != hipSuccess) return fail(AF_ENOMEM, "snapshot staging");
    if (hipHostMalloc(&s.host, s.n * sizeof(double), hipHostMallocDefault) != hipSuccess) {
      (void)hipFree(s.dev);
      return fail(AF_ENOMEM, "snapshot pinned buffer");
    }
    hipError_t e = sl.g.from_bricks(sl.base + sl.g.off_field[0], s.dev, sl.p.nx, sl.p.ny, sl.g.by, sl.g.bx,
                                    eng->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&ms, sl.base + sl.g.off_mean, sizeof(ms), hipMemcpyDeviceToHost, eng->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(eng->stream);   // 64 B: mean state now, field copy async
    if (e == hipSuccess) e = hipEventCreateWithFlags(&s.done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMemcpyAsync(s.host, s.dev, s.n * sizeof(double), hipMemcpyDeviceToHost, copy_stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, copy_stream);
    if (e != hipSuccess) {
      (void)hipFree(s.dev);
      (void)hipHostFree(s.host);
      return hip_fail(e, "snapshot");
    }
    s.M[0] = ms.M[0];
    s.M[1] = ms.M[1];
    snaps.push_back(s);
    return AF_OK;
  };

  st = AF_OK;
  while (st == AF_OK && (next < n || !res.empty())) {
    // 1. admission in arrival order
    double tnow = now();
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
      if (s2 == AF_ENOMEM && !res.empty()) break;   // wait for memory
      ++next;
      if (s2 != AF_OK) {
        results[q].status = s2;
        continue;
      }
      results[q].admit_s = now();
      Resident r;
      r.req = q;
      r.slot = slot;
      r.last = results[q].admit_s;
      res.push_back(r);
      if (rq.horizon <= 0) res.back().steps = 0;
    }
    max_resident = std::max(max_resident, (int)res.size());
    peak = std::max(peak, af_engine_bytes_in_use(eng));
    if (res.empty()) {
      if (next < n) {
        const double wait = reqs[order[next]].arrival_s - now();
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.005)));
      }
      if ((st = deliver(false)) != AF_OK) break;
      continue;
    }
    // 2. one tick for every resident run
    long long k = max_tick_steps;
    for (const Resident& r : res) {
      const af_request& rq = reqs[r.req];
      long long left = rq.horizon - r.steps;
      if (rq.snapshot_every > 0) left = std::min(left, rq.snapshot_every - r.steps % rq.snapshot_every);
      k = std::min(k, std::max(1LL, left));
    }
    std::vector<int> slots;
    for (const Resident& r : res) slots.push_back(r.slot);
    st = af_engine_step_slots(eng, slots.data(), (int)slots.size(), (int)k);
    // Every listed run has taken its k steps even when the watchdog reports
    // NaN.  A run whose state went NaN fails alone, as the reference stops
    // that run (.cu:850-861): status AF_ENAN, no finish time; the others go on.
    std::vector<char> bad(res.size(), 0);
    if (st == AF_ENAN) {
      st = AF_OK;
      for (size_t i = 0; i < res.size() && st == AF_OK; ++i) {
        const Slot& sl = eng->slots[res[i].slot];
        MeanState ms;
        hipError_t e = hipMemcpy(&ms, sl.base + sl.g.off_mean, sizeof(ms), hipMemcpyDeviceToHost);
        if (e != hipSuccess) st = hip_fail(e, "mean download");
        else bad[i] = ms.nan_step >= 0;
      }
    }
    if (st != AF_OK) break;
    ++ticks;
    tnow = now();
    // 3. latencies, snapshots, completions
    std::vector<Resident> keep;
    for (size_t i = 0; i < res.size(); ++i) {
      Resident& r = res[i];
      const af_request& rq = reqs[r.req];
      r.steps += k;
      const double per = (tnow - r.last) / (double)k;
      lat[r.req].push_back({per, k});
      lat_all.push_back({per, k});
      r.last = tnow;
      gps += (double)rq.params.nx * rq.params.ny * (double)k;
      if (bad[i]) {
        results[r.req].status = AF_ENAN;
        (void)deliver(true);   // its earlier snapshots precede its removal
        if ((st = af_engine_remove(eng, r.slot)) != AF_OK) break;
        continue;
      }
      if (rq.snapshot_every > 0 && r.steps % rq.snapshot_every == 0)
        if ((st = start_snapshot(r.req, r)) != AF_OK) break;
      if (r.steps >= rq.horizon) {
        const Slot& sl = eng->slots[r.slot];
        std::vector<double> tmp;
        double* out = (final_states && final_states[r.req]) ? final_states[r.req] : nullptr;
        if (!out) {
          tmp.resize((size_t)sl.g.ny * (sl.g.nx2 + 1) * 2);
          out = tmp.data();
        }
        if ((st = af_engine_get_state(eng, r.slot, out, results[r.req].M)) != AF_OK) break;
        results[r.req].finish_s = now();
        results[r.req].status = AF_OK;
        (void)deliver(true);   // snapshots of this run precede its removal
        if ((st = af_engine_remove(eng, r.slot)) != AF_OK) break;
      } else {
        keep.push_back(r);
      }
    }
    res.swap(keep);
    if (st == AF_OK) st = deliver(false);
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
  while (!snaps.empty()) {   // error path: release what is still queued
    (void)hipEventSynchronize(snaps.front().done);
    (void)hipFree(snaps.front().dev);
    (void)hipHostFree(snaps.front().host);
    (void)hipEventDestroy(snaps.front().done);
    snaps.pop_front();
  }
  (void)hipStreamDestroy(copy_stream);
  af_engine_destroy(eng);
  return st;
}

}  // extern "C"
