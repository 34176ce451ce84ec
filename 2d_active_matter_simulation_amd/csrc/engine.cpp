// Ensemble engine: many resident runs of mixed size and parameters on one
// device.  Runs of the same (precision, Nx, Ny) form a size class; one step of
// a class is the same 8 launches as a single run, each launch covering every
// run of the class (blockIdx -> (run, tile) through an array of per-run
// descriptors).  Classes are independent, so each gets its own stream and
// the steps of all classes are captured as one hipGraph that forks into a
// branch per class and joins at the end (one graph of one step, one of
// kGraphSteps steps in which each class runs its steps back to back):
// different sizes execute concurrently, no block ever waits on another, and
// each run's arithmetic is exactly that of the run stepped alone (same
// kernels, same per-run data, no cross-run reductions).
//
// Admission is by an HBM budget on the bytes of run state the engine holds.
// No reference counterpart: the reference runs one grid per process.
#include <algorithm>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "engine_impl.hpp"

using namespace af;

// Steps per multi-step graph (replaying one graph of 16 steps instead of 16
// one-step graphs removes the gaps between graph launches and the per-step
// joins of the classes).
constexpr int kGraphSteps = 16;

static void drop_graph(af_engine* e) {
  for (int i = 0; i < 2; ++i) {
    if (e->gexec[i]) (void)hipGraphExecDestroy(e->gexec[i]);
    if (e->graph[i]) (void)hipGraphDestroy(e->graph[i]);
    e->gexec[i] = nullptr;
    e->graph[i] = nullptr;
  }
  e->graph_slots.clear();
}

static ClassKey key_of(const Slot& s) { return ClassKey(s.p.precision, s.p.nx, s.p.ny, s.g.sched); }

extern "C" {

af_status af_engine_create(size_t hbm_budget, af_engine** out) {
  if (!out) return fail(AF_EINVAL, "out is NULL");
  af_engine* e = new af_engine();
  e->budget = hbm_budget;
  hipError_t he = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
  if (he != hipSuccess) { delete e; return hip_fail(he, "engine stream"); }
  *out = e;
  return AF_OK;
}

af_status af_engine_add(af_engine* e, const af_params* p, const double* omega_hat, const double M[2], int* slot) {
  if (!e || !p || !omega_hat || !M || !slot) return fail(AF_EINVAL, "NULL argument");
  af_status st = validate(p);
  if (st != AF_OK) return st;
  Geometry g = geometry(p, true);   // the engine's tuned table (batched launches)
  if (!g.rc || !g.cc) return fail(AF_EUNSUPPORTED, "no kernel for this size");
  if (e->in_use + g.bytes > e->budget) return fail(AF_ENOMEM, "HBM budget exceeded");
  st = nyquist_check(g, omega_hat);
  if (st != AF_OK) return st;
  // the class stream first: nothing is claimed if it cannot be created
  Class& c = e->classes[ClassKey(p->precision, p->nx, p->ny, g.sched)];
  hipError_t he;
  if (!c.stream) {
    he = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
    if (he != hipSuccess) { c.stream = nullptr; return hip_fail(he, "class stream"); }
  }
  unsigned char* base = nullptr;
  he = hipMalloc(&base, g.bytes);
  if (he != hipSuccess) return fail(AF_ENOMEM, std::string("device allocation: ") + hipGetErrorString(he));
  st = init_device_block(p, g, base, e->stream);
  if (st == AF_OK) st = load_state(g, base, omega_hat, M, e->stream);
  if (st != AF_OK) { (void)hipFree(base); return st; }
  int id = -1;
  for (size_t i = 0; i < e->slots.size(); ++i)
    if (!e->slots[i].used) { id = (int)i; break; }
  if (id < 0) { id = (int)e->slots.size(); e->slots.emplace_back(); }
  Slot& s = e->slots[id];
  s.used = true;
  s.p = *p;
  s.g = g;
  s.base = base;
  s.steps = 0;
  e->in_use += g.bytes;
  *slot = id;
  return AF_OK;
}

af_status af_engine_remove(af_engine* e, int slot) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used) return fail(AF_EINVAL, "bad slot");
  (void)hipStreamSynchronize(e->stream);
  for (auto& kv : e->classes) (void)hipStreamSynchronize(kv.second.stream);
  drop_graph(e);
  Slot& s = e->slots[slot];
  (void)hipFree(s.base);
  e->in_use -= s.g.bytes;
  s = Slot();
  return AF_OK;
}

static af_status build_descs(af_engine* e, const std::vector<int>& list,
                             std::vector<std::pair<ClassKey, std::vector<int>>>& groups) {
  std::map<ClassKey, std::vector<int>> by;
  for (int id : list) by[key_of(e->slots[id])].push_back(id);
  groups.assign(by.begin(), by.end());
  for (auto& gr : groups) {
    Class& c = e->classes[gr.first];
    const Geometry& g = e->slots[gr.second[0]].g;
    const size_t db = g.prec == 64 ? sizeof(InstDesc<double>) : sizeof(InstDesc<float>);
    if ((int)gr.second.size() > c.cap) {
      if (c.descs) (void)hipFree(c.descs);
      c.cap = std::max<int>((int)gr.second.size(), 2 * c.cap);
      hipError_t he = hipMalloc(&c.descs, (size_t)c.cap * db);
      if (he != hipSuccess) { c.cap = 0; c.descs = nullptr; return fail(AF_ENOMEM, "descriptor array"); }
    }
    c.desc_bytes = db;
    for (size_t i = 0; i < gr.second.size(); ++i) {
      const Slot& s = e->slots[gr.second[i]];
      hipError_t he = hipMemcpyAsync((unsigned char*)c.descs + i * db, s.base + s.g.off_desc, db,
                                     hipMemcpyDeviceToDevice, e->stream);
      if (he != hipSuccess) return hip_fail(he, "descriptor copy");
    }
  }
  hipError_t he = hipStreamSynchronize(e->stream);
  if (he != hipSuccess) return hip_fail(he, "descriptor sync");
  return AF_OK;
}

// `steps` steps of every class in `groups`: fork from the main stream, each
// class's steps back to back on its own stream, join.
static hipError_t enqueue_steps(af_engine* e, const std::vector<std::pair<ClassKey, std::vector<int>>>& groups,
                                int steps) {
  // fork: every class stream waits on the main stream, joins back at the end
  if (e->events.size() < groups.size() + 1) {
    for (size_t i = e->events.size(); i < groups.size() + 1; ++i) {
      hipEvent_t ev;
      hipError_t he = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
      if (he != hipSuccess) return he;
      e->events.push_back(ev);
    }
  }
  hipError_t he = hipEventRecord(e->events[0], e->stream);
  if (he != hipSuccess) return he;
  for (size_t i = 0; i < groups.size(); ++i) {
    Class& c = e->classes[groups[i].first];
    const Slot& s0 = e->slots[groups[i].second[0]];
    if ((he = hipStreamWaitEvent(c.stream, e->events[0], 0)) != hipSuccess) return he;
    for (int k = 0; k < steps; ++k)
      if ((he = launch_step(s0.g, c.descs, (int)groups[i].second.size(), s0.base + s0.g.off_pool, c.stream)) !=
          hipSuccess)
        return he;
    if ((he = hipEventRecord(e->events[i + 1], c.stream)) != hipSuccess) return he;
    if ((he = hipStreamWaitEvent(e->stream, e->events[i + 1], 0)) != hipSuccess) return he;
  }
  return hipSuccess;
}

af_status af_engine_step_slots(af_engine* e, const int* slots, int nslots, int nsteps) {
  if (!e || (nslots > 0 && !slots) || nsteps < 0) return fail(AF_EINVAL, "bad arguments");
  std::vector<int> list(slots, slots + nslots);
  std::sort(list.begin(), list.end());
  list.erase(std::unique(list.begin(), list.end()), list.end());
  for (int id : list)
    if (id < 0 || id >= (int)e->slots.size() || !e->slots[id].used) return fail(AF_EINVAL, "bad slot in list");
  if (list.empty() || nsteps == 0) return AF_OK;
  std::vector<std::pair<ClassKey, std::vector<int>>> groups;
  if (list != e->graph_slots || !e->gexec[0]) {
    drop_graph(e);
    af_status st = build_descs(e, list, groups);
    if (st != AF_OK) return st;
    for (int gi = 0; gi < 2; ++gi) {
      hipError_t he = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
      if (he == hipSuccess) {
        hipError_t le = enqueue_steps(e, groups, gi == 0 ? 1 : kGraphSteps);
        he = hipStreamEndCapture(e->stream, &e->graph[gi]);
        if (le != hipSuccess) he = le;
        if (he == hipSuccess) he = hipGraphInstantiate(&e->gexec[gi], e->graph[gi], nullptr, nullptr, 0);
      }
      if (he != hipSuccess) {
        (void)hipGetLastError();
        drop_graph(e);
        return hip_fail(he, "engine graph");
      }
    }
    e->graph_slots = list;
  }
  int done = 0;
  while (done < nsteps) {
    const int k = std::min(nsteps - done, kRingCap);
    for (int i = 0; i < k;) {
      const int n = (k - i >= kGraphSteps) ? kGraphSteps : 1;
      hipError_t he = hipGraphLaunch(e->gexec[n == 1 ? 0 : 1], e->stream);
      if (he != hipSuccess) return hip_fail(he, "engine step");
      i += n;
    }
    hipError_t he = hipStreamSynchronize(e->stream);
    if (he != hipSuccess) return hip_fail(he, "engine sync");
    for (int id : list) e->slots[id].steps += k;
    done += k;
  }
  // NaN watchdog per run
  for (int id : list) {
    MeanState ms;
    const Slot& s = e->slots[id];
    hipError_t he = hipMemcpy(&ms, s.base + s.g.off_mean, sizeof(ms), hipMemcpyDeviceToHost);
    if (he != hipSuccess) return hip_fail(he, "mean download");
    if (ms.nan_step >= 0)
      return fail(AF_ENAN, "slot " + std::to_string(id) + ": array contains nan values after " +
                               std::to_string(ms.nan_step) + " timesteps");
  }
  return AF_OK;
}

af_status af_engine_step(af_engine* e, int nsteps) {
  if (!e) return fail(AF_EINVAL, "engine is NULL");
  std::vector<int> all;
  for (size_t i = 0; i < e->slots.size(); ++i)
    if (e->slots[i].used) all.push_back((int)i);
  return af_engine_step_slots(e, all.data(), (int)all.size(), nsteps);
}

af_status af_engine_get_state(af_engine* e, int slot, double* omega_hat, double M[2]) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used || !omega_hat || !M)
    return fail(AF_EINVAL, "bad arguments");
  const Slot& s = e->slots[slot];
  return store_state(s.g, s.base, omega_hat, M, e->stream);
}

af_status af_engine_get_energy(af_engine* e, int slot, int n, double* energy) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used || n < 0 || (n > 0 && !energy))
    return fail(AF_EINVAL, "bad arguments");
  const Slot& s = e->slots[slot];
  if (n > s.steps || n > kRingCap) return fail(AF_EINVAL, "fewer energies available than requested");
  std::vector<double> ring(kRingCap);
  hipError_t he = hipMemcpy(ring.data(), s.base + s.g.off_energy, kRingCap * sizeof(double), hipMemcpyDeviceToHost);
  if (he != hipSuccess) return hip_fail(he, "energy download");
  for (int i = 0; i < n; ++i) energy[i] = ring[(size_t)((s.steps - n + i) % kRingCap)];
  return AF_OK;
}

af_status af_engine_diagnostics(af_engine* e, int slot, double* energy, double* enstrophy, double* spectrum,
                                int nshell) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used || nshell < 0 ||
      (nshell > 0 && !spectrum))
    return fail(AF_EINVAL, "bad arguments");
  const Slot& s = e->slots[slot];
  return diagnostics(s.p, s.g, s.base, e->stream, energy, enstrophy, spectrum, nshell);
}

long long af_engine_steps_done(const af_engine* e, int slot) {
  if (!e || slot < 0 || slot >= (int)e->slots.size() || !e->slots[slot].used) return -1;
  return e->slots[slot].steps;
}

size_t af_engine_bytes_in_use(const af_engine* e) { return e ? e->in_use : 0; }

int af_engine_count(const af_engine* e) {
  if (!e) return 0;
  int n = 0;
  for (const auto& s : e->slots) n += s.used ? 1 : 0;
  return n;
}

void af_engine_destroy(af_engine* e) {
  if (!e) return;
  (void)hipStreamSynchronize(e->stream);
  for (auto& kv : e->classes) (void)hipStreamSynchronize(kv.second.stream);
  drop_graph(e);
  for (auto& s : e->slots)
    if (s.used) (void)hipFree(s.base);
  for (auto& kv : e->classes) {
    if (kv.second.descs) (void)hipFree(kv.second.descs);
    if (kv.second.stream) (void)hipStreamDestroy(kv.second.stream);
  }
  for (auto ev : e->events) (void)hipEventDestroy(ev);
  (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"
