// Ensemble engine: many resident runs of mixed size and parameters on one
// device.  Runs of the same (precision, Nx, Ny, schedule) form a size class;
// one step of a class is the same 8 launches as a single run, each launch
// covering every run of the class (blockIdx -> (entry, tile) through the
// class's array of per-run descriptors).  Classes are independent: each has
// its own stream and its own step graphs, so different sizes execute
// concurrently, no block ever waits on another, and each run's arithmetic is
// exactly that of the run stepped alone (same kernels, same per-run data, no
// cross-run reductions).
//
// Membership changes edit the descriptor array in place on the class's
// stream (an empty entry is all zeros and its blocks exit at once), so a
// class's graphs are recaptured only when its array grows or is compacted.
// Blocks of retired runs are kept for reuse by runs of the same size.
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

namespace af {

static void drop_graphs(Class& c) {
  for (int l = 0; l < kEngGraphLevels; ++l) {
    if (c.gexec[l]) (void)hipGraphExecDestroy(c.gexec[l]);
    if (c.graph[l]) (void)hipGraphDestroy(c.graph[l]);
    c.gexec[l] = nullptr;
    c.graph[l] = nullptr;
  }
  c.graph_launched = -1;
}

Class* engine_class(af_engine* e, int slot) {
  auto it = e->classes.find(e->slots[slot].key);
  return it == e->classes.end() ? nullptr : &it->second;
}

// A device block of `bytes`: a retired run's block of the same size, else a
// new allocation (pooled blocks are released first if the budget requires).
static af_status take_block(af_engine* e, size_t bytes, unsigned char** out) {
  engine_reclaim(e);
  auto it = e->free_blocks.find(bytes);
  if (it != e->free_blocks.end()) {
    *out = it->second;
    e->free_blocks.erase(it);
    e->pooled -= bytes;
    return AF_OK;
  }
  while (!e->free_blocks.empty() && e->in_use + e->pooled + bytes > e->budget) {
    auto big = std::prev(e->free_blocks.end());
    (void)hipFree(big->second);
    e->pooled -= big->first;
    e->free_blocks.erase(big);
  }
  hipError_t he = hipMalloc(out, bytes);
  if (he != hipSuccess) return fail(AF_ENOMEM, std::string("device allocation: ") + hipGetErrorString(he));
  return AF_OK;
}

void engine_reclaim(af_engine* e) {
  std::vector<PendingBlock> keep;
  for (PendingBlock& b : e->pending) {
    if (hipEventQuery(b.done) == hipSuccess) {
      (void)hipEventDestroy(b.done);
      e->free_blocks.emplace(b.bytes, b.base);
      e->pooled += b.bytes;
      e->in_use -= b.bytes;
    } else {
      keep.push_back(b);
    }
  }
  e->pending.swap(keep);
}

// Creates the class of slot s's key (stream, descriptor array, twiddle pool).
static af_status make_class(af_engine* e, const Slot& s, Class& c) {
  c.g = s.g;
  c.desc_bytes = s.g.prec == 64 ? sizeof(InstDesc<double>) : sizeof(InstDesc<float>);
  hipError_t he = hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  if (he != hipSuccess) { c.stream = nullptr; return hip_fail(he, "class stream"); }
  const size_t pool_bytes = (size_t)2 * s.g.lmax * s.g.esz;
  c.cap = 16;
  if ((he = hipMalloc(&c.descs, (size_t)c.cap * c.desc_bytes)) != hipSuccess ||
      (he = hipMalloc(&c.pool, pool_bytes)) != hipSuccess)
    return fail(AF_ENOMEM, std::string("class arrays: ") + hipGetErrorString(he));
  if ((he = hipMemcpyAsync(c.pool, s.base + s.g.off_pool, pool_bytes, hipMemcpyDeviceToDevice, e->stream)) !=
          hipSuccess ||
      (he = hipStreamSynchronize(e->stream)) != hipSuccess)
    return hip_fail(he, "class twiddle pool");
  return AF_OK;
}

// Puts slot id into an empty entry of its class, or appends an entry.
static af_status class_insert(af_engine* e, Class& c, int id) {
  Slot& s = e->slots[id];
  int entry = -1;
  for (int i = 0; i < (int)c.entry_slot.size(); ++i)
    if (c.entry_slot[i] < 0) { entry = i; break; }
  hipError_t he;
  if (entry < 0) {
    entry = (int)c.entry_slot.size();
    if (entry == c.cap) {   // grow: copy the array on the class's stream, keep the old one until destroy
      unsigned char* bigger = nullptr;
      if ((he = hipMalloc(&bigger, (size_t)2 * c.cap * c.desc_bytes)) != hipSuccess)
        return fail(AF_ENOMEM, "descriptor array");
      if ((he = hipMemcpyAsync(bigger, c.descs, (size_t)c.cap * c.desc_bytes, hipMemcpyDeviceToDevice, c.stream)) !=
          hipSuccess) {
        (void)hipFree(bigger);
        return hip_fail(he, "descriptor array copy");
      }
      e->graveyard.push_back(c.descs);
      c.descs = bigger;
      c.cap *= 2;
      drop_graphs(c);   // they point at the old array
    }
    c.entry_slot.push_back(-1);
  }
  if ((he = hipMemcpyAsync(c.descs + (size_t)entry * c.desc_bytes, s.base + s.g.off_desc, c.desc_bytes,
                           hipMemcpyDeviceToDevice, c.stream)) != hipSuccess)
    return hip_fail(he, "descriptor copy");
  c.entry_slot[entry] = id;
  s.entry = entry;
  c.members += 1;
  return AF_OK;
}

af_status engine_retire(af_engine* e, int id) {
  Slot& s = e->slots[id];
  Class* c = engine_class(e, id);
  if (!c) return fail(AF_EINVAL, "slot has no class");
  hipError_t he = hipMemsetAsync(c->descs + (size_t)s.entry * c->desc_bytes, 0, c->desc_bytes, c->stream);
  PendingBlock b{nullptr, s.base, s.g.bytes};
  if (he == hipSuccess) he = hipEventCreateWithFlags(&b.done, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventRecord(b.done, c->stream);
  if (he != hipSuccess) return hip_fail(he, "retire");
  e->pending.push_back(b);
  c->entry_slot[s.entry] = -1;
  c->members -= 1;
  s = Slot();
  return AF_OK;
}

// Drops trailing empty entries; once a quarter of the rest are empty, moves
// the last runs into the holes (descriptor copies on the class's stream).
static af_status compact(af_engine* e, Class& c) {
  auto trim = [&]() {
    while (!c.entry_slot.empty() && c.entry_slot.back() < 0) c.entry_slot.pop_back();
  };
  trim();
  const int n = (int)c.entry_slot.size();
  if ((n - c.members) * 4 > n) {
    for (int i = 0; i < (int)c.entry_slot.size(); ++i) {
      if (c.entry_slot[i] >= 0) continue;
      const int last = (int)c.entry_slot.size() - 1;
      if (last <= i) break;
      const int id = c.entry_slot[last];
      hipError_t he = hipMemcpyAsync(c.descs + (size_t)i * c.desc_bytes, c.descs + (size_t)last * c.desc_bytes,
                                     c.desc_bytes, hipMemcpyDeviceToDevice, c.stream);
      if (he == hipSuccess) he = hipMemsetAsync(c.descs + (size_t)last * c.desc_bytes, 0, c.desc_bytes, c.stream);
      if (he != hipSuccess) return hip_fail(he, "compaction");
      c.entry_slot[i] = id;
      c.entry_slot[last] = -1;
      e->slots[id].entry = i;
      trim();
    }
  }
  c.launched = (int)c.entry_slot.size();
  return AF_OK;
}

static bool capture_level(Class& c, int level) {
  hipError_t e = hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t le = hipSuccess;
    for (int i = 0; i < (1 << level) && le == hipSuccess; ++i) le = launch_step(c.g, c.descs, c.launched, c.pool, c.stream);
    e = hipStreamEndCapture(c.stream, &c.graph[level]);
    if (le == hipSuccess && e == hipSuccess) e = hipGraphInstantiate(&c.gexec[level], c.graph[level], nullptr, nullptr, 0);
    else if (e == hipSuccess) e = le;
  }
  if (e != hipSuccess || !c.gexec[level]) {
    (void)hipGetLastError();
    if (c.graph[level]) (void)hipGraphDestroy(c.graph[level]);
    c.graph[level] = nullptr;
    c.gexec[level] = nullptr;
    return false;
  }
  return true;
}

// Compaction and a mean-state mirror for every entry, before a launch.
static af_status class_prepare(af_engine* e, Class& c) {
  af_status st = compact(e, c);
  if (st != AF_OK) return st;
  if (c.mirror_cap < c.launched) {
    hipError_t he = hipStreamSynchronize(c.stream);
    if (he == hipSuccess && c.mirror) he = hipHostFree(c.mirror);
    c.mirror = nullptr;
    c.mirror_cap = 0;
    if (he == hipSuccess) he = hipHostMalloc((void**)&c.mirror, (size_t)c.cap * sizeof(MeanState), hipHostMallocCoherent);
    if (he != hipSuccess) { c.mirror = nullptr; return fail(AF_ENOMEM, "mean-state mirror"); }
    c.mirror_cap = c.cap;
  }
  return AF_OK;
}

af_status engine_enqueue(af_engine* e, Class& c, int k) {
  af_status st = class_prepare(e, c);
  if (st != AF_OK) return st;
  if (c.launched == 0 || k <= 0) return AF_OK;
  if (c.graph_launched != c.launched) {
    // the previous graphs may still be running on the class's stream
    hipError_t he = hipStreamSynchronize(c.stream);
    if (he != hipSuccess) return hip_fail(he, "class sync");
    drop_graphs(c);
    c.graph_launched = c.launched;
  }
  auto enqueue_level = [&](int level) -> hipError_t {
    if (c.gexec[level] || capture_level(c, level)) return hipGraphLaunch(c.gexec[level], c.stream);
    hipError_t he = hipSuccess;   // capture failed: direct launches
    for (int i = 0; i < (1 << level) && he == hipSuccess; ++i)
      he = launch_step(c.g, c.descs, c.launched, c.pool, c.stream);
    return he;
  };
  const int top = kEngGraphLevels - 1;
  hipError_t he = hipSuccess;
  int left = k;
  while (left >= (1 << top) && he == hipSuccess) { he = enqueue_level(top); left -= 1 << top; }
  for (int l = top - 1; l >= 0 && he == hipSuccess; --l)
    if (left & (1 << l)) he = enqueue_level(l);
  if (he == hipSuccess) he = gather_means(c.g.prec, c.descs, c.launched, c.mirror, c.stream);
  if (he != hipSuccess) return hip_fail(he, "engine step");
  return AF_OK;
}

// ---- single-launch mode ----------------------------------------------------

static void drop_mix_graphs(af_engine* e) {
  for (int l = 0; l < kEngGraphLevels; ++l) {
    if (e->mix_exec[l]) (void)hipGraphExecDestroy(e->mix_exec[l]);
    if (e->mix_graph[l]) (void)hipGraphDestroy(e->mix_graph[l]);
    e->mix_exec[l] = nullptr;
    e->mix_graph[l] = nullptr;
  }
  e->mix = nullptr;
  e->mix_key.clear();
}

static bool same_schedule(const Schedule& a, const Schedule& b) {
  return a.nx2 == b.nx2 && a.ny == b.ny && a.by == b.by && a.bx == b.bx && a.row_r == b.row_r &&
         a.row_g == b.row_g && a.col_r == b.col_r && a.col_gf == b.col_gf && a.col_gi == b.col_gi &&
         a.col_w == b.col_w && a.pp == b.pp;
}

// The instantiated mix made of exactly the classes `keys` (its classes in
// launch order into `order`), or null.
static const MultiCfg* find_mix(af_engine* e, const std::vector<ClassKey>& keys, std::vector<Class*>& order) {
  int n = 0;
  const MultiCfg* m = multi_cfgs(&n);
  for (int i = 0; i < n; ++i) {
    const MultiCfg& cfg = m[i];
    if (cfg.nclass != (int)keys.size()) continue;
    std::vector<Class*> ord;
    for (int k = 0; k < cfg.nclass; ++k) {
      Class* hit = nullptr;
      for (const ClassKey& key : keys) {
        const Schedule* s = std::get<3>(key);
        if (std::get<0>(key) == cfg.precision && s && same_schedule(*s, cfg.cls[k])) {
          hit = &e->classes.at(key);
          break;
        }
      }
      if (!hit) break;
      ord.push_back(hit);
    }
    if ((int)ord.size() == cfg.nclass) {
      order = ord;
      return &cfg;
    }
  }
  return nullptr;
}

// k steps of every run of the mix's classes, each pass one mixed launch on
// the engine's stream, then the classes' mean-state gathers.  Asynchronous.
static af_status mix_enqueue(af_engine* e, const MultiCfg* mix, const std::vector<Class*>& cls, int k) {
  for (Class* c : cls) {
    af_status st = class_prepare(e, *c);
    if (st != AF_OK) return st;
  }
  hipError_t he = hipSuccess;
  for (Class* c : cls)   // descriptor edits on the class streams land before the mixed launches
    if (he == hipSuccess) he = hipStreamSynchronize(c->stream);
  if (he != hipSuccess) return hip_fail(he, "class sync");
  MultiArgs ra{}, ca{};
  std::vector<const void*> key;
  for (size_t i = 0; i < cls.size(); ++i) {
    const Class& c = *cls[i];
    ra.descs[i] = ca.descs[i] = c.descs;
    ra.pool[i] = ca.pool[i] = c.pool;
    ra.tiles[i] = c.g.row_tiles;
    ca.tiles[i] = c.g.col_tiles;
    ra.nblocks[i] = c.launched * c.g.row_tiles;
    ca.nblocks[i] = c.launched * c.g.col_tiles;
    key.push_back(c.descs);
    key.push_back(reinterpret_cast<const void*>((size_t)c.launched));
  }
  if (e->mix != mix || e->mix_key != key) {
    drop_mix_graphs(e);
    e->mix = mix;
    e->mix_key = key;
  }
  auto steps = [&](int n) -> hipError_t {
    hipError_t r = hipSuccess;
    for (int i = 0; i < n && r == hipSuccess; ++i)
      for (int stage = 1; stage <= 4 && r == hipSuccess; ++stage) {
        r = mix->row(ra, e->stream);
        if (r == hipSuccess) r = mix->col(ca, stage, e->stream);
      }
    return r;
  };
  auto level = [&](int l) -> hipError_t {
    if (!e->mix_exec[l]) {
      hipError_t r = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
      if (r == hipSuccess) {
        const hipError_t lr = steps(1 << l);
        r = hipStreamEndCapture(e->stream, &e->mix_graph[l]);
        if (lr == hipSuccess && r == hipSuccess) r = hipGraphInstantiate(&e->mix_exec[l], e->mix_graph[l], nullptr, nullptr, 0);
        else if (r == hipSuccess) r = lr;
      }
      if (r != hipSuccess || !e->mix_exec[l]) {   // capture failed: direct launches
        (void)hipGetLastError();
        if (e->mix_graph[l]) (void)hipGraphDestroy(e->mix_graph[l]);
        e->mix_graph[l] = nullptr;
        e->mix_exec[l] = nullptr;
        return steps(1 << l);
      }
    }
    return hipGraphLaunch(e->mix_exec[l], e->stream);
  };
  const int top = kEngGraphLevels - 1;
  int left = k;
  while (left >= (1 << top) && he == hipSuccess) { he = level(top); left -= 1 << top; }
  for (int l = top - 1; l >= 0 && he == hipSuccess; --l)
    if (left & (1 << l)) he = level(l);
  for (Class* c : cls)
    if (he == hipSuccess) he = gather_means(c->g.prec, c->descs, c->launched, c->mirror, e->stream);
  if (he != hipSuccess) return hip_fail(he, "mixed step");
  return AF_OK;
}

}  // namespace af

using namespace af;

static bool valid_slot(const af_engine* e, int slot) {
  return e && slot >= 0 && slot < (int)e->slots.size() && e->slots[slot].used;
}

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
  engine_reclaim(e);
  if (e->in_use + g.bytes > e->budget) return fail(AF_ENOMEM, "HBM budget exceeded");
  st = nyquist_check(g, omega_hat);
  if (st != AF_OK) return st;
  unsigned char* base = nullptr;
  st = take_block(e, g.bytes, &base);
  if (st != AF_OK) return st;
  st = init_device_block(p, g, base, e->stream);
  if (st == AF_OK) st = load_state(g, base, omega_hat, M, e->stream);
  if (st != AF_OK) {
    e->free_blocks.emplace(g.bytes, base);
    e->pooled += g.bytes;
    return st;
  }
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
  s.key = ClassKey(p->precision, p->nx, p->ny, g.sched);
  e->in_use += g.bytes;
  auto it = e->classes.find(s.key);
  if (it == e->classes.end()) {
    Class c;
    st = make_class(e, s, c);
    if (st == AF_OK) {
      it = e->classes.emplace(s.key, c).first;
    } else {
      if (c.descs) (void)hipFree(c.descs);
      if (c.pool) (void)hipFree(c.pool);
      if (c.stream) (void)hipStreamDestroy(c.stream);
    }
  }
  if (st == AF_OK) st = class_insert(e, it->second, id);
  if (st != AF_OK) {   // give the block back; the slot stays free
    e->free_blocks.emplace(g.bytes, base);
    e->pooled += g.bytes;
    e->in_use -= g.bytes;
    s = Slot();
    return st;
  }
  *slot = id;
  return AF_OK;
}

af_status af_engine_remove(af_engine* e, int slot) {
  if (!valid_slot(e, slot)) return fail(AF_EINVAL, "bad slot");
  Class* c = engine_class(e, slot);
  af_status st = engine_retire(e, slot);
  if (st != AF_OK) return st;
  hipError_t he = hipStreamSynchronize(c->stream);
  if (he != hipSuccess) return hip_fail(he, "remove sync");
  engine_reclaim(e);
  return AF_OK;
}

af_status af_engine_step_slots(af_engine* e, const int* slots, int nslots, int nsteps) {
  if (!e || (nslots > 0 && !slots) || nsteps < 0) return fail(AF_EINVAL, "bad arguments");
  std::vector<int> list(slots, slots + nslots);
  std::sort(list.begin(), list.end());
  list.erase(std::unique(list.begin(), list.end()), list.end());
  for (int id : list)
    if (!valid_slot(e, id)) return fail(AF_EINVAL, "bad slot in list");
  if (list.empty() || nsteps == 0) return AF_OK;
  // classes involved; members not listed sit out (their entries are zeroed
  // for the call and restored afterwards)
  std::map<ClassKey, std::vector<int>> by;
  for (int id : list) by[e->slots[id].key].push_back(id);
  std::vector<std::pair<Class*, std::vector<int>>> idle;   // (class, slots sitting out)
  hipError_t he = hipSuccess;
  for (auto& kv : by) {
    Class& c = e->classes.at(kv.first);
    if ((int)kv.second.size() == c.members) continue;
    std::vector<int> out;
    for (int id : c.entry_slot)
      if (id >= 0 && !std::binary_search(kv.second.begin(), kv.second.end(), id)) out.push_back(id);
    for (int id : out)
      if (he == hipSuccess)
        he = hipMemsetAsync(c.descs + (size_t)e->slots[id].entry * c.desc_bytes, 0, c.desc_bytes, c.stream);
    idle.emplace_back(&c, out);
  }
  af_status st = he == hipSuccess ? AF_OK : hip_fail(he, "sit-out descriptors");
  std::vector<Class*> mix_order;
  const MultiCfg* mix = nullptr;
  if (e->single_launch) {
    std::vector<ClassKey> keys;
    for (auto& kv : by) keys.push_back(kv.first);
    mix = find_mix(e, keys, mix_order);
  }
  int done = 0;
  while (st == AF_OK && done < nsteps) {
    const int k = std::min(nsteps - done, kRingCap);
    if (mix) {
      st = mix_enqueue(e, mix, mix_order, k);
      if (st == AF_OK && (he = hipStreamSynchronize(e->stream)) != hipSuccess) st = hip_fail(he, "engine sync");
    } else {
      for (auto& kv : by)
        if (st == AF_OK) st = engine_enqueue(e, e->classes.at(kv.first), k);
      for (auto& kv : by)
        if (st == AF_OK && (he = hipStreamSynchronize(e->classes.at(kv.first).stream)) != hipSuccess)
          st = hip_fail(he, "engine sync");
    }
    if (st == AF_OK)
      for (int id : list) e->slots[id].steps += k;
    done += k;
  }
  for (auto& ci : idle)   // restore the entries of the runs that sat out
    for (int id : ci.second) {
      const Slot& s = e->slots[id];
      he = hipMemcpyAsync(ci.first->descs + (size_t)s.entry * ci.first->desc_bytes, s.base + s.g.off_desc,
                          ci.first->desc_bytes, hipMemcpyDeviceToDevice, ci.first->stream);
      if (he != hipSuccess && st == AF_OK) st = hip_fail(he, "descriptor restore");
    }
  for (auto& ci : idle)
    if ((he = hipStreamSynchronize(ci.first->stream)) != hipSuccess && st == AF_OK) st = hip_fail(he, "restore sync");
  if (st != AF_OK) return st;
  // NaN watchdog per run, from the mirrors the last chunk's gathers wrote
  for (int id : list) {
    const Class* c = engine_class(e, id);
    const MeanState& ms = c->mirror[e->slots[id].entry];
    if (ms.nan_step >= 0)
      return fail(AF_ENAN, "slot " + std::to_string(id) + ": array contains nan values after " +
                               std::to_string(ms.nan_step) + " timesteps");
  }
  return AF_OK;
}

af_status af_engine_set_single_launch(af_engine* e, int on) {
  if (!e) return fail(AF_EINVAL, "engine is NULL");
  e->single_launch = on != 0;
  return AF_OK;
}

int af_engine_single_launch_active(af_engine* e) {
  if (!e || !e->single_launch) return 0;
  std::vector<ClassKey> keys;
  for (const auto& kv : e->classes)
    if (kv.second.members > 0) keys.push_back(kv.first);
  std::vector<Class*> order;
  return find_mix(e, keys, order) ? 1 : 0;
}

af_status af_engine_step(af_engine* e, int nsteps) {
  if (!e) return fail(AF_EINVAL, "engine is NULL");
  std::vector<int> all;
  for (size_t i = 0; i < e->slots.size(); ++i)
    if (e->slots[i].used) all.push_back((int)i);
  return af_engine_step_slots(e, all.data(), (int)all.size(), nsteps);
}

af_status af_engine_get_state(af_engine* e, int slot, double* omega_hat, double M[2]) {
  if (!valid_slot(e, slot) || !omega_hat || !M) return fail(AF_EINVAL, "bad arguments");
  const Slot& s = e->slots[slot];
  hipError_t he = hipStreamSynchronize(engine_class(e, slot)->stream);
  if (he != hipSuccess) return hip_fail(he, "class sync");
  return store_state(s.g, s.base, omega_hat, M, e->stream);
}

af_status af_engine_get_energy(af_engine* e, int slot, int n, double* energy) {
  if (!valid_slot(e, slot) || n < 0 || (n > 0 && !energy)) return fail(AF_EINVAL, "bad arguments");
  const Slot& s = e->slots[slot];
  if (n > s.steps || n > kRingCap) return fail(AF_EINVAL, "fewer energies available than requested");
  hipError_t he = hipStreamSynchronize(engine_class(e, slot)->stream);
  if (he != hipSuccess) return hip_fail(he, "class sync");
  std::vector<double> ring(kRingCap);
  he = hipMemcpy(ring.data(), s.base + s.g.off_energy, kRingCap * sizeof(double), hipMemcpyDeviceToHost);
  if (he != hipSuccess) return hip_fail(he, "energy download");
  for (int i = 0; i < n; ++i) energy[i] = ring[(size_t)((s.steps - n + i) % kRingCap)];
  return AF_OK;
}

af_status af_engine_diagnostics(af_engine* e, int slot, double* energy, double* enstrophy, double* spectrum,
                                int nshell) {
  if (!valid_slot(e, slot) || nshell < 0 || (nshell > 0 && !spectrum)) return fail(AF_EINVAL, "bad arguments");
  const Slot& s = e->slots[slot];
  hipError_t he = hipStreamSynchronize(engine_class(e, slot)->stream);
  if (he != hipSuccess) return hip_fail(he, "class sync");
  return diagnostics(s.p, s.g, s.base, e->stream, energy, enstrophy, spectrum, nshell);
}

long long af_engine_steps_done(const af_engine* e, int slot) {
  if (!valid_slot(e, slot)) return -1;
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
  drop_mix_graphs(e);
  for (auto& kv : e->classes) {
    Class& c = kv.second;
    drop_graphs(c);
    if (c.descs) (void)hipFree(c.descs);
    if (c.pool) (void)hipFree(c.pool);
    if (c.mirror) (void)hipHostFree(c.mirror);
    if (c.stream) (void)hipStreamDestroy(c.stream);
  }
  for (auto& s : e->slots)
    if (s.used) (void)hipFree(s.base);
  for (auto& b : e->pending) {
    (void)hipEventDestroy(b.done);
    (void)hipFree(b.base);
  }
  for (auto& kv : e->free_blocks) (void)hipFree(kv.second);
  for (void* p : e->graveyard) (void)hipFree(p);
  (void)hipStreamDestroy(e->stream);
  delete e;
}

}  // extern "C"
