// Schedule lookup over the per-precision launch tables, with the tuned tables
// next to libactflow.so (written by autotune.py) read once: schedules.txt for
// single runs, schedules_batched.txt for runs stepped by the ensemble engine.
#include "registry.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

namespace af {

thread_local LaunchEvents g_launch_events;


const RowCfg* row_cfgs_f32(int* n);
const ColCfg* col_cfgs_f32(int* n);
const Schedule* schedules_f32(int* n);
ConvLaunch to_bricks_f32();
IfacLaunch ifac_init_f32();
ConvLaunch from_bricks_f32();
const RowCfg* row_cfgs_f64(int* n);
const ColCfg* col_cfgs_f64(int* n);
const Schedule* schedules_f64(int* n);
ConvLaunch to_bricks_f64();
ConvLaunch from_bricks_f64();

namespace {

using Key = std::tuple<int, int, int>;   // precision, nx2, ny
std::mutex g_mu;
std::map<Key, const Schedule*> g_selected;
std::map<Key, const Schedule*> g_tuned;           // single runs (schedules.txt)
std::map<Key, const Schedule*> g_tuned_batched;   // ensemble engine (schedules_batched.txt)
bool g_tuned_loaded = false;

const Schedule* table(int precision, int* n) {
  return precision == 64 ? schedules_f64(n) : schedules_f32(n);
}

std::vector<const Schedule*> list_for(int precision, int nx2, int ny) {
  int n = 0;
  const Schedule* s = table(precision, &n);
  std::vector<const Schedule*> out;
  for (int i = 0; i < n; ++i)
    if (s[i].nx2 == nx2 && s[i].ny == ny) {
      if (s[i].is_default) out.insert(out.begin(), &s[i]); else out.push_back(&s[i]);
    }
  return out;
}

void load_table(const std::string& path, std::map<Key, const Schedule*>& into) {
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof line, f)) {
    if (line[0] == '#') continue;
    // prec nx ny by bx row_r row_g col_r col_gf col_gi [col_w [options]]
    // (col_w defaults to bx, options to 0)
    int prec, nx, ny, v[9];
    const int got = std::sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d %d", &prec, &nx, &ny, &v[0], &v[1], &v[2],
                                &v[3], &v[4], &v[5], &v[6], &v[7], &v[8]);
    if (got < 10) continue;
    if (got == 10 || v[7] <= 0) v[7] = v[1];
    if (got < 12) v[8] = 0;
    for (const Schedule* s : list_for(prec, nx / 2, ny))
      if (s->by == v[0] && s->bx == v[1] && s->row_r == v[2] && s->row_g == v[3] && s->col_r == v[4] &&
          s->col_gf == v[5] && s->col_gi == v[6] && s->col_w == v[7] && s->pp == v[8])
        into[Key(prec, nx / 2, ny)] = s;
  }
  std::fclose(f);
}

void load_tuned_locked() {
  if (g_tuned_loaded) return;
  g_tuned_loaded = true;
  Dl_info info;
  if (!dladdr((void*)&load_tuned_locked, &info) || !info.dli_fname) return;
  std::string dir(info.dli_fname);
  const size_t slash = dir.rfind('/');
  dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
  load_table(dir + "/schedules.txt", g_tuned);
  load_table(dir + "/schedules_batched.txt", g_tuned_batched);
}

}  // namespace

int schedule_count(int precision, int nx2, int ny) { return (int)list_for(precision, nx2, ny).size(); }

const Schedule* schedule_at(int precision, int nx2, int ny, int index) {
  auto l = list_for(precision, nx2, ny);
  return (index >= 0 && index < (int)l.size()) ? l[index] : nullptr;
}

const Schedule* active_schedule(int precision, int nx2, int ny, bool batched) {
  std::lock_guard<std::mutex> lk(g_mu);
  const Key k(precision, nx2, ny);
  auto it = g_selected.find(k);
  if (it != g_selected.end()) return it->second;
  load_tuned_locked();
  if (batched) {
    auto bt = g_tuned_batched.find(k);
    if (bt != g_tuned_batched.end()) return bt->second;
  }
  auto jt = g_tuned.find(k);
  if (jt != g_tuned.end()) return jt->second;
  auto l = list_for(precision, nx2, ny);
  return l.empty() ? nullptr : l[0];
}

bool select_schedule(int precision, int nx2, int ny, int index) {
  if (index < 0) {   // back to the tuned table / default
    std::lock_guard<std::mutex> lk(g_mu);
    g_selected.erase(Key(precision, nx2, ny));
    return true;
  }
  const Schedule* s = schedule_at(precision, nx2, ny, index);
  if (!s) return false;
  std::lock_guard<std::mutex> lk(g_mu);
  g_selected[Key(precision, nx2, ny)] = s;
  return true;
}

const RowCfg* find_row_cfg(int precision, const Schedule* s) {
  int n = 0;
  const RowCfg* c = precision == 64 ? row_cfgs_f64(&n) : row_cfgs_f32(&n);
  const int rows = (s->pp & kOptHalfRows) ? s->by / 2 : s->by;   // rows per row block
  for (int i = 0; i < n; ++i)
    if (c[i].nx2 == s->nx2 && c[i].rb == rows && c[i].r == s->row_r && c[i].bx == s->bx && c[i].g == s->row_g &&
        c[i].pp == (s->pp & kRowOptions))
      return &c[i];
  return nullptr;
}

const ColCfg* find_col_cfg(int precision, const Schedule* s) {
  int n = 0;
  const ColCfg* c = precision == 64 ? col_cfgs_f64(&n) : col_cfgs_f32(&n);
  for (int i = 0; i < n; ++i)
    if (c[i].ny == s->ny && c[i].w == s->col_w && c[i].bx == s->bx && c[i].r == s->col_r && c[i].by == s->by &&
        c[i].gf == s->col_gf && c[i].gi == s->col_gi && c[i].pp == (s->pp & kColOptions))
      return &c[i];
  return nullptr;
}

ConvLaunch to_bricks_launch(int precision) { return precision == 64 ? to_bricks_f64() : to_bricks_f32(); }
IfacLaunch ifac_init_launch() { return ifac_init_f32(); }
ConvLaunch from_bricks_launch(int precision) { return precision == 64 ? from_bricks_f64() : from_bricks_f32(); }

}  // namespace af
