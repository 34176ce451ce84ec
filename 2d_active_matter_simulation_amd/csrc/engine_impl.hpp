// Internals of the ensemble engine shared with the serving loop.
#pragma once
#include <map>
#include <tuple>
#include <vector>

#include "run_impl.hpp"

namespace af {

constexpr int kEngGraphLevels = 5;   // step graphs of 1, 2, 4, 8 and 16 steps

using ClassKey = std::tuple<int, int, int, const Schedule*>;  // precision, nx, ny, schedule

struct Slot {
  bool used = false;
  af_params p;
  Geometry g;
  unsigned char* base = nullptr;
  long long steps = 0;
  ClassKey key;
  int entry = -1;   // position in its class's descriptor array
};

// Runs of one (precision, Nx, Ny, schedule).  Their descriptors sit in one
// device array, edited in place on the class's stream: a run that leaves
// zeroes its entry (the kernels skip such an entry), a run that joins fills an
// empty entry or is appended.  The step graphs cover the first `launched`
// entries and are recaptured only when that number changes (an append, or a
// compaction once a quarter of the entries are empty).
struct Class {
  Geometry g;                          // launch configuration of the class
  hipStream_t stream = nullptr;
  unsigned char* descs = nullptr;      // device descriptor array, cap entries
  int cap = 0;
  size_t desc_bytes = 0;
  unsigned char* pool = nullptr;       // the class's own copy of the twiddle pool
  std::vector<int> entry_slot;         // slot of each entry, -1 = empty
  int launched = 0;                    // entries covered by a launch
  int graph_launched = -1;             // `launched` of the captured graphs
  hipGraph_t graph[kEngGraphLevels] = {};
  hipGraphExec_t gexec[kEngGraphLevels] = {};
  MeanState* mirror = nullptr;         // pinned, device-written: mean state per entry
  int mirror_cap = 0;
  int members = 0;
};

// A retired run's block, reusable once the class stream has passed its
// removal (event).
struct PendingBlock {
  hipEvent_t done;
  unsigned char* base;
  size_t bytes;
};

}  // namespace af

struct af_engine {
  size_t budget = 0;
  size_t in_use = 0;                    // bytes of runs held (pending blocks included)
  std::vector<af::Slot> slots;
  std::map<af::ClassKey, af::Class> classes;
  hipStream_t stream = nullptr;         // set-up work: state loads and copies
  std::multimap<size_t, unsigned char*> free_blocks;   // reusable device blocks, by size
  size_t pooled = 0;                    // bytes in free_blocks
  std::vector<af::PendingBlock> pending;
  std::vector<void*> graveyard;         // replaced descriptor arrays, freed at destroy
  // single-launch mode: when the classes stepped together form an
  // instantiated mix (af::multi_cfgs), each pass is one mixed launch over all
  // of them on `stream` instead of one launch per class on the class streams
  bool single_launch = false;
  const af::MultiCfg* mix = nullptr;     // the mix of the captured graphs
  std::vector<const void*> mix_key;      // descriptor arrays and entry counts they cover
  hipGraph_t mix_graph[af::kEngGraphLevels] = {};
  hipGraphExec_t mix_exec[af::kEngGraphLevels] = {};
};

namespace af {

// Enqueues k >= 1 steps of every run of class c on its stream (compacting
// and recapturing first if needed), then a gather of the runs' mean states
// into c.mirror.  Asynchronous; the class's stream must not be capturing.
af_status engine_enqueue(af_engine* e, Class& c, int k);
// Removes a run: its entry is zeroed on its class's stream and its block is
// reusable once the stream has passed that point.  Asynchronous.
af_status engine_retire(af_engine* e, int slot);
// Moves the blocks of retired runs whose removal has completed to the pool.
void engine_reclaim(af_engine* e);
Class* engine_class(af_engine* e, int slot);
// The gather kernel behind Class::mirror (diag.hip).
hipError_t gather_means(int prec, const void* descs, int n, MeanState* out, hipStream_t s);

}  // namespace af
