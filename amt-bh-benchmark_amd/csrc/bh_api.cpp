// Host side of the C ABI (include/bh_gpu.h): the context, its grow-only
// device buffers, and the per-step orchestration of the kernels.
//
// One step (step_particles, integrate_ttg.h:88-107, grouped branch):
//   system table upload -> bounds -> keys + digit histograms -> 8 onesweep
//   radix passes -> gather (Morton-ordered particles + FP64 walk copy) ->
//   tree counts -> node emit -> links/centres -> upward pass (per level) ->
//   targets and walk units -> walk with fused kick/drift.
// Everything is ordered on the context's stream and nothing in between
// needs the host: the kernels read the node, level and unit counts on the
// device. The host waits once, at the end, and reads back one block of flag
// and count words. The node arrays and the number of upward levels are
// planned from what the context has seen (and from n); a tree that turns out
// larger or deeper sets an abort word, every kernel after the node count
// returns at once, and the step runs again with the arrays grown (its input
// is never written, so running it again is exact). Because a step makes no
// host decision midway, the whole launch sequence can be captured once into
// a HIP graph and replayed (bh_set_graph_mode).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "bh_internal.h"

using namespace bhk;

static_assert(kGroupCap == BH_MAX_WALK_TILE && kMinTile == BH_MIN_WALK_TILE, "walk tile limits");

namespace {

// The allocation-free entry point's working memory: one fixed array per GPU
// (every call on that GPU shares it), reached through the address that
// hipGetSymbolAddress returns for the current device.
constexpr size_t kNoAllocBytes = size_t(BH_NOALLOC_SCRATCH_MIB) << 20;
constexpr uint32_t kNoAllocSlots = 512;  // walk workgroups of the allocation-free step
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
hipError_t scratch_address(int device, void** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch_ptr[device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = ptr;
  }
  *out = g_scratch_ptr[device];
  return hipSuccess;
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
};

// Every array a step uses besides the particles in and out.
struct Scratch {
  uint64_t *keys_a = nullptr, *keys_b = nullptr;
  uint32_t *vals_a = nullptr, *vals_b = nullptr;
  uint32_t *hist = nullptr, *dbase = nullptr, *lookback = nullptr;
  double *lohi = nullptr, *partial = nullptr;
  uint32_t* misc = nullptr;              // kMiscWords
  unsigned long long* totals = nullptr;  // [2] interaction totals
  double4* posm = nullptr;
  uint8_t *L = nullptr, *H = nullptr;
  uint32_t *tcounts = nullptr, *tprefix = nullptr;
  uint16_t* sid = nullptr;
  uint8_t* tables = nullptr;
  TreeDev t{};
  uint32_t *mark = nullptr, *targets = nullptr, *gmark = nullptr, *groups = nullptr, *unit_n = nullptr;
  uint32_t *defer = nullptr, *deep_frames = nullptr;
  double* gacc = nullptr;
};

// What the host decides before a step is enqueued.
struct Plan {
  uint32_t n = 0;         // particles of all systems
  uint32_t S = 0;         // non-empty systems
  uint32_t n_chunks = 0;  // bounds blocks
  uint32_t mlx = 0;       // largest max_leaf
  int sid_passes = 0;     // radix passes over the system id
  bool unsoftened = false;
  size_t table_bytes = 0, off_chunks = 0, off_first = 0, off_hist = 0;
  uint64_t node_estimate = 0;
  uint32_t cap_nodes = 0;  // node array capacity the kernels may use
  uint32_t depth_plan = kDepthCap;
  bool deep = true;        // walk stack frames beyond LDS
  uint32_t slots = 0;      // walk workgroups
  bool walk = true;        // false: tree build only
  bool carry_out = false;  // the walk writes the compact positions of its output (resident steps)
  bool carried_in = false; // the input is the last resident step's output: build from its compact positions
};

// One captured step, keyed by everything its launches depend on.
struct GraphKey {
  const void* in;
  const void* out;
  uint64_t gen;
  uint32_t n, S, n_chunks, mlx, cap_nodes, depth_plan, slots;
  uint32_t table_bytes;
  uint8_t deep, unsoftened, precision, mode, carry_out, carried_in;
  bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof *this) == 0; }
};
struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec;
  uint64_t used;
};
constexpr size_t kGraphCache = 256;

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // events: 0 step start, 1 sorted, 2 tree built, 3 targets chosen, 4 walk start, 5 walk (step) end
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool have_stats = false;
  bool last_graph = false;  // the last step ran as a graph: only its total time is known
  uint32_t reruns = 0;      // times the last step ran again with larger arrays
  std::string err;
  int precision = BH_FP32;
  int walk_mode = BH_WALK_GROUPED;
  int graph_mode = BH_GRAPH_OFF;
  uint32_t walk_tile = kGroupCap;
  uint32_t stack_limit = kLevels;  // lowered only by bh_debug_walk_stack_limit
  void* arena = nullptr;  // this device's allocation-free scratch

  DevBuf<bh_particle> pa, pb;    // particle ping-pong of the resident state
  int cur = 0;                   // which of pa/pb holds the resident state
  uint64_t state_n = 0;
  // {x, y, z, m} of the last resident step's output and its non-finite flag,
  // valid (carried_ok) until anything else runs or the state is exposed
  DevBuf<double> carry;
  DevBuf<uint32_t> carry_flag;
  bool carried_ok = false;
  const void* carried_out = nullptr;
  uint64_t carried_n = 0;
  DevBuf<bh_particle> tmp_in, tmp_out;  // host-path staging (bh_step, bh_step_batch_host, bh_build_tree)
  DevBuf<uint64_t> keys_a, keys_b;
  DevBuf<uint32_t> vals_a, vals_b;
  DevBuf<uint32_t> hist, dbase, lookback;
  DevBuf<double> lohi, partial;
  DevBuf<uint32_t> misc;
  DevBuf<unsigned long long> totals;
  DevBuf<double4> posm;
  DevBuf<uint8_t> L, H;
  DevBuf<uint32_t> tcounts, tprefix;
  DevBuf<uint32_t> mark, targets, gmark, groups, unit_n, defer, deep_frames;
  DevBuf<double> gacc, nbox;
  DevBuf<uint32_t> begin, count, cs, first_child, parent, leaf_idx, leaf_node, leaf_first, level_off;
  DevBuf<uint8_t> depth, leaf, nchild;
  DevBuf<double4> ctr;
  DevBuf<NodeRec> rec;
  DevBuf<double> mass, com, iab, q64;
  DevBuf<uint8_t> tables;
  DevBuf<uint16_t> sid;
  // direct sum and energies
  DevBuf<double4> dsrc;
  DevBuf<uint32_t> didx, dcnt;
  DevBuf<double> dpart, dacc;

  uint32_t* host_misc = nullptr;  // pinned: kMiscWords flag/count words, then the two totals
  uint8_t* host_tables = nullptr; // pinned system table staging
  size_t host_tables_cap = 0;
  uint64_t gen = 1;               // bumped whenever a buffer a graph may hold moves

  // plans learnt from earlier steps
  uint32_t cap_nodes = 0;        // node capacity asked for (grow-only)
  bool depth_known = false;
  uint32_t depth_hint = kDepthCap;

  std::vector<GraphEntry> graphs;
  uint64_t graph_clock = 0;

  // last tree (for export and bh_accel) and last step's report
  bool has_tree = false;
  Plan plan{};
  Scratch sc{};
  uint32_t n = 0, n_nodes = 0, n_leaves = 0, max_depth = 0, n_targets = 0, n_groups = 0;
  uint32_t max_leaf = 0, group_size = 0;
  unsigned long long last_totals[2] = {0, 0};
  bh_particle* sorted = nullptr;    // device: Morton-ordered particles of the last tree
  uint64_t* keys_sorted = nullptr;  // device: sorted keys of the last tree
  uint32_t* perm_sorted = nullptr;  // device: sorted position -> input index
};

// One system of a batch as the host sees it.
struct HostSys {
  uint64_t n;
  bh_force_params fp;
  uint32_t max_leaf, group_size, target_size;
  double kick_dt, drift_dt;
  uint32_t tile;  // 0: the context's
};

bool set_err(Ctx* c, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  c->err = buf;
  return false;
}

#define HIP_OK(c, expr)                                                                         \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      set_err((c), "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return BH_ERR_HIP;                                                                        \
    }                                                                                           \
  } while (0)

#define TRY(expr)                 \
  do {                            \
    int rc_ = (expr);             \
    if (rc_ != BH_OK) return rc_; \
  } while (0)

void drop_graphs(Ctx* c) {
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  c->graphs.clear();
}

template <typename T>
int ensure(Ctx* c, DevBuf<T>& b, size_t n, const char* what) {
  if (n == 0) n = 1;
  if (b.cap >= n) return BH_OK;
  // a graph may still refer to the old buffer: forget every graph first
  drop_graphs(c);
  ++c->gen;
  if (b.p) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(b.p);
  }
  b.p = nullptr;
  b.cap = 0;
  const size_t want = n + n / 8;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(&b.p), want * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    set_err(c, "device allocation of %zu bytes for %s failed: %s", want * sizeof(T), what, hipGetErrorString(e));
    return BH_ERR_OOM;
  }
  b.cap = want;
  return BH_OK;
}

template <typename T>
void release(DevBuf<T>& b) {
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
}

int check_args(Ctx* c, uint64_t n, const bh_force_params* fp, uint32_t max_leaf, uint32_t group_size) {
  if (n > kMaxN) {
    set_err(c, "n = %llu exceeds the supported %llu particles", (unsigned long long)n, (unsigned long long)kMaxN);
    return BH_ERR_TOO_LARGE;
  }
  if (max_leaf > kMaxLeafCap) {
    set_err(c, "max_leaf = %u exceeds the supported %u", max_leaf, kMaxLeafCap);
    return BH_ERR_TOO_LARGE;
  }
  if (group_size == 0) {
    set_err(c, "group_size 0 selects the per-node pass, which this library does not provide");
    return BH_ERR_INVALID;
  }
  if (fp && (!(fp->theta >= 0.0) || !(fp->eps2 >= 0.0) || !std::isfinite(fp->G) || !std::isfinite(fp->theta) ||
             !std::isfinite(fp->eps2))) {
    set_err(c, "invalid force parameters (theta %g, eps2 %g, G %g)", fp->theta, fp->eps2, fp->G);
    return BH_ERR_INVALID;
  }
  return BH_OK;
}

// Pinned staging of the system table; with no_grow (the allocation-free
// step) it must already be large enough.
int host_table_room(Ctx* c, size_t bytes, bool no_grow) {
  if (c->host_tables_cap >= bytes) return BH_OK;
  if (no_grow) {
    set_err(c, "the system table needs %zu bytes of pinned staging, %zu were set up", bytes, c->host_tables_cap);
    return BH_ERR_TOO_LARGE;
  }
  drop_graphs(c);  // graphs copy the table from this buffer
  ++c->gen;
  if (c->host_tables) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipHostFree(c->host_tables);
  }
  c->host_tables = nullptr;
  c->host_tables_cap = 0;
  const size_t want = bytes + bytes / 2;
  if (hipHostMalloc(reinterpret_cast<void**>(&c->host_tables), want, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    set_err(c, "pinned host allocation of %zu bytes failed", want);
    return BH_ERR_OOM;
  }
  c->host_tables_cap = want;
  return BH_OK;
}

// The batch's system table: per-system parameters, chunks of <= kChunk
// particles for the per-system bounds, and the system-id digit counts of the
// sort's last passes, written to the pinned staging (the step uploads it
// with one copy). Empty systems take no part: the device table lists the
// others, in order, and a particle's system id indexes that list.
int plan_systems(Ctx* c, const HostSys* sys, uint32_t S, bool no_grow, Plan* p) {
  if (S == 0 || S > kMaxSystems) {
    set_err(c, "a batch holds 1..%u systems, %u given", kMaxSystems, S);
    return BH_ERR_INVALID;
  }
  uint64_t n = 0, chunks = 0;
  uint32_t mlx = 0, live = 0;
  bool unsoftened = false;
  double nodes = 0.0;
  for (uint32_t k = 0; k < S; ++k) {
    TRY(check_args(c, sys[k].n, &sys[k].fp, sys[k].max_leaf, sys[k].group_size));
    if (sys[k].tile && (sys[k].tile < kMinTile || sys[k].tile > kGroupCap)) {
      set_err(c, "walk tile %u outside %u..%u", sys[k].tile, kMinTile, kGroupCap);
      return BH_ERR_INVALID;
    }
    if (sys[k].n == 0) continue;
    ++live;
    n += sys[k].n;
    chunks += (sys[k].n + kChunk - 1) / kChunk;
    mlx = std::max(mlx, sys[k].max_leaf);
    unsoftened |= static_cast<float>(sys[k].fp.eps2) == 0.0f;  // zero in the FP32 walk too
    // nodes per particle of Plummer, Hernquist and uniform trees: at most
    // min(1.6, 4 / max_leaf), with a quarter to spare
    nodes += 1.25 * static_cast<double>(sys[k].n) * std::min(1.6, 4.0 / std::max(1u, sys[k].max_leaf)) + 64.0;
  }
  if (n > kMaxN) {
    set_err(c, "batch of %llu particles exceeds the supported %llu", (unsigned long long)n, (unsigned long long)kMaxN);
    return BH_ERR_TOO_LARGE;
  }
  *p = Plan{};
  p->n = static_cast<uint32_t>(n);
  if (n == 0) return BH_OK;
  p->S = live;
  p->n_chunks = static_cast<uint32_t>(chunks);
  p->mlx = mlx;
  p->sid_passes = live == 1 ? 0 : (live <= 256 ? 1 : 2);
  p->unsoftened = unsoftened;
  p->node_estimate = static_cast<uint64_t>(nodes);
  p->off_chunks = sizeof(SysDev) * live;
  p->off_first = p->off_chunks + sizeof(uint4) * chunks;
  p->off_hist = (p->off_first + 4 * (live + 1) + 15) / 16 * 16;
  p->table_bytes = p->off_hist + 2 * 256 * 4;
  TRY(host_table_room(c, p->table_bytes, no_grow));
  std::memset(c->host_tables, 0, p->table_bytes);
  SysDev* sd = reinterpret_cast<SysDev*>(c->host_tables);
  uint4* ch = reinterpret_cast<uint4*>(c->host_tables + p->off_chunks);
  uint32_t* first = reinterpret_cast<uint32_t*>(c->host_tables + p->off_first);
  uint32_t* hist = reinterpret_cast<uint32_t*>(c->host_tables + p->off_hist);
  uint64_t off = 0, ci = 0;
  for (uint32_t k = 0, src = 0; k < live; ++k, ++src) {
    while (sys[src].n == 0) ++src;
    const HostSys& h = sys[src];
    sd[k].theta2 = h.fp.theta * h.fp.theta;
    sd[k].eps2 = h.fp.eps2;
    sd[k].G = h.fp.G;
    sd[k].kick_dt = h.kick_dt;
    sd[k].drift_dt = h.drift_dt;
    sd[k].off = static_cast<uint32_t>(off);
    sd[k].n = static_cast<uint32_t>(h.n);
    sd[k].max_leaf = h.max_leaf;
    sd[k].tsm = std::min(h.target_size, h.group_size);
    sd[k].use_quad = h.fp.use_quadrupole != 0;
    sd[k].tile = h.tile ? h.tile : c->walk_tile;
    first[k] = static_cast<uint32_t>(ci);
    for (uint64_t b = 0; b < h.n; b += kChunk)
      ch[ci++] = make_uint4(k, static_cast<uint32_t>(off + b), static_cast<uint32_t>(off + std::min<uint64_t>(h.n, b + kChunk)), 0);
    hist[k & 255] += static_cast<uint32_t>(h.n);
    hist[256 + (k >> 8)] += static_cast<uint32_t>(h.n);
    off += h.n;
  }
  first[live] = static_cast<uint32_t>(ci);
  // single-system queries (export, accel) read the first system's shape
  for (uint32_t k = 0; k < S; ++k) {
    if (sys[k].n == 0) continue;
    c->max_leaf = sys[k].max_leaf;
    c->group_size = sys[k].group_size;
    break;
  }
  return BH_OK;
}

// Walk workgroups: enough to fill the GPU (4 waves on every SIMD), fewer for
// small inputs; always a multiple of 8 (one share per XCD).
uint32_t walk_slots(uint32_t n, uint32_t limit) {
  uint32_t s = (n + 7) / 8;
  s = std::min(s, limit);
  s = std::max(s, 8u);
  return (s + 7) / 8 * 8;
}

// The plan's sizes that depend on what the context has learnt.
void plan_shape(Ctx* c, Plan* p, bool walk) {
  p->walk = walk;
  const uint64_t want = std::max<uint64_t>(p->node_estimate, c->cap_nodes);
  p->cap_nodes = static_cast<uint32_t>(std::min<uint64_t>(want, 1ull << 30));
  p->depth_plan = c->depth_known ? c->depth_hint : static_cast<uint32_t>(kDepthCap);
  p->deep = p->depth_plan + 1 > kShallowFrames;
  p->slots = walk_slots(p->n, kWalkSlots);
}

// Grows the context's buffers to the plan and points `b` at them.
int provision(Ctx* c, const Plan& p, Scratch* b) {
  const uint32_t n = p.n, m = p.cap_nodes;
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  TRY(ensure(c, c->keys_a, n, "keys"));
  TRY(ensure(c, c->keys_b, n, "keys"));
  TRY(ensure(c, c->vals_a, n, "indices"));
  TRY(ensure(c, c->vals_b, n, "indices"));
  TRY(ensure(c, c->hist, 10 * 256, "histograms"));
  TRY(ensure(c, c->dbase, 10 * 256, "digit offsets"));
  TRY(ensure(c, c->lookback, radix_lookback_words(n), "look-back"));
  TRY(ensure(c, c->lohi, 6ull * p.S, "bounds"));
  TRY(ensure(c, c->partial, 6ull * p.n_chunks, "bounds"));
  TRY(ensure(c, c->misc, kMiscWords, "flags"));
  TRY(ensure(c, c->totals, 2, "totals"));
  TRY(ensure(c, c->posm, n, "walk copy"));
  TRY(ensure(c, c->L, n, "tree build"));
  TRY(ensure(c, c->H, n, "tree build"));
  TRY(ensure(c, c->tcounts, uint64_t(tiles) * kTreeCols, "tree build"));
  TRY(ensure(c, c->tprefix, uint64_t(tiles) * kTreeCols, "tree build"));
  TRY(ensure(c, c->level_off, kLevels + 2, "tree levels"));
  TRY(ensure(c, c->sid, n, "system ids"));
  TRY(ensure(c, c->tables, p.table_bytes, "system table"));
  TRY(ensure(c, c->begin, m, "nodes"));
  TRY(ensure(c, c->count, m, "nodes"));
  TRY(ensure(c, c->cs, m, "nodes"));
  TRY(ensure(c, c->first_child, m, "nodes"));
  TRY(ensure(c, c->parent, m, "nodes"));
  TRY(ensure(c, c->leaf_idx, m, "nodes"));
  TRY(ensure(c, c->depth, m, "nodes"));
  TRY(ensure(c, c->leaf, m, "nodes"));
  TRY(ensure(c, c->nchild, m, "nodes"));
  TRY(ensure(c, c->ctr, m, "nodes"));
  TRY(ensure(c, c->rec, m, "nodes"));
  TRY(ensure(c, c->mass, m, "nodes"));
  TRY(ensure(c, c->com, 3ull * m, "nodes"));
  TRY(ensure(c, c->iab, 6ull * m, "nodes"));
  TRY(ensure(c, c->q64, 8ull * m, "nodes"));
  TRY(ensure(c, c->nbox, 6ull * m, "nodes"));
  TRY(ensure(c, c->leaf_node, m, "leaves"));
  TRY(ensure(c, c->leaf_first, m, "leaves"));
  if (p.carry_out || p.carried_in) {
    TRY(ensure(c, c->carry, 4ull * n, "carried positions"));
    TRY(ensure(c, c->carry_flag, 1, "carried positions"));
  }
  if (p.walk) {
    TRY(ensure(c, c->mark, n, "targets"));
    TRY(ensure(c, c->targets, n, "targets"));
    TRY(ensure(c, c->gmark, n, "groups"));
    TRY(ensure(c, c->groups, n, "groups"));
    TRY(ensure(c, c->unit_n, m, "walk units"));
    TRY(ensure(c, c->gacc, 3ull * n, "group sums"));
    TRY(ensure(c, c->defer, uint64_t(p.slots) * kDefer, "deferred nodes"));
    if (p.deep) TRY(ensure(c, c->deep_frames, uint64_t(p.slots) * (kLevels - kShallowFrames) * kWave, "walk stack"));
  }
  Scratch& s = *b;
  s = Scratch{};
  s.keys_a = c->keys_a.p;
  s.keys_b = c->keys_b.p;
  s.vals_a = c->vals_a.p;
  s.vals_b = c->vals_b.p;
  s.hist = c->hist.p;
  s.dbase = c->dbase.p;
  s.lookback = c->lookback.p;
  s.lohi = c->lohi.p;
  s.partial = c->partial.p;
  s.misc = c->misc.p;
  s.totals = c->totals.p;
  s.posm = c->posm.p;
  s.L = c->L.p;
  s.H = c->H.p;
  s.tcounts = c->tcounts.p;
  s.tprefix = c->tprefix.p;
  s.sid = c->sid.p;
  s.tables = c->tables.p;
  TreeDev& t = s.t;
  t.begin = c->begin.p;
  t.count = c->count.p;
  t.depth = c->depth.p;
  t.leaf = c->leaf.p;
  t.cs = c->cs.p;
  t.first_child = c->first_child.p;
  t.nchild = c->nchild.p;
  t.parent = c->parent.p;
  t.leaf_idx = c->leaf_idx.p;
  t.ctr = c->ctr.p;
  t.mass = c->mass.p;
  t.com = c->com.p;
  t.iab = c->iab.p;
  t.rec = c->rec.p;
  t.q64 = c->q64.p;
  t.nbox = c->nbox.p;
  t.leaf_node = c->leaf_node.p;
  t.leaf_first = c->leaf_first.p;
  t.level_off = c->level_off.p;
  t.misc = c->misc.p;
  t.cap = m;
  if (p.walk) {
    s.mark = c->mark.p;
    s.targets = c->targets.p;
    s.gmark = c->gmark.p;
    s.groups = c->groups.p;
    s.unit_n = c->unit_n.p;
    s.gacc = c->gacc.p;
    s.defer = c->defer.p;
    s.deep_frames = p.deep ? c->deep_frames.p : nullptr;
  }
  return BH_OK;
}

// Cuts the plan's arrays out of a fixed block of device memory, in the
// order and sizes of provision(), with the node capacity set to whatever is
// left. Fails (returns false) when the per-particle arrays do not fit.
bool carve(uint8_t* base, size_t bytes, Plan* p, Scratch* b) {
  size_t off = 0;
  bool ok = true;
  auto take = [&](size_t sz) -> uint8_t* {
    off = (off + 255) / 256 * 256;
    uint8_t* q = base + off;
    if (off + sz > bytes) ok = false;
    off += sz;
    return q;
  };
  const uint32_t n = p->n;
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  Scratch& s = *b;
  s = Scratch{};
  s.keys_a = reinterpret_cast<uint64_t*>(take(8ull * n));
  s.keys_b = reinterpret_cast<uint64_t*>(take(8ull * n));
  s.vals_a = reinterpret_cast<uint32_t*>(take(4ull * n));
  s.vals_b = reinterpret_cast<uint32_t*>(take(4ull * n));
  s.hist = reinterpret_cast<uint32_t*>(take(4 * 10 * 256));
  s.dbase = reinterpret_cast<uint32_t*>(take(4 * 10 * 256));
  s.lookback = reinterpret_cast<uint32_t*>(take(4 * radix_lookback_words(n)));
  s.lohi = reinterpret_cast<double*>(take(48ull * p->S));
  s.partial = reinterpret_cast<double*>(take(48ull * p->n_chunks));
  s.misc = reinterpret_cast<uint32_t*>(take(4 * kMiscWords));
  s.totals = reinterpret_cast<unsigned long long*>(take(16));
  s.posm = reinterpret_cast<double4*>(take(32ull * n));
  s.L = take(n);
  s.H = take(n);
  s.tcounts = reinterpret_cast<uint32_t*>(take(4ull * tiles * kTreeCols));
  s.tprefix = reinterpret_cast<uint32_t*>(take(4ull * tiles * kTreeCols));
  s.sid = reinterpret_cast<uint16_t*>(take(2ull * n));
  s.tables = take(p->table_bytes);
  s.mark = reinterpret_cast<uint32_t*>(take(4ull * n));
  s.targets = reinterpret_cast<uint32_t*>(take(4ull * n));
  s.gmark = reinterpret_cast<uint32_t*>(take(4ull * n));
  s.groups = reinterpret_cast<uint32_t*>(take(4ull * n));
  s.gacc = reinterpret_cast<double*>(take(24ull * n));
  s.defer = reinterpret_cast<uint32_t*>(take(4ull * p->slots * kDefer));
  s.deep_frames = reinterpret_cast<uint32_t*>(take(4ull * p->slots * (kLevels - kShallowFrames) * kWave));
  TreeDev& t = s.t;
  t.level_off = reinterpret_cast<uint32_t*>(take(4 * (kLevels + 2)));
  t.misc = s.misc;
  if (!ok) return false;
  // node arrays: 327 bytes per node, plus alignment of the 19 arrays
  const size_t per_node = 4 * 6 + 3 + 32 + 64 + 8 + 24 + 48 + 64 + 48 + 4 * 3;
  off = (off + 255) / 256 * 256;
  const size_t left = bytes > off + 19 * 256 ? bytes - off - 19 * 256 : 0;
  const uint64_t m = std::min<uint64_t>(left / per_node, 1ull << 30);
  if (m < 64) return false;
  t.begin = reinterpret_cast<uint32_t*>(take(4 * m));
  t.count = reinterpret_cast<uint32_t*>(take(4 * m));
  t.cs = reinterpret_cast<uint32_t*>(take(4 * m));
  t.first_child = reinterpret_cast<uint32_t*>(take(4 * m));
  t.parent = reinterpret_cast<uint32_t*>(take(4 * m));
  t.leaf_idx = reinterpret_cast<uint32_t*>(take(4 * m));
  t.depth = take(m);
  t.leaf = take(m);
  t.nchild = take(m);
  t.ctr = reinterpret_cast<double4*>(take(32 * m));
  t.rec = reinterpret_cast<NodeRec*>(take(64 * m));
  t.mass = reinterpret_cast<double*>(take(8 * m));
  t.com = reinterpret_cast<double*>(take(24 * m));
  t.iab = reinterpret_cast<double*>(take(48 * m));
  t.q64 = reinterpret_cast<double*>(take(64 * m));
  t.nbox = reinterpret_cast<double*>(take(48 * m));
  t.leaf_node = reinterpret_cast<uint32_t*>(take(4 * m));
  t.leaf_first = reinterpret_cast<uint32_t*>(take(4 * m));
  s.unit_n = reinterpret_cast<uint32_t*>(take(4 * m));
  t.cap = static_cast<uint32_t>(m);
  p->cap_nodes = t.cap;
  return ok;
}

struct Sorted {
  uint64_t* keys = nullptr;
  uint32_t* perm = nullptr;
};

// build_flat_tree (flat_build_ttg.h:276-281) of every system of the plan,
// d_in[0..n) into the scratch; the Morton-ordered particles (each system in
// its own range) go to d_out, here or (walk_writes_out) by the step's walk,
// which reads them from d_in. `events`: record the phase events.
int enqueue_build(Ctx* c, const Plan& p, const Scratch& b, const bh_particle* d_in, bh_particle* d_out, bool events,
                  Sorted* out, bool walk_writes_out) {
  hipStream_t s = c->stream;
  const uint32_t n = p.n;
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  const uint4* chunks = reinterpret_cast<const uint4*>(b.tables + p.off_chunks);
  const uint32_t* chunk_first = reinterpret_cast<const uint32_t*>(b.tables + p.off_first);
  const SysDev* sys = reinterpret_cast<const SysDev*>(b.tables);
  const uint64_t lb_words = radix_lookback_words(n);
  HIP_OK(c, hipMemcpyAsync(b.tables, c->host_tables, p.table_bytes, hipMemcpyHostToDevice, s));
  {
    // the step's flag words, totals, key digit histograms and look-back
    // words, in one launch
    ZeroRanges z;
    z.p[0] = b.misc;
    z.n[0] = kMiscWords;
    z.p[1] = reinterpret_cast<uint32_t*>(b.totals);
    z.n[1] = 4;
    z.p[2] = b.hist;
    z.n[2] = 8 * 256;
    z.p[3] = b.lookback;
    z.n[3] = lb_words;
    HIP_OK(c, launch_zero(z, s));
  }
  if (p.sid_passes)
    HIP_OK(c, hipMemcpyAsync(b.hist + 8 * 256, c->host_tables + p.off_hist, 2 * 256 * 4, hipMemcpyHostToDevice, s));
  HIP_OK(c, launch_fill_sid(chunks, p.n_chunks, b.sid, s));
  if (p.carried_in) {
    HIP_OK(c, launch_bounds(c->carry.p, n, c->carry_flag.p, chunks, p.n_chunks, chunk_first, p.S, b.partial,
                            b.misc + kMiscNonfinite, b.lohi, s));
    HIP_OK(c, launch_keys_hist(c->carry.p, n, b.sid, b.lohi, b.keys_a, b.hist, s));
  } else {
    HIP_OK(c, launch_bounds(d_in, chunks, p.n_chunks, chunk_first, p.S, b.partial, b.misc + kMiscNonfinite, b.lohi, s));
    HIP_OK(c, launch_keys_hist(d_in, n, b.sid, b.lohi, b.keys_a, b.hist, s));
  }
  HIP_OK(c, launch_radix_sort(b.keys_a, b.vals_a, b.keys_b, b.vals_b, n, p.sid_passes, b.sid, b.hist, b.dbase,
                              b.lookback, b.lookback + (lb_words - 10), &out->keys, &out->perm, s));
  if (p.carried_in && walk_writes_out)
    HIP_OK(c, launch_gather(c->carry.p, out->perm, b.posm, n, s));
  else
    HIP_OK(c, launch_gather(d_in, walk_writes_out ? nullptr : d_out, out->perm, b.posm, n, s));
  if (events) HIP_OK(c, hipEventRecord(c->ev[1], s));
  HIP_OK(c, launch_tree_count(out->keys, n, b.sid, sys, p.mlx, b.L, b.H, b.tcounts, s));
  HIP_OK(c, launch_tree_scan(b.tcounts, tiles, b.tprefix, b.misc + kMiscLevels, s));
  HIP_OK(c, launch_tree_emit(b.L, b.H, n, b.tprefix, b.t, p.depth_plan, s));
  HIP_OK(c, launch_tree_finish(out->keys, b.posm, n, b.sid, p.S, b.lohi, p.depth_plan, b.t, s));
  if (events) HIP_OK(c, hipEventRecord(c->ev[2], s));
  return BH_OK;
}

// run_grouped_pass (grouped_ttg.h:297-323) over every system of the tree in
// the scratch, with each system's parameters from the system table.
int enqueue_walk(Ctx* c, const Plan& p, const Scratch& b, bh_particle* parts, double* d_acc, uint32_t* d_npp,
                 uint32_t* d_npc, bool events, const bh_particle* src = nullptr, const uint32_t* perm = nullptr) {
  hipStream_t s = c->stream;
  const uint32_t n = p.n;
  const SysDev* sys = reinterpret_cast<const SysDev*>(b.tables);
  const bool grouped = c->walk_mode == BH_WALK_GROUPED;
  // a walk on its own (bh_accel) clears what an earlier walk left
  HIP_OK(c, launch_fill_u32(b.misc + kMiscIncomplete, 0u, 1, s));
  HIP_OK(c, launch_fill_u32(b.misc + kMiscQueue, 0u, 8, s));
  HIP_OK(c, launch_fill_u32(reinterpret_cast<uint32_t*>(b.totals), 0u, 4, s));
  HIP_OK(c, launch_targets(b.t, n, b.sid, sys, false, b.mark, b.tcounts, b.tprefix, b.targets, b.misc + kMiscTargets,
                           nullptr, s));
  // per-target mode: the walk units are the targets themselves
  HIP_OK(c, launch_targets(b.t, n, b.sid, sys, grouped, b.gmark, b.tcounts, b.tprefix, b.groups, b.misc + kMiscGroups,
                           grouped ? b.unit_n : nullptr, s));
  if (events) HIP_OK(c, hipEventRecord(c->ev[3], s));
  WalkArgs a;
  a.rec = b.t.rec;
  a.q64 = b.t.q64;
  a.ctr = b.t.ctr;
  a.begin = b.t.begin;
  a.count = b.t.count;
  a.posm = b.posm;
  a.groups = b.groups;
  a.unit_n = grouped ? b.unit_n : nullptr;
  a.tmark = b.mark;
  a.sid = b.sid;
  a.sys = sys;
  a.gacc = b.gacc;
  a.nbox = b.t.nbox;
  a.defer = b.defer;
  a.deep_frames = p.deep ? b.deep_frames : nullptr;
  a.misc = b.misc;
  a.n_nodes = p.cap_nodes;
  a.n_parts = n;
  a.stack_limit = c->stack_limit;
  a.parts = parts;
  a.src = src;
  a.perm = perm;
  a.carry = nullptr;
  a.carry_bad = nullptr;
  if (parts && p.carry_out) {
    // read by this step's build already (carried_in): cleared for the walk's
    HIP_OK(c, launch_fill_u32(c->carry_flag.p, 0u, 1, s));
    a.carry = c->carry.p;
    a.carry_bad = c->carry_flag.p;
  }
  a.acc = d_acc;
  a.n_pp = d_npp;
  a.n_pc = d_npc;
  a.totals = b.totals;
  if (events) HIP_OK(c, hipEventRecord(c->ev[4], s));
  HIP_OK(c, launch_walk(a, p.slots, p.deep, c->precision == BH_FP64, p.unsoftened, s));
  if (events) HIP_OK(c, hipEventRecord(c->ev[5], s));
  return BH_OK;
}

// Copies the flag/count words and the totals to the pinned report and
// waits for the stream: the step's only synchronisation.
int finish(Ctx* c, const Scratch& b) {
  HIP_OK(c, hipMemcpyAsync(c->host_misc, b.misc, kMiscWords * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipMemcpyAsync(c->host_misc + kMiscWords, b.totals, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

GraphKey graph_key(const Ctx* c, const Plan& p, const void* in, const void* out) {
  GraphKey k;
  std::memset(&k, 0, sizeof k);
  k.in = in;
  k.out = out;
  k.gen = c->gen;
  k.n = p.n;
  k.S = p.S;
  k.n_chunks = p.n_chunks;
  k.mlx = p.mlx;
  k.cap_nodes = p.cap_nodes;
  k.depth_plan = p.depth_plan;
  k.slots = p.slots;
  k.table_bytes = static_cast<uint32_t>(p.table_bytes);
  k.deep = p.deep;
  k.unsoftened = p.unsoftened;
  k.precision = static_cast<uint8_t>(c->precision);
  k.mode = static_cast<uint8_t>(c->walk_mode);
  k.carry_out = p.carry_out ? 1 : 0;
  k.carried_in = p.carried_in ? 1 : 0;
  return k;
}

// The step as one graph launch: replays the graph captured for the same
// launches, or captures and instantiates it first (the least recently used
// of kGraphCache graphs makes room).
int launch_graph(Ctx* c, const Plan& p, const Scratch& b, const bh_particle* d_in, bh_particle* d_out, Sorted* sorted) {
  const GraphKey key = graph_key(c, p, d_in, d_out);
  GraphEntry* hit = nullptr;
  for (auto& g : c->graphs)
    if (g.key == key) hit = &g;
  // the sort's output buffers follow from the pass count alone
  const int passes = 8 + p.sid_passes;
  sorted->keys = (passes % 2 == 0) ? b.keys_a : b.keys_b;
  sorted->perm = (passes % 2 == 0) ? b.vals_a : b.vals_b;
  if (!hit) {
    HIP_OK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    Sorted sd;
    int rc = enqueue_build(c, p, b, d_in, d_out, false, &sd, true);
    if (rc == BH_OK) rc = enqueue_walk(c, p, b, d_out, nullptr, nullptr, nullptr, false, d_in, sd.perm);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (rc != BH_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    HIP_OK(c, ec);
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    HIP_OK(c, ei);
    if (sd.keys != sorted->keys || sd.perm != sorted->perm) {
      (void)hipGraphExecDestroy(exec);
      set_err(c, "radix sort output buffers differ from the pass-count rule");
      return BH_ERR_HIP;
    }
    if (c->graphs.size() >= kGraphCache) {
      auto lru = std::min_element(c->graphs.begin(), c->graphs.end(),
                                  [](const GraphEntry& x, const GraphEntry& y) { return x.used < y.used; });
      (void)hipGraphExecDestroy(lru->exec);
      c->graphs.erase(lru);
    }
    c->graphs.push_back(GraphEntry{key, exec, 0});
    hit = &c->graphs.back();
  }
  hit->used = ++c->graph_clock;
  HIP_OK(c, hipEventRecord(c->ev[0], c->stream));
  HIP_OK(c, hipGraphLaunch(hit->exec, c->stream));
  HIP_OK(c, hipEventRecord(c->ev[5], c->stream));
  return BH_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
    (void)hipGetLastError();
    return 0.f;
  }
  return ms;
}

// Keeps the report of a finished step or build.
void take_report(Ctx* c, const Plan& p, const Scratch& b, const Sorted& sd, bh_particle* d_out, bool walked) {
  const uint32_t* m = c->host_misc;
  uint64_t nodes = 0;
  for (int d = 0; d < kLevels; ++d) nodes += m[kMiscLevels + d];
  c->n = p.n;
  c->n_nodes = static_cast<uint32_t>(nodes);
  c->n_leaves = m[kMiscLevels + kLevels];
  c->max_depth = m[kMiscMaxDepth];
  if (walked) {
    c->n_targets = m[kMiscTargets];
    c->n_groups = m[kMiscGroups];
    std::memcpy(c->last_totals, m + kMiscWords, sizeof c->last_totals);
  }
  c->plan = p;
  c->sc = b;
  c->sorted = d_out;
  c->keys_sorted = sd.keys;
  c->perm_sorted = sd.perm;
}

// A step (walk = true) or tree build of the systems planned in the pinned
// table, on the context's buffers: eagerly or as a graph, then the report.
// A tree that outgrew the plan is built again with room for it.
int run(Ctx* c, Plan p, const bh_particle* d_in, bh_particle* d_out, bool walk, bh_step_timing* timing) {
  c->have_stats = false;
  c->has_tree = false;
  c->carried_ok = false;
  for (int attempt = 0;; ++attempt) {
    c->reruns = static_cast<uint32_t>(attempt);
    plan_shape(c, &p, walk);
    Scratch b;
    TRY(provision(c, p, &b));
    Sorted sd;
    const bool graph = walk && c->graph_mode == BH_GRAPH_ON;
    if (graph) {
      TRY(launch_graph(c, p, b, d_in, d_out, &sd));
    } else {
      HIP_OK(c, hipEventRecord(c->ev[0], c->stream));
      TRY(enqueue_build(c, p, b, d_in, d_out, true, &sd, walk));
      if (walk) TRY(enqueue_walk(c, p, b, d_out, nullptr, nullptr, nullptr, true, d_in, sd.perm));
    }
    TRY(finish(c, b));
    const uint32_t* m = c->host_misc;
    if (m[kMiscNonfinite]) {
      set_err(c, "non-finite position, velocity or mass in the input");
      return BH_ERR_NONFINITE;
    }
    const uint32_t abort = m[kMiscAbort];
    if (abort) {
      if (attempt >= 3) {
        set_err(c, "the tree outgrew its planned arrays on every attempt (abort word %u)", abort);
        return BH_ERR_HIP;
      }
      if (abort & kAbortNodes) {
        uint64_t need = m[kMiscLevels + kLevels];
        uint64_t nodes = 0;
        for (int d = 0; d < kLevels; ++d) nodes += m[kMiscLevels + d];
        need = std::max(need, nodes);
        c->cap_nodes = static_cast<uint32_t>(std::min<uint64_t>(need + need / 8 + 64, 1ull << 30));
      }
      if (abort & kAbortDepth) {
        c->depth_hint = std::max(c->depth_known ? c->depth_hint : 0u, m[kMiscMaxDepth]);
        c->depth_known = true;
      }
      continue;
    }
    // later steps plan for the deepest tree seen
    c->depth_hint = c->depth_known ? std::max(c->depth_hint, m[kMiscMaxDepth]) : m[kMiscMaxDepth];
    c->depth_known = true;
    take_report(c, p, b, sd, d_out, walk);
    c->has_tree = true;
    c->have_stats = walk;
    c->last_graph = graph;
    if (timing) {
      if (graph) {
        timing->tree_build_ms = 0.0;
        timing->traversal_ms = elapsed(c->ev[0], c->ev[5]);  // the whole step: a graph has no inner events
      } else {
        timing->tree_build_ms = elapsed(c->ev[0], c->ev[2]);
        timing->traversal_ms = walk ? elapsed(c->ev[2], c->ev[5]) : 0.0;
      }
      timing->exchange_ms = 0.0;
    }
    if (walk && m[kMiscIncomplete]) {
      set_err(c, "incomplete tree: the walk met an interior node it could not open");
      return BH_ERR_INCOMPLETE;
    }
    if (walk && p.carry_out) {
      c->carried_ok = true;
      c->carried_out = d_out;
      c->carried_n = p.n;
    }
    return BH_OK;
  }
}

// One step of a batch: systems sys[0..S) own consecutive ranges of d_in.
// resident: a step of the resident state; its walk keeps the output's
// compact positions, which the next resident step builds from when its input
// is that output and nothing else ran in between.
int step_batch(Ctx* c, const bh_particle* d_in, bh_particle* d_out, const HostSys* sys, uint32_t S, bh_step_timing* timing,
               bool resident = false) {
  if (timing) *timing = bh_step_timing{0.0, 0.0, 0.0};
  HIP_OK(c, hipSetDevice(c->device));
  Plan p;
  TRY(plan_systems(c, sys, S, false, &p));
  if (p.n == 0) return BH_OK;
  p.carry_out = resident;
  p.carried_in = resident && c->carried_ok && c->carried_out == d_in && c->carried_n == p.n;
  return run(c, p, d_in, d_out, true, timing);
}

int step_device(Ctx* c, const bh_particle* d_in, bh_particle* d_out, uint64_t n, const bh_force_params* fp,
                uint32_t max_leaf, uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt,
                bh_step_timing* timing, bool resident = false) {
  TRY(check_args(c, n, fp, max_leaf, group_size));
  if (n == 0) {
    if (timing) *timing = bh_step_timing{0.0, 0.0, 0.0};
    return BH_OK;
  }
  const HostSys one{n, *fp, max_leaf, group_size, target_size, kick_dt, drift_dt, 0};
  return step_batch(c, d_in, d_out, &one, 1, timing, resident);
}

std::vector<HostSys> host_systems(const bh_system* systems, uint32_t S) {
  std::vector<HostSys> v(S);
  for (uint32_t k = 0; k < S; ++k)
    v[k] = HostSys{systems[k].n,         systems[k].fp,      systems[k].max_leaf,
                   systems[k].group_size, systems[k].target_size, systems[k].kick_dt,
                   systems[k].drift_dt, systems[k].walk_tile};
  return v;
}

bool overlap(const bh_particle* a, const bh_particle* b, uint64_t n) { return n && a + n > b && b + n > a; }

Ctx* as(bh_ctx* c) { return reinterpret_cast<Ctx*>(c); }

}  // namespace

int bhk::ctx_device(const bh_ctx* c) { return c ? reinterpret_cast<const Ctx*>(c)->device : 0; }

namespace {

// Reference node numbering (flat_build_ttg.h:194-273): BFS backbone down to
// subtree roots (count <= 4096 or leaf), then each subtree's other nodes BFS
// in subtree-root order; group roots in push order (flat_tree.h:190-225).
void reference_order(const std::vector<uint32_t>& count, const std::vector<uint8_t>& leaf,
                     const std::vector<uint32_t>& first_child, const std::vector<uint8_t>& nchild, uint32_t group_size,
                     std::vector<uint32_t>& ref_of, std::vector<uint32_t>& groups) {
  const size_t m = count.size();
  ref_of.assign(m, 0);
  groups.clear();
  auto group_root = [&](uint32_t u, uint32_t parent_count) {
    return (leaf[u] || count[u] <= group_size) && parent_count > group_size;
  };
  std::vector<uint32_t> bb{0};
  ref_of[0] = 0;
  if (group_root(0, std::numeric_limits<uint32_t>::max())) groups.push_back(0);
  std::vector<uint32_t> sroots;
  for (size_t i = 0; i < bb.size(); ++i) {
    const uint32_t u = bb[i];
    if (count[u] <= 4096 || leaf[u]) {
      sroots.push_back(u);
      continue;
    }
    for (uint32_t k = 0; k < nchild[u]; ++k) {
      const uint32_t ch = first_child[u] + k;
      ref_of[ch] = static_cast<uint32_t>(bb.size());
      bb.push_back(ch);
      if (group_root(ch, count[u])) groups.push_back(ch);
    }
  }
  uint32_t next = static_cast<uint32_t>(bb.size());
  std::vector<uint32_t> sub;
  for (uint32_t r : sroots) {
    sub.clear();
    sub.push_back(r);
    for (size_t k = 0; k < sub.size(); ++k) {
      const uint32_t u = sub[k];
      if (leaf[u]) continue;
      for (uint32_t q = 0; q < nchild[u]; ++q) {
        const uint32_t ch = first_child[u] + q;
        ref_of[ch] = next++;
        sub.push_back(ch);
        if (group_root(ch, count[u])) groups.push_back(ch);
      }
    }
  }
}

}  // namespace

extern "C" {

uint64_t bh_max_particles(void) { return kMaxN; }

int bh_ctx_create(bh_ctx** out, int device) {
  if (!out) return BH_ERR_INVALID;
  *out = nullptr;
  Ctx* c = new Ctx();
  c->device = device;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev || device >= kMaxDevices) {
    (void)hipGetLastError();
    delete c;
    return BH_ERR_HIP;
  }
  // pinned staging sized so that a one-system table of the allocation-free
  // step never has to grow
  const size_t table0 = 64 << 10;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->host_misc), (kMiscWords + 4) * sizeof(uint32_t), hipHostMallocDefault) !=
          hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->host_tables), table0, hipHostMallocDefault) != hipSuccess ||
      scratch_address(device, &c->arena) != hipSuccess) {
    (void)hipGetLastError();
    if (c->host_tables) (void)hipHostFree(c->host_tables);
    if (c->host_misc) (void)hipHostFree(c->host_misc);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return BH_ERR_HIP;
  }
  c->host_tables_cap = table0;
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      (void)hipGetLastError();
      for (auto& f : c->ev)
        if (f) (void)hipEventDestroy(f);
      (void)hipHostFree(c->host_tables);
      (void)hipHostFree(c->host_misc);
      (void)hipStreamDestroy(c->stream);
      delete c;
      return BH_ERR_HIP;
    }
  }
  *out = reinterpret_cast<bh_ctx*>(c);
  return BH_OK;
}

void bh_ctx_destroy(bh_ctx* h) {
  Ctx* c = as(h);
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graphs(c);
  release(c->pa);
  release(c->carry);
  release(c->carry_flag);
  release(c->pb);
  release(c->tmp_in);
  release(c->tmp_out);
  release(c->keys_a);
  release(c->keys_b);
  release(c->vals_a);
  release(c->vals_b);
  release(c->hist);
  release(c->dbase);
  release(c->lookback);
  release(c->lohi);
  release(c->partial);
  release(c->misc);
  release(c->totals);
  release(c->posm);
  release(c->L);
  release(c->H);
  release(c->tcounts);
  release(c->tprefix);
  release(c->mark);
  release(c->targets);
  release(c->gmark);
  release(c->groups);
  release(c->unit_n);
  release(c->defer);
  release(c->deep_frames);
  release(c->gacc);
  release(c->nbox);
  release(c->begin);
  release(c->count);
  release(c->cs);
  release(c->first_child);
  release(c->parent);
  release(c->leaf_idx);
  release(c->leaf_node);
  release(c->leaf_first);
  release(c->level_off);
  release(c->depth);
  release(c->leaf);
  release(c->nchild);
  release(c->ctr);
  release(c->rec);
  release(c->mass);
  release(c->com);
  release(c->iab);
  release(c->q64);
  release(c->tables);
  release(c->sid);
  release(c->dsrc);
  release(c->didx);
  release(c->dcnt);
  release(c->dpart);
  release(c->dacc);
  if (c->host_tables) (void)hipHostFree(c->host_tables);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->host_misc) (void)hipHostFree(c->host_misc);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* bh_last_error(const bh_ctx* h) {
  const Ctx* c = reinterpret_cast<const Ctx*>(h);
  return c ? c->err.c_str() : "null context";
}

int bh_set_precision(bh_ctx* h, int precision) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (precision != BH_FP32 && precision != BH_FP64) {
    set_err(c, "unknown precision %d", precision);
    return BH_ERR_INVALID;
  }
  c->precision = precision;
  return BH_OK;
}

int bh_set_walk_mode(bh_ctx* h, int mode) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (mode != BH_WALK_GROUPED && mode != BH_WALK_PER_TARGET) {
    set_err(c, "unknown walk mode %d", mode);
    return BH_ERR_INVALID;
  }
  c->walk_mode = mode;
  return BH_OK;
}

int bh_set_graph_mode(bh_ctx* h, int mode) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (mode != BH_GRAPH_OFF && mode != BH_GRAPH_ON) {
    set_err(c, "unknown graph mode %d", mode);
    return BH_ERR_INVALID;
  }
  if (mode == BH_GRAPH_OFF) {
    (void)hipSetDevice(c->device);
    drop_graphs(c);
  }
  c->graph_mode = mode;
  return BH_OK;
}

int bh_set_walk_tile(bh_ctx* h, uint32_t tile) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (tile < kMinTile || tile > kGroupCap) {
    set_err(c, "walk tile %u outside %u..%u", tile, kMinTile, kGroupCap);
    return BH_ERR_INVALID;
  }
  c->walk_tile = tile;
  return BH_OK;
}

int bh_debug_walk_stack_limit(bh_ctx* h, uint32_t levels) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (levels == 0 || levels > static_cast<uint32_t>(kLevels)) {
    set_err(c, "walk stack limit %u outside 1..%d", levels, kLevels);
    return BH_ERR_INVALID;
  }
  drop_graphs(c);  // the limit is a launch argument
  c->stack_limit = levels;
  return BH_OK;
}

int bh_graph_count(bh_ctx* h, uint32_t* count) {
  Ctx* c = as(h);
  if (!c || !count) return BH_ERR_INVALID;
  *count = static_cast<uint32_t>(c->graphs.size());
  return BH_OK;
}

int bh_step_device(bh_ctx* h, const bh_particle* d_in, bh_particle* d_out, uint64_t n, const bh_force_params* fp,
                   uint32_t max_leaf, uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt,
                   bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !fp || (n && (!d_in || !d_out))) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  if (overlap(d_in, d_out, n)) {
    set_err(c, "d_in and d_out overlap");
    return BH_ERR_INVALID;
  }
  return step_device(c, d_in, d_out, n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, timing);
}

int bh_step_device_noalloc(bh_ctx* h, const bh_particle* d_in, bh_particle* d_out, uint64_t n, const bh_force_params* fp,
                           uint32_t max_leaf, uint32_t group_size, uint32_t target_size, double kick_dt,
                           double drift_dt, bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !fp || (n && (!d_in || !d_out))) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  if (timing) *timing = bh_step_timing{0.0, 0.0, 0.0};
  if (n > BH_NOALLOC_MAX_N) {
    set_err(c, "n = %llu exceeds the allocation-free step's limit of %u particles", (unsigned long long)n,
            BH_NOALLOC_MAX_N);
    return BH_ERR_TOO_LARGE;
  }
  if (overlap(d_in, d_out, n)) {
    set_err(c, "d_in and d_out overlap");
    return BH_ERR_INVALID;
  }
  TRY(check_args(c, n, fp, max_leaf, group_size));
  if (n == 0) return BH_OK;
  HIP_OK(c, hipSetDevice(c->device));
  const HostSys one{n, *fp, max_leaf, group_size, target_size, kick_dt, drift_dt, 0};
  Plan p;
  TRY(plan_systems(c, &one, 1, true, &p));
  // every level and the deep stack are planned from the start: the only
  // thing that can outgrow the fixed scratch is the node count
  p.walk = true;
  p.depth_plan = kDepthCap;
  p.deep = true;
  p.slots = walk_slots(p.n, kNoAllocSlots);
  Scratch b;
  if (!carve(static_cast<uint8_t*>(c->arena), kNoAllocBytes, &p, &b)) {
    set_err(c, "n = %llu does not fit the allocation-free step's %zu-byte scratch", (unsigned long long)n, kNoAllocBytes);
    return BH_ERR_TOO_LARGE;
  }
  // the tree lives in the shared scratch: it is not this context's to export
  c->has_tree = false;
  c->have_stats = false;
  c->reruns = 0;
  Sorted sd;
  HIP_OK(c, hipEventRecord(c->ev[0], c->stream));
  TRY(enqueue_build(c, p, b, d_in, d_out, true, &sd, true));
  TRY(enqueue_walk(c, p, b, d_out, nullptr, nullptr, nullptr, true, d_in, sd.perm));
  TRY(finish(c, b));
  const uint32_t* m = c->host_misc;
  if (m[kMiscNonfinite]) {
    set_err(c, "non-finite position, velocity or mass in the input");
    return BH_ERR_NONFINITE;
  }
  if (m[kMiscAbort]) {
    uint64_t nodes = 0;
    for (int d = 0; d < kLevels; ++d) nodes += m[kMiscLevels + d];
    set_err(c, "the tree needs %llu nodes and %u leaves; the allocation-free scratch holds %u",
            (unsigned long long)nodes, m[kMiscLevels + kLevels], p.cap_nodes);
    return BH_ERR_TOO_LARGE;
  }
  c->n_targets = m[kMiscTargets];
  c->n_groups = m[kMiscGroups];
  std::memcpy(c->last_totals, m + kMiscWords, sizeof c->last_totals);
  uint64_t nodes = 0;
  for (int d = 0; d < kLevels; ++d) nodes += m[kMiscLevels + d];
  c->n_nodes = static_cast<uint32_t>(nodes);
  c->n_leaves = m[kMiscLevels + kLevels];
  c->have_stats = true;
  c->last_graph = false;
  if (timing) {
    timing->tree_build_ms = elapsed(c->ev[0], c->ev[2]);
    timing->traversal_ms = elapsed(c->ev[2], c->ev[5]);
  }
  if (m[kMiscIncomplete]) {
    set_err(c, "incomplete tree: the walk met an interior node it could not open");
    return BH_ERR_INCOMPLETE;
  }
  return BH_OK;
}

int bh_step(bh_ctx* h, bh_particle* particles, uint64_t n, const bh_force_params* fp, uint32_t max_leaf,
            uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt, bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !fp || (n && !particles)) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  TRY(check_args(c, n, fp, max_leaf, group_size));
  if (n == 0) {
    if (timing) *timing = bh_step_timing{0.0, 0.0, 0.0};
    return BH_OK;
  }
  HIP_OK(c, hipSetDevice(c->device));
  TRY(ensure(c, c->tmp_in, n, "particles"));
  TRY(ensure(c, c->tmp_out, n, "particles"));
  HIP_OK(c, hipMemcpyAsync(c->tmp_in.p, particles, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  TRY(step_device(c, c->tmp_in.p, c->tmp_out.p, n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, timing));
  HIP_OK(c, hipMemcpyAsync(particles, c->tmp_out.p, n * sizeof(bh_particle), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_step_batch(bh_ctx* h, const bh_particle* d_in, bh_particle* d_out, const bh_system* systems, uint32_t n_systems,
                  bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !systems || !d_in || !d_out) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  const std::vector<HostSys> v = host_systems(systems, n_systems);
  uint64_t n = 0;
  for (const auto& q : v) n += q.n;
  if (overlap(d_in, d_out, n)) {
    set_err(c, "d_in and d_out overlap");
    return BH_ERR_INVALID;
  }
  return step_batch(c, d_in, d_out, v.data(), n_systems, timing);
}

int bh_step_batch_host(bh_ctx* h, bh_particle* particles, const bh_system* systems, uint32_t n_systems,
                       bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !systems || !particles) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  const std::vector<HostSys> v = host_systems(systems, n_systems);
  uint64_t n = 0;
  for (const auto& q : v) n += q.n;
  if (n > kMaxN) {
    set_err(c, "batch of %llu particles exceeds the supported %llu", (unsigned long long)n, (unsigned long long)kMaxN);
    return BH_ERR_TOO_LARGE;
  }
  HIP_OK(c, hipSetDevice(c->device));
  TRY(ensure(c, c->tmp_in, n, "particles"));
  TRY(ensure(c, c->tmp_out, n, "particles"));
  if (n) HIP_OK(c, hipMemcpyAsync(c->tmp_in.p, particles, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  TRY(step_batch(c, c->tmp_in.p, c->tmp_out.p, v.data(), n_systems, timing));
  if (n) HIP_OK(c, hipMemcpyAsync(particles, c->tmp_out.p, n * sizeof(bh_particle), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_state_upload(bh_ctx* h, const bh_particle* host, uint64_t n) {
  Ctx* c = as(h);
  if (!c || (n && !host)) return BH_ERR_INVALID;
  if (n > kMaxN) {
    set_err(c, "n = %llu exceeds the supported %llu particles", (unsigned long long)n, (unsigned long long)kMaxN);
    return BH_ERR_TOO_LARGE;
  }
  HIP_OK(c, hipSetDevice(c->device));
  TRY(ensure(c, c->pa, n, "particles"));
  TRY(ensure(c, c->pb, n, "particles"));
  c->carried_ok = false;
  if (n) HIP_OK(c, hipMemcpyAsync(c->pa.p, host, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  c->cur = 0;
  c->state_n = n;
  return BH_OK;
}

int bh_state_step(bh_ctx* h, const bh_force_params* fp, uint32_t max_leaf, uint32_t group_size, uint32_t target_size,
                  double kick_dt, double drift_dt, bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !fp) return BH_ERR_INVALID;
  if (c->state_n == 0) {
    set_err(c, "no resident state: upload particles with bh_state_upload first");
    return BH_ERR_INVALID;
  }
  bh_particle* in = c->cur == 0 ? c->pa.p : c->pb.p;
  bh_particle* out = c->cur == 0 ? c->pb.p : c->pa.p;
  TRY(step_device(c, in, out, c->state_n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, timing, true));
  c->cur ^= 1;
  return BH_OK;
}

int bh_state_step_batch(bh_ctx* h, const bh_system* systems, uint32_t n_systems, bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !systems) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  if (c->state_n == 0) {
    set_err(c, "no resident state: upload particles with bh_state_upload first");
    return BH_ERR_INVALID;
  }
  const std::vector<HostSys> v = host_systems(systems, n_systems);
  uint64_t n = 0;
  for (const auto& q : v) n += q.n;
  if (n != c->state_n) {
    set_err(c, "the systems hold %llu particles, the resident state %llu", (unsigned long long)n,
            (unsigned long long)c->state_n);
    return BH_ERR_INVALID;
  }
  bh_particle* in = c->cur == 0 ? c->pa.p : c->pb.p;
  bh_particle* out = c->cur == 0 ? c->pb.p : c->pa.p;
  TRY(step_batch(c, in, out, v.data(), n_systems, timing, true));
  c->cur ^= 1;
  return BH_OK;
}

bh_particle* bh_state_device_ptr(bh_ctx* h) {
  Ctx* c = as(h);
  if (!c) return nullptr;
  // the caller may write the state: the next step reads it afresh
  c->carried_ok = false;
  return c->cur == 0 ? c->pa.p : c->pb.p;
}

int bh_state_download(bh_ctx* h, bh_particle* host, uint64_t n) {
  Ctx* c = as(h);
  if (!c || (n && !host)) return BH_ERR_INVALID;
  if (n != c->state_n) {
    set_err(c, "resident state holds %llu particles, %llu requested", (unsigned long long)c->state_n,
            (unsigned long long)n);
    return BH_ERR_INVALID;
  }
  HIP_OK(c, hipSetDevice(c->device));
  if (n)
    HIP_OK(c, hipMemcpyAsync(host, bh_state_device_ptr(h), n * sizeof(bh_particle), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_build_tree(bh_ctx* h, const bh_particle* host, uint64_t n, uint32_t max_leaf, uint32_t group_size) {
  Ctx* c = as(h);
  if (!c || (n && !host)) return BH_ERR_INVALID;
  TRY(check_args(c, n, nullptr, max_leaf, group_size));
  c->has_tree = false;
  if (n == 0) {
    set_err(c, "empty input");
    return BH_ERR_INVALID;
  }
  HIP_OK(c, hipSetDevice(c->device));
  TRY(ensure(c, c->tmp_in, n, "particles"));
  TRY(ensure(c, c->tmp_out, n, "particles"));
  HIP_OK(c, hipMemcpyAsync(c->tmp_in.p, host, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  const bh_force_params dflt{0.5, 0.0, 1.0, 1, 0};
  const HostSys one{n, dflt, max_leaf, group_size, group_size, 0.0, 0.0, 0};
  Plan p;
  TRY(plan_systems(c, &one, 1, false, &p));
  return run(c, p, c->tmp_in.p, c->tmp_out.p, false, nullptr);
}

int bh_tree_sizes(bh_ctx* h, uint64_t* n_nodes, uint64_t* n_groups, double* dom4) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (!c->has_tree) {
    set_err(c, "no tree built");
    return BH_ERR_NO_TREE;
  }
  HIP_OK(c, hipSetDevice(c->device));
  const TreeDev& t = c->sc.t;
  if (n_nodes) *n_nodes = c->n_nodes;
  if (n_groups) {
    const uint32_t m = c->n_nodes;
    std::vector<uint32_t> count(m), parent(m);
    std::vector<uint8_t> leaf(m);
    HIP_OK(c, hipMemcpy(count.data(), t.count, m * 4ull, hipMemcpyDeviceToHost));
    HIP_OK(c, hipMemcpy(parent.data(), t.parent, m * 4ull, hipMemcpyDeviceToHost));
    HIP_OK(c, hipMemcpy(leaf.data(), t.leaf, m, hipMemcpyDeviceToHost));
    uint64_t g = 0;
    for (uint32_t u = 0; u < m; ++u) {
      const uint32_t pc = u == 0 ? std::numeric_limits<uint32_t>::max() : count[parent[u]];
      if ((leaf[u] || count[u] <= c->group_size) && pc > c->group_size) ++g;
    }
    *n_groups = g;
  }
  if (dom4) {
    double4 r;
    HIP_OK(c, hipMemcpy(&r, t.ctr, sizeof(double4), hipMemcpyDeviceToHost));
    dom4[0] = r.x;
    dom4[1] = r.y;
    dom4[2] = r.z;
    dom4[3] = r.w;
  }
  return BH_OK;
}

int bh_tree_export(bh_ctx* h, int layout, uint64_t* keys, uint32_t* perm, bh_particle* sorted, bh_walk_node* nodes,
                   double* center, double* iab, uint32_t* groups) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (!c->has_tree) {
    set_err(c, "no tree built");
    return BH_ERR_NO_TREE;
  }
  if (layout != BH_LAYOUT_REFERENCE && layout != BH_LAYOUT_BFS) {
    set_err(c, "unknown layout %d", layout);
    return BH_ERR_INVALID;
  }
  if (c->plan.S != 1) {
    set_err(c, "tree export covers single-system trees");
    return BH_ERR_INVALID;
  }
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  const TreeDev& t = c->sc.t;
  const uint32_t n = c->n, m = c->n_nodes;
  if (keys) HIP_OK(c, hipMemcpy(keys, c->keys_sorted, uint64_t(n) * 8, hipMemcpyDeviceToHost));
  if (perm) HIP_OK(c, hipMemcpy(perm, c->perm_sorted, uint64_t(n) * 4, hipMemcpyDeviceToHost));
  if (sorted) HIP_OK(c, hipMemcpy(sorted, c->sorted, uint64_t(n) * sizeof(bh_particle), hipMemcpyDeviceToHost));
  std::vector<uint32_t> begin(m), count(m), first_child(m), parent(m);
  std::vector<uint8_t> depth(m), leaf(m), nchild(m);
  std::vector<double4> ctr(m);
  std::vector<double> mass(m), com(3ull * m), ib(6ull * m);
  HIP_OK(c, hipMemcpy(begin.data(), t.begin, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(count.data(), t.count, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(first_child.data(), t.first_child, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(parent.data(), t.parent, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(depth.data(), t.depth, m, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(leaf.data(), t.leaf, m, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(nchild.data(), t.nchild, m, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(ctr.data(), t.ctr, m * sizeof(double4), hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(mass.data(), t.mass, m * 8ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(com.data(), t.com, m * 24ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(ib.data(), t.iab, m * 48ull, hipMemcpyDeviceToHost));
  std::vector<uint32_t> ref_of, grp;
  if (layout == BH_LAYOUT_REFERENCE) {
    reference_order(count, leaf, first_child, nchild, c->group_size, ref_of, grp);
  } else {
    ref_of.resize(m);
    for (uint32_t u = 0; u < m; ++u) {
      ref_of[u] = u;
      const uint32_t pc = u == 0 ? std::numeric_limits<uint32_t>::max() : count[parent[u]];
      if ((leaf[u] || count[u] <= c->group_size) && pc > c->group_size) grp.push_back(u);
    }
  }
  for (uint32_t u = 0; u < m; ++u) {
    const uint32_t r = ref_of[u];
    if (nodes) {
      bh_walk_node w;
      std::memset(&w, 0, sizeof w);
      for (int a = 0; a < 3; ++a) w.com[a] = com[3ull * u + a];
      w.mass = mass[u];
      w.half_width = ctr[u].w;
      w.first_child = leaf[u] ? 0u : ref_of[first_child[u]];
      w.particle_begin = begin[u];
      w.particle_count = count[u];
      w.num_children = leaf[u] ? 0 : nchild[u];
      w.leaf = leaf[u];
      w.depth = depth[u];
      nodes[r] = w;
    }
    if (center) {
      center[3ull * r + 0] = ctr[u].x;
      center[3ull * r + 1] = ctr[u].y;
      center[3ull * r + 2] = ctr[u].z;
    }
    if (iab)
      for (int a = 0; a < 6; ++a) iab[6ull * r + a] = ib[6ull * u + a];
  }
  if (groups)
    for (size_t g = 0; g < grp.size(); ++g) groups[g] = ref_of[grp[g]];
  return BH_OK;
}

int bh_accel(bh_ctx* h, const bh_force_params* fp, uint32_t target_size, double* acc, uint32_t* n_pp, uint32_t* n_pc,
             uint64_t* totals) {
  Ctx* c = as(h);
  if (!c || !fp) return BH_ERR_INVALID;
  if (!c->has_tree) {
    set_err(c, "no tree built");
    return BH_ERR_NO_TREE;
  }
  TRY(check_args(c, c->n, fp, c->max_leaf, c->group_size));
  if (c->plan.S != 1) {
    set_err(c, "bh_accel covers single-system trees");
    return BH_ERR_INVALID;
  }
  HIP_OK(c, hipSetDevice(c->device));
  const uint32_t n = c->n;
  TRY(ensure(c, c->dacc, 3ull * n, "accelerations"));
  TRY(ensure(c, c->dcnt, 2ull * n, "interaction counts"));
  // the walk's buffers for the built tree (the build planned none)
  Plan p = c->plan;
  p.walk = true;
  Scratch b;
  TRY(provision(c, p, &b));
  if (b.t.begin != c->sc.t.begin || b.keys_a != c->sc.keys_a || b.misc != c->sc.misc) {
    set_err(c, "the tree's buffers moved since it was built");
    return BH_ERR_NO_TREE;
  }
  // the walk's parameters for the built tree: rewrite its system entry
  HIP_OK(c, hipStreamSynchronize(c->stream));
  SysDev* sd = reinterpret_cast<SysDev*>(c->host_tables);
  sd->theta2 = fp->theta * fp->theta;
  sd->eps2 = fp->eps2;
  sd->G = fp->G;
  sd->kick_dt = 0.0;
  sd->drift_dt = 0.0;
  sd->tsm = std::min(target_size, c->group_size);
  sd->use_quad = fp->use_quadrupole != 0;
  p.unsoftened = static_cast<float>(fp->eps2) == 0.0f;
  HIP_OK(c, hipMemcpyAsync(b.tables, c->host_tables, sizeof(SysDev), hipMemcpyHostToDevice, c->stream));
  c->have_stats = false;
  c->reruns = 0;
  HIP_OK(c, hipEventRecord(c->ev[0], c->stream));
  HIP_OK(c, hipEventRecord(c->ev[1], c->stream));
  HIP_OK(c, hipEventRecord(c->ev[2], c->stream));
  TRY(enqueue_walk(c, p, b, nullptr, c->dacc.p, c->dcnt.p, c->dcnt.p + n, true));
  TRY(finish(c, b));
  const uint32_t* m = c->host_misc;
  c->n_targets = m[kMiscTargets];
  c->n_groups = m[kMiscGroups];
  std::memcpy(c->last_totals, m + kMiscWords, sizeof c->last_totals);
  c->have_stats = true;
  c->last_graph = false;
  if (acc) HIP_OK(c, hipMemcpy(acc, c->dacc.p, 24ull * n, hipMemcpyDeviceToHost));
  if (n_pp) HIP_OK(c, hipMemcpy(n_pp, c->dcnt.p, 4ull * n, hipMemcpyDeviceToHost));
  if (n_pc) HIP_OK(c, hipMemcpy(n_pc, c->dcnt.p + n, 4ull * n, hipMemcpyDeviceToHost));
  if (totals) {
    totals[0] = c->last_totals[0];
    totals[1] = c->last_totals[1];
  }
  if (m[kMiscIncomplete]) {
    set_err(c, "incomplete tree: the walk met an interior node it could not open");
    return BH_ERR_INCOMPLETE;
  }
  return BH_OK;
}

int bh_last_step_stats(bh_ctx* h, bh_step_stats* st) {
  Ctx* c = as(h);
  if (!c || !st) return BH_ERR_INVALID;
  if (!c->have_stats) {
    set_err(c, "no step or walk has run on this context");
    return BH_ERR_NO_TREE;
  }
  HIP_OK(c, hipSetDevice(c->device));
  if (c->last_graph) {
    st->sort_ms = st->tree_ms = st->targets_ms = st->walk_ms = 0.0;
  } else {
    st->sort_ms = elapsed(c->ev[0], c->ev[1]);
    st->tree_ms = elapsed(c->ev[1], c->ev[2]);
    st->targets_ms = elapsed(c->ev[2], c->ev[3]);
    st->walk_ms = elapsed(c->ev[4], c->ev[5]);
  }
  st->step_ms = elapsed(c->ev[0], c->ev[5]);
  st->n_pp = c->last_totals[0];
  st->n_pc = c->last_totals[1];
  st->n_nodes = c->n_nodes;
  st->n_leaves = c->n_leaves;
  st->n_targets = c->n_targets;
  st->n_units = c->n_groups;
  st->graph = c->last_graph ? 1 : 0;
  st->reruns = c->reruns;
  return BH_OK;
}

int bh_direct_sum(bh_ctx* h, const bh_particle* host, uint64_t n, const uint32_t* idx, uint64_t m, double eps2, double G,
                  double* acc) {
  Ctx* c = as(h);
  if (!c || !host || !acc) return BH_ERR_INVALID;
  if (n > kMaxN || m > kMaxN) {
    set_err(c, "too many particles");
    return BH_ERR_TOO_LARGE;
  }
  if (!idx) m = n;
  if (n == 0 || m == 0) return BH_OK;
  HIP_OK(c, hipSetDevice(c->device));
  std::vector<double4> src(n);
  for (uint64_t i = 0; i < n; ++i) src[i] = make_double4(host[i].pos[0], host[i].pos[1], host[i].pos[2], host[i].mass);
  if (idx)
    for (uint64_t k = 0; k < m; ++k)
      if (idx[k] >= n) {
        set_err(c, "target index %u out of range", idx[k]);
        return BH_ERR_INVALID;
      }
  const uint32_t splits = direct_splits(static_cast<uint32_t>(n), static_cast<uint32_t>(m));
  TRY(ensure(c, c->dsrc, n, "direct sum"));
  TRY(ensure(c, c->didx, m, "direct sum"));
  TRY(ensure(c, c->dpart, 3ull * m * splits, "direct sum"));
  TRY(ensure(c, c->dacc, 3ull * m, "direct sum"));
  HIP_OK(c, hipMemcpyAsync(c->dsrc.p, src.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->stream));
  if (idx) HIP_OK(c, hipMemcpyAsync(c->didx.p, idx, m * 4, hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, launch_direct(c->dsrc.p, static_cast<uint32_t>(n), idx ? c->didx.p : nullptr, static_cast<uint32_t>(m), eps2, G,
                          c->dpart.p, splits, c->dacc.p, c->stream));
  HIP_OK(c, hipMemcpyAsync(acc, c->dacc.p, 24 * m, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_energies(bh_ctx* h, const bh_particle* host, uint64_t n, double G, double* kinetic, double* potential) {
  Ctx* c = as(h);
  if (!c || (n && !host)) return BH_ERR_INVALID;
  if (n > kMaxN) {
    set_err(c, "too many particles");
    return BH_ERR_TOO_LARGE;
  }
  if (kinetic) {
    // integrate.h:27-31, summed in particle order on the host (O(n))
    double e = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
      const double v2 = host[i].vel[0] * host[i].vel[0] + host[i].vel[1] * host[i].vel[1] + host[i].vel[2] * host[i].vel[2];
      e += 0.5 * host[i].mass * v2;
    }
    *kinetic = e;
  }
  if (!potential) return BH_OK;
  *potential = 0.0;
  if (n < 2) return BH_OK;
  HIP_OK(c, hipSetDevice(c->device));
  std::vector<double4> src(n);
  for (uint64_t i = 0; i < n; ++i) src[i] = make_double4(host[i].pos[0], host[i].pos[1], host[i].pos[2], host[i].mass);
  const uint32_t blocks = static_cast<uint32_t>((n + 255) / 256);
  uint32_t splits = 2048 / blocks;
  if (splits < 1) splits = 1;
  if (splits > 512) splits = 512;
  while (splits > 1 && n / splits < 256) splits >>= 1;
  TRY(ensure(c, c->dsrc, n, "energies"));
  TRY(ensure(c, c->dpart, uint64_t(blocks) * splits, "energies"));
  HIP_OK(c, hipMemcpyAsync(c->dsrc.p, src.data(), n * sizeof(double4), hipMemcpyHostToDevice, c->stream));
  HIP_OK(c, launch_potential(c->dsrc.p, static_cast<uint32_t>(n), c->dpart.p, splits, c->stream));
  std::vector<double> part(uint64_t(blocks) * splits);
  HIP_OK(c, hipMemcpyAsync(part.data(), c->dpart.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  // the kernel counts every unordered pair once
  double sum = 0.0;
  for (double v : part) sum += v;
  *potential = -G * sum;
  return BH_OK;
}

}  // extern "C"
