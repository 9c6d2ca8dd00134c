// Host side of the C ABI (include/bh_gpu.h): the context, its grow-only
// device buffers, and the per-step orchestration of the kernels.
//
// One step (step_particles, integrate_ttg.h:88-107, grouped branch):
//   bounds -> keys + digit histograms -> 8 onesweep radix passes -> gather
//   (Morton-ordered particles + FP64 walk copy) -> tree counts -> [one
//   16-word read-back: node/leaf totals, non-finite flag] -> node emit ->
//   links/centres -> upward pass (per level) -> targets -> walk with fused
//   kick/drift. Everything is ordered on the context's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "bh_internal.h"

using namespace bhk;

namespace {

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // events: 0 step start, 1 sorted, 2 tree built, 3 targets chosen, 4 walk start, 5 walk end
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool have_stats = false;
  std::string err;
  int precision = BH_FP32;
  int walk_mode = BH_WALK_GROUPED;

  DevBuf<bh_particle> pa, pb;   // particle ping-pong (resident state / step output)
  int cur = 0;                  // which of pa/pb holds the resident state
  uint64_t state_n = 0;
  DevBuf<bh_particle> tmp_in;   // host-upload staging for bh_step / bh_build_tree
  DevBuf<uint64_t> keys_a, keys_b;
  DevBuf<uint32_t> vals_a, vals_b;
  DevBuf<uint32_t> hist, dbase, lookback;
  DevBuf<double> lohi, partial;
  DevBuf<uint32_t> misc;        // [0] non-finite flag, [8..32) tree totals, [40] n_targets
  DevBuf<unsigned long long> totals;
  DevBuf<double4> posm;
  DevBuf<uint8_t> L, H;
  DevBuf<uint32_t> tcounts, tprefix;
  DevBuf<uint32_t> mark, targets;   // walk targets: marks per particle, list
  DevBuf<uint32_t> gmark, groups;   // walk groups: marks per particle, list
  DevBuf<uint32_t> defer;           // deferred nodes per group
  DevBuf<uint32_t> deep_frames;     // walk stack frames of deep trees, per group
  DevBuf<double> gacc;              // group sums per particle
  DevBuf<double> nbox;              // particle box per node
  DevBuf<uint32_t> unit_n;          // walk units packed from single-target groups, per node
  // tree (level-order BFS)
  DevBuf<uint32_t> begin, count, cs, first_child, parent, leaf_idx, leaf_node, leaf_first, level_off;
  DevBuf<uint8_t> depth, leaf, nchild;
  DevBuf<double4> ctr;
  DevBuf<NodeRec> rec;
  DevBuf<double> mass, com, iab, q64;
  // direct sum
  DevBuf<double4> dsrc;
  DevBuf<uint32_t> didx;
  DevBuf<double> dpart, dacc;

  uint32_t*  host_misc = nullptr;   // pinned
  // system table of the current batch: device copy and pinned staging
  DevBuf<uint8_t> tables;           // SysDev[S] | uint4 chunks[C] | uint32 chunk_first[S + 1] | uint32 sid hist[2][256]
  uint8_t* host_tables = nullptr;
  size_t host_tables_cap = 0;
  DevBuf<uint16_t> sid;
  uint32_t n_sys = 0, n_chunks = 0, mlx = 0;
  int sid_passes = 0;
  bool any_unsoftened = false;
  size_t off_chunks = 0, off_chunk_first = 0, off_sid_hist = 0;
  // last built tree
  bool has_tree = false;
  uint32_t n = 0, n_nodes = 0, n_leaves = 0, max_depth = 0;
  uint32_t max_leaf = 0, group_size = 0;
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

template <typename T>
int ensure(Ctx* c, DevBuf<T>& b, size_t n, const char* what) {
  if (n == 0) n = 1;
  if (b.cap >= n) return BH_OK;
  if (b.p) (void)hipFree(b.p);
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

#define TRY(expr)            \
  do {                       \
    int rc_ = (expr);        \
    if (rc_ != BH_OK) return rc_; \
  } while (0)


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

TreeDev tree_view(Ctx* c) {
  TreeDev t;
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
  return t;
}

// The batch's system table: per-system parameters, chunks of <= kChunk
// particles for the per-system bounds, and the system-id digit counts of the
// sort's last passes. Staged in pinned memory, uploaded with one copy.
int prepare_systems(Ctx* c, const HostSys* sys, uint32_t S, uint64_t* n_total) {
  if (S == 0 || S > kMaxSystems) {
    set_err(c, "a batch holds 1..%u systems, %u given", kMaxSystems, S);
    return BH_ERR_INVALID;
  }
  // Empty systems take no part: the device table lists the others, in
  // order, and a particle's system id indexes that list.
  uint64_t n = 0, chunks = 0;
  uint32_t mlx = 0, live = 0;
  bool unsoftened = false;
  for (uint32_t k = 0; k < S; ++k) {
    TRY(check_args(c, sys[k].n, &sys[k].fp, sys[k].max_leaf, sys[k].group_size));
    if (sys[k].n == 0) continue;
    ++live;
    n += sys[k].n;
    chunks += (sys[k].n + kChunk - 1) / kChunk;
    mlx = std::max(mlx, sys[k].max_leaf);
    unsoftened |= static_cast<float>(sys[k].fp.eps2) == 0.0f;  // zero in the FP32 walk too
  }
  if (n > kMaxN) {
    set_err(c, "batch of %llu particles exceeds the supported %llu", (unsigned long long)n, (unsigned long long)kMaxN);
    return BH_ERR_TOO_LARGE;
  }
  *n_total = n;
  if (n == 0) return BH_OK;
  const uint32_t S_all = S;
  S = live;
  const size_t off_chunks = sizeof(SysDev) * S;
  const size_t off_first = off_chunks + sizeof(uint4) * chunks;
  const size_t off_hist = (off_first + 4 * (S + 1) + 15) / 16 * 16;
  const size_t bytes = off_hist + 2 * 256 * 4;
  if (c->host_tables_cap < bytes) {
    if (c->host_tables) (void)hipHostFree(c->host_tables);
    c->host_tables = nullptr;
    c->host_tables_cap = 0;
    const size_t want = bytes + bytes / 2;
    if (hipHostMalloc(reinterpret_cast<void**>(&c->host_tables), want, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      set_err(c, "pinned host allocation of %zu bytes failed", want);
      return BH_ERR_OOM;
    }
    c->host_tables_cap = want;
  }
  std::memset(c->host_tables, 0, bytes);
  SysDev* sd = reinterpret_cast<SysDev*>(c->host_tables);
  uint4* ch = reinterpret_cast<uint4*>(c->host_tables + off_chunks);
  uint32_t* first = reinterpret_cast<uint32_t*>(c->host_tables + off_first);
  uint32_t* hist = reinterpret_cast<uint32_t*>(c->host_tables + off_hist);
  uint64_t off = 0, ci = 0;
  for (uint32_t k = 0, src = 0; k < S; ++k, ++src) {
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
    first[k] = static_cast<uint32_t>(ci);
    for (uint64_t b = 0; b < h.n; b += kChunk)
      ch[ci++] = make_uint4(k, static_cast<uint32_t>(off + b), static_cast<uint32_t>(off + std::min<uint64_t>(h.n, b + kChunk)), 0);
    hist[k & 255] += static_cast<uint32_t>(h.n);
    hist[256 + (k >> 8)] += static_cast<uint32_t>(h.n);
    off += h.n;
  }
  first[S] = static_cast<uint32_t>(ci);
  TRY(ensure(c, c->tables, bytes, "system table"));
  TRY(ensure(c, c->sid, n, "system ids"));
  HIP_OK(c, hipMemcpyAsync(c->tables.p, c->host_tables, bytes, hipMemcpyHostToDevice, c->stream));
  c->n_sys = S;
  c->n_chunks = static_cast<uint32_t>(chunks);
  c->mlx = mlx;
  c->sid_passes = S == 1 ? 0 : (S <= 256 ? 1 : 2);
  c->any_unsoftened = unsoftened;
  c->off_chunks = off_chunks;
  c->off_chunk_first = off_first;
  c->off_sid_hist = off_hist;
  // single-system queries (export, accel) read the first system's shape
  for (uint32_t k = 0; k < S_all; ++k) {
    if (sys[k].n == 0) continue;
    c->max_leaf = sys[k].max_leaf;
    c->group_size = sys[k].group_size;
    break;
  }
  return BH_OK;
}

const SysDev* sys_dev(Ctx* c) { return reinterpret_cast<const SysDev*>(c->tables.p); }

// build_flat_tree (flat_build_ttg.h:276-281) of every system of the prepared
// batch, d_in[0..n) into the context; the Morton-ordered particles (each
// system in its own range) go to d_out.
int build(Ctx* c, const bh_particle* d_in, bh_particle* d_out, uint32_t n) {
  hipStream_t s = c->stream;
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  TRY(ensure(c, c->keys_a, n, "keys"));
  TRY(ensure(c, c->keys_b, n, "keys"));
  TRY(ensure(c, c->vals_a, n, "indices"));
  TRY(ensure(c, c->vals_b, n, "indices"));
  TRY(ensure(c, c->hist, 10 * 256, "histograms"));
  TRY(ensure(c, c->dbase, 10 * 256, "digit offsets"));
  TRY(ensure(c, c->lookback, radix_lookback_words(n), "look-back"));
  TRY(ensure(c, c->lohi, 6ull * c->n_sys, "bounds"));
  TRY(ensure(c, c->partial, 6ull * c->n_chunks, "bounds"));
  TRY(ensure(c, c->misc, 64, "flags"));
  TRY(ensure(c, c->posm, n, "walk copy"));
  TRY(ensure(c, c->L, n, "tree build"));
  TRY(ensure(c, c->H, n, "tree build"));
  TRY(ensure(c, c->tcounts, uint64_t(tiles) * kTreeCols, "tree build"));
  TRY(ensure(c, c->tprefix, uint64_t(tiles) * kTreeCols, "tree build"));
  TRY(ensure(c, c->level_off, kLevels + 2, "tree levels"));

  const uint4* chunks = reinterpret_cast<const uint4*>(c->tables.p + c->off_chunks);
  const uint32_t* chunk_first = reinterpret_cast<const uint32_t*>(c->tables.p + c->off_chunk_first);
  HIP_OK(c, launch_fill_u32(c->misc.p, 0u, 64, s));
  HIP_OK(c, launch_fill_u32(c->hist.p, 0u, 8 * 256, s));
  if (c->sid_passes)
    HIP_OK(c, hipMemcpyAsync(c->hist.p + 8 * 256, c->host_tables + c->off_sid_hist, 2 * 256 * 4, hipMemcpyHostToDevice, s));
  HIP_OK(c, launch_fill_u32(c->lookback.p, 0u, radix_lookback_words(n), s));
  HIP_OK(c, launch_fill_sid(chunks, c->n_chunks, c->sid.p, s));
  HIP_OK(c, launch_bounds(d_in, chunks, c->n_chunks, chunk_first, c->n_sys, c->partial.p, c->misc.p, c->lohi.p, s));
  HIP_OK(c, launch_keys_hist(d_in, n, c->sid.p, c->lohi.p, c->keys_a.p, c->hist.p, s));
  const uint64_t lb_words = radix_lookback_words(n);
  uint64_t* keys_sorted = nullptr;
  uint32_t* perm = nullptr;
  HIP_OK(c, launch_radix_sort(c->keys_a.p, c->vals_a.p, c->keys_b.p, c->vals_b.p, n, c->sid_passes, c->sid.p, c->hist.p,
                              c->dbase.p, c->lookback.p, c->lookback.p + (lb_words - 10), &keys_sorted, &perm, s));
  HIP_OK(c, launch_gather(d_in, d_out, perm, c->posm.p, n, s));
  HIP_OK(c, hipEventRecord(c->ev[1], s));
  HIP_OK(c, launch_tree_count(keys_sorted, n, c->sid.p, sys_dev(c), c->mlx, c->L.p, c->H.p, c->tcounts.p, s));
  HIP_OK(c, launch_tree_scan(c->tcounts.p, tiles, c->tprefix.p, c->misc.p + 8, s));
  HIP_OK(c, hipMemcpyAsync(c->host_misc, c->misc.p, 64 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(c, hipStreamSynchronize(s));
  if (c->host_misc[0] != 0) {
    set_err(c, "non-finite position, velocity or mass in the input");
    return BH_ERR_NONFINITE;
  }
  uint64_t nodes = 0;
  c->max_depth = 0;
  for (int d = 0; d < kLevels; ++d) {
    nodes += c->host_misc[8 + d];
    if (c->host_misc[8 + d]) c->max_depth = static_cast<uint32_t>(d);
  }
  const uint32_t leaves = c->host_misc[8 + kLevels];
  if (nodes >= (1ull << 29)) {
    set_err(c, "tree of %llu nodes exceeds the supported 2^29", (unsigned long long)nodes);
    return BH_ERR_TOO_LARGE;
  }
  const uint32_t m = static_cast<uint32_t>(nodes);
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

  TRY(ensure(c, c->leaf_node, leaves, "leaves"));
  TRY(ensure(c, c->leaf_first, leaves, "leaves"));
  const TreeDev t = tree_view(c);
  HIP_OK(c, launch_tree_emit(c->L.p, c->H.p, n, c->tprefix.p, c->misc.p + 8, t, s));
  uint32_t level_nodes[kLevels];
  for (int d = 0; d < kLevels; ++d) level_nodes[d] = c->host_misc[8 + d];
  HIP_OK(c, launch_tree_finish(keys_sorted, c->posm.p, n, c->sid.p, c->max_depth, c->lohi.p, m, leaves, level_nodes, t, s));
  HIP_OK(c, hipEventRecord(c->ev[2], s));
  c->n = n;
  c->n_nodes = m;
  c->n_leaves = leaves;
  c->sorted = d_out;
  c->keys_sorted = keys_sorted;
  c->perm_sorted = perm;
  c->has_tree = true;
  return BH_OK;
}

// run_grouped_pass (grouped_ttg.h:297-323) over every system of the
// context's tree, with each system's parameters from the system table.
int walk(Ctx* c, bh_particle* parts, double* d_acc, uint32_t* d_npp, uint32_t* d_npc) {
  hipStream_t s = c->stream;
  const uint32_t n = c->n;
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  TRY(ensure(c, c->mark, n, "targets"));
  TRY(ensure(c, c->targets, n, "targets"));
  TRY(ensure(c, c->gmark, n, "groups"));
  TRY(ensure(c, c->groups, n, "groups"));
  TRY(ensure(c, c->gacc, 3ull * n, "group sums"));
  TRY(ensure(c, c->totals, 2, "totals"));
  HIP_OK(c, launch_targets(tree_view(c), c->n_nodes, n, c->sid.p, sys_dev(c), false, c->mark.p, c->tcounts.p,
                           c->tprefix.p, c->targets.p, c->misc.p + 40, nullptr, s));
  // per-target mode: the walk groups are the targets themselves
  const bool grouped = c->walk_mode == BH_WALK_GROUPED;
  if (grouped) TRY(ensure(c, c->unit_n, c->n_nodes, "walk units"));
  HIP_OK(c, launch_targets(tree_view(c), c->n_nodes, n, c->sid.p, sys_dev(c), grouped, c->gmark.p, c->tcounts.p,
                           c->tprefix.p, c->groups.p, c->misc.p + 41, grouped ? c->unit_n.p : nullptr, s));
  HIP_OK(c, launch_fill_u32(reinterpret_cast<uint32_t*>(c->totals.p), 0u, 4, s));
  HIP_OK(c, hipEventRecord(c->ev[3], s));
  HIP_OK(c, hipMemcpyAsync(c->host_misc + 40, c->misc.p + 40, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  HIP_OK(c, hipStreamSynchronize(s));
  (void)tiles;
  const uint32_t n_groups = c->host_misc[41];
  TRY(ensure(c, c->defer, uint64_t(n_groups) * kDefer, "deferred nodes"));
  WalkArgs a;
  a.rec = c->rec.p;
  a.q64 = c->q64.p;
  a.ctr = c->ctr.p;
  a.begin = c->begin.p;
  a.count = c->count.p;
  a.posm = c->posm.p;
  a.groups = c->groups.p;
  a.unit_n = grouped ? c->unit_n.p : nullptr;
  a.tmark = c->mark.p;
  a.sid = c->sid.p;
  a.sys = sys_dev(c);
  a.gacc = c->gacc.p;
  a.nbox = c->nbox.p;
  a.defer = c->defer.p;
  a.n_groups = n_groups;
  a.n_nodes = c->n_nodes;
  a.max_depth = c->max_depth;
  a.deep_frames = nullptr;
  if (c->max_depth + 1 > kShallowFrames) {
    TRY(ensure(c, c->deep_frames, uint64_t(n_groups) * (kLevels - kShallowFrames) * kWave, "walk stack"));
    a.deep_frames = c->deep_frames.p;
  }
  a.n_parts = c->n;
  a.parts = parts;
  a.acc = d_acc;
  a.n_pp = d_npp;
  a.n_pc = d_npc;
  a.totals = c->totals.p;
  HIP_OK(c, hipEventRecord(c->ev[4], s));
  HIP_OK(c, launch_walk(a, c->precision == BH_FP64, c->any_unsoftened, s));
  HIP_OK(c, hipEventRecord(c->ev[5], s));
  c->have_stats = true;
  return BH_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// One step of a batch: systems sys[0..S) own consecutive ranges of d_in.
int step_batch(Ctx* c, const bh_particle* d_in, bh_particle* d_out, const HostSys* sys, uint32_t S, bh_step_timing* timing) {
  if (timing) *timing = bh_step_timing{0.0, 0.0, 0.0};
  uint64_t n = 0;
  HIP_OK(c, hipSetDevice(c->device));
  TRY(prepare_systems(c, sys, S, &n));
  if (n == 0) return BH_OK;
  c->have_stats = false;
  HIP_OK(c, hipEventRecord(c->ev[0], c->stream));
  TRY(build(c, d_in, d_out, static_cast<uint32_t>(n)));
  TRY(walk(c, d_out, nullptr, nullptr, nullptr));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (timing) {
    timing->tree_build_ms = elapsed(c->ev[0], c->ev[2]);
    timing->traversal_ms = elapsed(c->ev[2], c->ev[5]);
    timing->exchange_ms = 0.0;
  }
  return BH_OK;
}

int step_device(Ctx* c, const bh_particle* d_in, bh_particle* d_out, uint64_t n, const bh_force_params* fp,
                uint32_t max_leaf, uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt,
                bh_step_timing* timing) {
  TRY(check_args(c, n, fp, max_leaf, group_size));
  if (n == 0) {
    if (timing) *timing = bh_step_timing{0.0, 0.0, 0.0};
    return BH_OK;
  }
  const HostSys one{n, *fp, max_leaf, group_size, target_size, kick_dt, drift_dt};
  return step_batch(c, d_in, d_out, &one, 1, timing);
}

std::vector<HostSys> host_systems(const bh_system* systems, uint32_t S) {
  std::vector<HostSys> v(S);
  for (uint32_t k = 0; k < S; ++k)
    v[k] = HostSys{systems[k].n, systems[k].fp, systems[k].max_leaf, systems[k].group_size, systems[k].target_size,
                   systems[k].kick_dt, systems[k].drift_dt};
  return v;
}

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
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
    (void)hipGetLastError();
    delete c;
    return BH_ERR_HIP;
  }
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->host_misc), 64 * sizeof(uint32_t), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return BH_ERR_HIP;
  }
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
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
  release(c->pa);
  release(c->pb);
  release(c->tmp_in);
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
  release(c->defer);
  release(c->deep_frames);
  release(c->gacc);
  release(c->nbox);
  release(c->unit_n);
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

  release(c->dsrc);
  release(c->didx);
  release(c->dpart);
  release(c->dacc);
  release(c->tables);
  release(c->sid);
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

int bh_step_device(bh_ctx* h, const bh_particle* d_in, bh_particle* d_out, uint64_t n, const bh_force_params* fp,
                   uint32_t max_leaf, uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt,
                   bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !fp || (n && (!d_in || !d_out))) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  if (n && d_in + n > d_out && d_out + n > d_in) {
    set_err(c, "d_in and d_out overlap");
    return BH_ERR_INVALID;
  }
  return step_device(c, d_in, d_out, n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, timing);
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
  TRY(ensure(c, c->pb, n, "particles"));
  HIP_OK(c, hipMemcpyAsync(c->tmp_in.p, particles, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  TRY(step_device(c, c->tmp_in.p, c->pb.p, n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, timing));
  HIP_OK(c, hipMemcpyAsync(particles, c->pb.p, n * sizeof(bh_particle), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  c->state_n = 0;  // pb no longer holds a resident state
  return BH_OK;
}

int bh_step_batch(bh_ctx* h, const bh_particle* d_in, bh_particle* d_out, const bh_system* systems, uint32_t n_systems,
                  bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !systems || !d_in || !d_out) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
  const std::vector<HostSys> v = host_systems(systems, n_systems);
  uint64_t n = 0;
  for (const auto& q : v) n += q.n;
  if (d_in + n > d_out && d_out + n > d_in) {
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
  TRY(ensure(c, c->pb, n, "particles"));
  if (n) HIP_OK(c, hipMemcpyAsync(c->tmp_in.p, particles, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  TRY(step_batch(c, c->tmp_in.p, c->pb.p, v.data(), n_systems, timing));
  if (n) HIP_OK(c, hipMemcpyAsync(particles, c->pb.p, n * sizeof(bh_particle), hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  c->state_n = 0;
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
  bh_particle* in = c->cur == 0 ? c->pa.p : c->pb.p;
  bh_particle* out = c->cur == 0 ? c->pb.p : c->pa.p;
  TRY(step_device(c, in, out, c->state_n, fp, max_leaf, group_size, target_size, kick_dt, drift_dt, timing));
  c->cur ^= 1;
  return BH_OK;
}

int bh_state_step_batch(bh_ctx* h, const bh_system* systems, uint32_t n_systems, bh_step_timing* timing) {
  Ctx* c = as(h);
  if (!c || !systems) return c ? (set_err(c, "null argument"), BH_ERR_INVALID) : BH_ERR_INVALID;
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
  TRY(step_batch(c, in, out, v.data(), n_systems, timing));
  c->cur ^= 1;
  return BH_OK;
}

bh_particle* bh_state_device_ptr(bh_ctx* h) {
  Ctx* c = as(h);
  if (!c) return nullptr;
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
  TRY(ensure(c, c->pb, n, "particles"));
  HIP_OK(c, hipMemcpyAsync(c->tmp_in.p, host, n * sizeof(bh_particle), hipMemcpyHostToDevice, c->stream));
  c->have_stats = false;
  const bh_force_params dflt{0.5, 0.0, 1.0, 1, 0};
  const HostSys one{n, dflt, max_leaf, group_size, group_size, 0.0, 0.0};
  uint64_t total = 0;
  TRY(prepare_systems(c, &one, 1, &total));
  HIP_OK(c, hipEventRecord(c->ev[0], c->stream));
  TRY(build(c, c->tmp_in.p, c->pb.p, static_cast<uint32_t>(n)));
  c->state_n = 0;
  HIP_OK(c, hipStreamSynchronize(c->stream));
  return BH_OK;
}

int bh_tree_sizes(bh_ctx* h, uint64_t* n_nodes, uint64_t* n_groups, double* dom4) {
  Ctx* c = as(h);
  if (!c) return BH_ERR_INVALID;
  if (!c->has_tree) {
    set_err(c, "no tree built");
    return BH_ERR_NO_TREE;
  }
  HIP_OK(c, hipSetDevice(c->device));
  if (n_nodes) *n_nodes = c->n_nodes;
  if (n_groups) {
    const uint32_t m = c->n_nodes;
    std::vector<uint32_t> count(m), parent(m);
    std::vector<uint8_t> leaf(m);
    HIP_OK(c, hipMemcpy(count.data(), c->count.p, m * 4, hipMemcpyDeviceToHost));
    HIP_OK(c, hipMemcpy(parent.data(), c->parent.p, m * 4, hipMemcpyDeviceToHost));
    HIP_OK(c, hipMemcpy(leaf.data(), c->leaf.p, m, hipMemcpyDeviceToHost));
    uint64_t g = 0;
    for (uint32_t u = 0; u < m; ++u) {
      const uint32_t pc = u == 0 ? std::numeric_limits<uint32_t>::max() : count[parent[u]];
      if ((leaf[u] || count[u] <= c->group_size) && pc > c->group_size) ++g;
    }
    *n_groups = g;
  }
  if (dom4) {
    double4 r;
    HIP_OK(c, hipMemcpy(&r, c->ctr.p, sizeof(double4), hipMemcpyDeviceToHost));
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
  if (c->n_sys != 1) {
    set_err(c, "tree export covers single-system trees");
    return BH_ERR_INVALID;
  }
  HIP_OK(c, hipSetDevice(c->device));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  const uint32_t n = c->n, m = c->n_nodes;
  if (keys) HIP_OK(c, hipMemcpy(keys, c->keys_sorted, uint64_t(n) * 8, hipMemcpyDeviceToHost));
  if (perm) HIP_OK(c, hipMemcpy(perm, c->perm_sorted, uint64_t(n) * 4, hipMemcpyDeviceToHost));
  if (sorted) HIP_OK(c, hipMemcpy(sorted, c->sorted, uint64_t(n) * sizeof(bh_particle), hipMemcpyDeviceToHost));
  std::vector<uint32_t> begin(m), count(m), first_child(m), parent(m);
  std::vector<uint8_t> depth(m), leaf(m), nchild(m);
  std::vector<double4> ctr(m);
  std::vector<double> mass(m), com(3ull * m), ib(6ull * m);
  HIP_OK(c, hipMemcpy(begin.data(), c->begin.p, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(count.data(), c->count.p, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(first_child.data(), c->first_child.p, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(parent.data(), c->parent.p, m * 4ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(depth.data(), c->depth.p, m, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(leaf.data(), c->leaf.p, m, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(nchild.data(), c->nchild.p, m, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(ctr.data(), c->ctr.p, m * sizeof(double4), hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(mass.data(), c->mass.p, m * 8ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(com.data(), c->com.p, m * 24ull, hipMemcpyDeviceToHost));
  HIP_OK(c, hipMemcpy(ib.data(), c->iab.p, m * 48ull, hipMemcpyDeviceToHost));
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
  if (c->n_sys != 1) {
    set_err(c, "bh_accel covers single-system trees");
    return BH_ERR_INVALID;
  }
  HIP_OK(c, hipSetDevice(c->device));
  const uint32_t n = c->n;
  DevBuf<double>& da = c->dacc;
  TRY(ensure(c, da, 3ull * n, "accelerations"));
  DevBuf<uint32_t>& di = c->didx;
  TRY(ensure(c, di, 2ull * n, "interaction counts"));
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
  c->any_unsoftened = static_cast<float>(fp->eps2) == 0.0f;
  HIP_OK(c, hipMemcpyAsync(c->tables.p, c->host_tables, sizeof(SysDev), hipMemcpyHostToDevice, c->stream));
  TRY(walk(c, nullptr, da.p, di.p, di.p + n));
  if (acc) HIP_OK(c, hipMemcpyAsync(acc, da.p, 24ull * n, hipMemcpyDeviceToHost, c->stream));
  if (n_pp) HIP_OK(c, hipMemcpyAsync(n_pp, di.p, 4ull * n, hipMemcpyDeviceToHost, c->stream));
  if (n_pc) HIP_OK(c, hipMemcpyAsync(n_pc, di.p + n, 4ull * n, hipMemcpyDeviceToHost, c->stream));
  unsigned long long t2[2] = {0, 0};
  HIP_OK(c, hipMemcpyAsync(t2, c->totals.p, sizeof t2, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  if (totals) {
    totals[0] = t2[0];
    totals[1] = t2[1];
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
  unsigned long long t2[2] = {0, 0};
  HIP_OK(c, hipMemcpyAsync(t2, c->totals.p, sizeof t2, hipMemcpyDeviceToHost, c->stream));
  HIP_OK(c, hipStreamSynchronize(c->stream));
  st->sort_ms = elapsed(c->ev[0], c->ev[1]);
  st->tree_ms = elapsed(c->ev[1], c->ev[2]);
  st->targets_ms = elapsed(c->ev[2], c->ev[3]);
  st->walk_ms = elapsed(c->ev[4], c->ev[5]);
  st->n_pp = t2[0];
  st->n_pc = t2[1];
  st->n_nodes = c->n_nodes;
  st->n_leaves = c->n_leaves;
  st->n_targets = c->host_misc[40];
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
  double sum = 0.0;
  for (double v : part) sum += v;
  *potential = -0.5 * G * sum;
  return BH_OK;
}

}  // extern "C"
