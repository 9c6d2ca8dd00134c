// Internal interfaces between the host orchestration (bh_api.cpp) and the
// CDNA4 kernels (bh_sort.hip, bh_tree.hip, bh_walk.hip). Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bh_gpu.h"

namespace bhk {

constexpr int kWave = 64;
constexpr int kDepthCap = 21;           // flat_tree.h:102 (21 octal digits)
constexpr int kLevels = kDepthCap + 1;  // depths 0..21
constexpr uint32_t kMaxLeafCap = 1024;  // halo of the tree-build tile
constexpr uint32_t kSortTile = 4096;    // keys per onesweep workgroup
constexpr uint32_t kTreeTile = 4096;    // particles per tree-build workgroup
constexpr int kTreeCols = 24;           // per tile: 22 level counts, leaf count, pad

// Flag/value word of the decoupled look-back (30-bit counts).
constexpr uint32_t kLbAgg = 1u << 30;
constexpr uint32_t kLbPre = 2u << 30;
constexpr uint32_t kLbVal = (1u << 30) - 1;

// One system of a batch: its particles are [off, off + n) of the batch
// arrays, both in input and in sorted order (systems stay contiguous).
struct SysDev {
  double theta2;
  double eps2;
  double G;
  double kick_dt;
  double drift_dt;
  uint32_t off;
  uint32_t n;
  uint32_t max_leaf;
  uint32_t tsm;      // min(target_size, group_size)
  int32_t use_quad;
  uint32_t tile;     // walk tile: particles a walk unit holds (kMinTile..kGroupCap)
};
static_assert(sizeof(SysDev) == 64, "SysDev layout");

constexpr uint32_t kChunk = 4096;    // particles per bounds block
constexpr uint32_t kMaxSystems = 65535;
constexpr uint64_t kMaxN = (1ull << 30) - 1;  // particles per step (all systems)
constexpr uint32_t kGroupCap = 128;  // largest walk tile: a walk unit's particles fill at most two slots per lane
constexpr uint32_t kMinTile = 16;    // smallest walk tile
constexpr uint32_t kDefer = 512;     // deferred nodes a group may hand its targets
constexpr uint32_t kShallowFrames = 12;  // walk stack of trees with at most this many levels
constexpr uint32_t kWalkSlots = 4096;    // walk workgroups at most: 256 CUs x 4 SIMDs x 4 waves

// The step's flag and count words (`misc`), cleared when a step starts and
// read back once when it ends. Kernels after the tree's node count test the
// abort word and return at once when it is set, so a step whose tree did not
// fit the planned arrays or levels leaves its output untouched and is run
// again with larger ones.
constexpr int kMiscNonfinite = 0;   // some position, velocity or mass is NaN or infinite
constexpr int kMiscAbort = 1;       // kAbortNodes | kAbortDepth
constexpr int kMiscIncomplete = 2;  // the walk met an interior node it could not open
constexpr int kMiscMaxDepth = 3;    // deepest level of the tree
constexpr int kMiscLevels = 8;      // [8, 8 + kLevels): nodes per level; [8 + kLevels]: leaves
constexpr int kMiscTargets = 40;    // walk targets
constexpr int kMiscGroups = 41;     // walk units
constexpr int kMiscQueue = 48;      // [48, 56): walk-unit queue heads, one per XCD
constexpr int kMiscWords = 64;
constexpr uint32_t kAbortNodes = 1;  // more nodes or leaves than the node arrays hold
constexpr uint32_t kAbortDepth = 2;  // deeper than the levels the upward pass launched

// The device a context was created on.
int ctx_device(const bh_ctx* ctx);

// What the walk reads of a node, in one 64-byte record: the MAC data (FP64
// com and squared edge), the link (leaf: first particle and count | 1 << 31;
// interior: first child and child count) and the FP32 mass and traceless
// quadrupole (Qzz = -(Qxx + Qyy)).
struct alignas(64) NodeRec {
  double cx, cy, cz, edge2;
  uint32_t first, count;
  float m, qxx, qyy, qxy, qxz, qyz;
};
static_assert(sizeof(NodeRec) == 64, "NodeRec layout");

// Device view of one tree (level-order BFS numbering).
struct TreeDev {
  uint32_t* begin;       // [M] first particle (Morton position)
  uint32_t* count;       // [M] particles
  uint8_t* depth;        // [M]
  uint8_t* leaf;         // [M]
  uint32_t* cs;          // [M] child start: level-order index of the first child slot
  uint32_t* first_child; // [M] (0 for leaves)
  uint8_t* nchild;       // [M]
  uint32_t* parent;      // [M] (UINT32_MAX for the root)
  uint32_t* leaf_idx;    // [M] index into the leaf list (leaves only)
  double4* ctr;          // [M] geometric centre xyz, half width
  double* mass;          // [M]
  double* com;           // [3M]
  double* iab;           // [6M] xx,yy,zz,xy,xz,yz about the com
  NodeRec* rec;          // [M] the walk's node record
  double* q64;           // [8M] mass, Qxx, Qyy, Qzz, Qxy, Qxz, Qyz, 0
  double* nbox;          // [6M] FP64 box of the node's particles: lo x, y, z, hi x, y, z
  uint32_t* leaf_node;   // [leaves] node of each leaf, Morton order
  uint32_t* leaf_first;  // [leaves] first particle of each leaf
  uint32_t* level_off;   // [kLevels + 2] level start offsets, then total, total
  uint32_t* misc;        // [kMiscWords] the step's flags and counts
  uint32_t cap;          // node (and leaf) capacity of the arrays above
};

struct WalkArgs {
  const NodeRec* rec;
  const double* q64;
  const double4* ctr;
  const uint32_t* begin;
  const uint32_t* count;
  const double4* posm;     // x, y, z, m in Morton order
  const uint32_t* groups;  // walk units (first node of each), Morton order
  const uint32_t* unit_n;  // [n_nodes] particles | 1 << 31 of a packed unit at its first node, else 0 (nullable)
  const uint32_t* tmark;   // [n] target node + 1 at each target's first particle, else 0
  const uint16_t* sid;     // system of each particle
  const SysDev* sys;       // system table; the root of system s is node s
  const double* nbox;      // [6 n_nodes] FP64 particle box of every node (lo xyz, hi xyz)
  double* gacc;            // [3n] scratch: a group's shared-list sums per particle
  uint32_t* defer;         // [slots * kDefer] scratch: deferred nodes, per walk workgroup
  uint32_t* deep_frames;   // [slots * (kLevels - kShallowFrames) * 64] stack frames below the LDS ones
  uint32_t* misc;          // the step's flags; unit count and queue heads
  uint32_t n_nodes;        // index clamps (the node capacity): a node or particle
  uint32_t n_parts;        // index never leaves its array, whatever the tree holds
  uint32_t stack_limit;    // frames the walk may use (kLevels; lower only to test the incomplete-walk report)
  bh_particle* parts;      // Morton-order particles to kick/drift (nullable)
  const bh_particle* src;  // with parts: the unsorted input, read at perm (else parts is read in place)
  const uint32_t* perm;    // [n] Morton position -> input index (with src)
  double* carry;           // with parts: [4 n] the new x, y, z and m (blocks of n) for the next step's build (nullable)
  uint32_t* carry_bad;     // with carry: set when a new position, velocity or mass is not finite
  double* acc;             // [3n] G * acceleration (nullable)
  uint32_t* n_pp;          // [n] (nullable)
  uint32_t* n_pc;          // [n] (nullable)
  unsigned long long* totals;  // [2]
};

// ---- bh_sort.hip
// Up to kMax device ranges of 32-bit words to zero (unused: n = 0).
struct ZeroRanges {
  static constexpr int kMax = 4;
  uint32_t* p[kMax] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t n[kMax] = {0, 0, 0, 0};
};
hipError_t launch_zero(const ZeroRanges& r, hipStream_t s);

// chunks[c] = {system, begin, end, 0}: blocks of <= kChunk particles of one
// system; chunk_first[s] = first chunk of system s (chunk_first[S] = total).
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);
hipError_t launch_fill_sid(const uint4* chunks, uint32_t n_chunks, uint16_t* sid, hipStream_t s);
hipError_t launch_bounds(const bh_particle* p, const uint4* chunks, uint32_t n_chunks, const uint32_t* chunk_first,
                         uint32_t n_sys, double* partial, uint32_t* nonfinite, double* lohi, hipStream_t s);
hipError_t launch_keys_hist(const bh_particle* p, uint32_t n, const uint16_t* sid, const double* lohi, uint64_t* keys,
                            uint32_t* hist, hipStream_t s);
// The same from the compact copy a resident step's walk wrote for its
// output (the next step's input): blocks of n x, y, z and m, of which these
// read x, y, z (24 of the records' 64 bytes). carried_bad: that walk's
// non-finite flag, added to nonfinite.
hipError_t launch_bounds(const double* pos, uint32_t n, const uint32_t* carried_bad, const uint4* chunks,
                         uint32_t n_chunks, const uint32_t* chunk_first, uint32_t n_sys, double* partial,
                         uint32_t* nonfinite, double* lohi, hipStream_t s);
hipError_t launch_keys_hist(const double* pos, uint32_t n, const uint16_t* sid, const double* lohi, uint64_t* keys,
                            uint32_t* hist, hipStream_t s);
// 8 passes over the 63-bit keys, then sid_passes (0-2) over the system id
// taken from sid[index]: a stable sort of each system's (key, index) pairs
// in place. hist holds 10 x 256 digit counts (passes 8-9: the systems).
hipError_t launch_radix_sort(uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b, uint32_t n,
                             int sid_passes, const uint16_t* sid, uint32_t* hist, uint32_t* digit_base,
                             uint32_t* lookback, uint32_t* tile_counters, uint64_t** keys_out, uint32_t** vals_out,
                             hipStream_t s);
// The FP64 walk copy (x, y, z, m) in Morton order; with out, also the
// Morton-ordered particles (a step's walk writes those itself).
hipError_t launch_gather(const bh_particle* in, bh_particle* out, const uint32_t* perm, double4* posm, uint32_t n,
                         hipStream_t s);
// The same walk copy from a carried compact copy (blocks of n x, y, z and
// m) in the order perm indexes.
hipError_t launch_gather(const double* pos, const uint32_t* perm, double4* posm, uint32_t n, hipStream_t s);
uint64_t radix_lookback_words(uint32_t n);  // 10 passes of look-back words + 10 tile counters

// ---- bh_tree.hip
hipError_t launch_tree_count(const uint64_t* keys, uint32_t n, const uint16_t* sid, const SysDev* sys,
                             uint32_t max_leaf_max, uint8_t* L, uint8_t* H, uint32_t* tile_counts, hipStream_t s);
hipError_t launch_tree_scan(const uint32_t* tile_counts, uint32_t tiles, uint32_t* tile_prefix, uint32_t* totals,
                            hipStream_t s);
// Numbers the nodes and writes their topology; when the tree has more nodes
// or leaves than t.cap it writes nothing and sets kAbortNodes, when it is
// deeper than depth_plan it sets kAbortDepth.
hipError_t launch_tree_emit(const uint8_t* L, const uint8_t* H, uint32_t n, const uint32_t* tile_prefix, TreeDev t,
                            uint32_t depth_plan, hipStream_t s);
// Links, geometry and the upward pass: every leaf in one launch, then the
// interior nodes of levels depth_plan - 1..0 (one launch per level, each
// grid-striding over the level's range read on the device; n_sys systems
// bound a level's node count by n_sys * 8^d).
hipError_t launch_tree_finish(const uint64_t* keys, const double4* posm, uint32_t n, const uint16_t* sid, uint32_t n_sys,
                              const double* lohi, uint32_t depth_plan, TreeDev t, hipStream_t s);
// Marks (node + 1 at its first particle) and lists the topmost nodes with
// leaf || count <= thr: thr = the system's tsm (walk targets) or, with
// groups set, max(tsm, tile) (walk groups; tile = the system's walk tile).
// With groups and unit_n, runs of sibling single-target groups are packed
// into walk units of at most tile particles (unit_n per node: particles |
// 1 << 31 at a unit's first node, else 0).
hipError_t launch_targets(TreeDev t, uint32_t n, const uint16_t* sid, const SysDev* sys, bool groups, uint32_t* mark,
                          uint32_t* tile_counts, uint32_t* tile_prefix, uint32_t* targets, uint32_t* n_targets,
                          uint32_t* unit_n, hipStream_t s);

// ---- bh_walk.hip
// slots workgroups (a multiple of 8, at most kWalkSlots) take the units from
// per-XCD queues; deep selects the variant whose stack also uses deep_frames
// (trees of more than kShallowFrames levels).
hipError_t launch_walk(const WalkArgs& a, uint32_t slots, bool deep, int fp64, bool any_unsoftened, hipStream_t s);
hipError_t launch_direct(const double4* src, uint32_t n, const uint32_t* tidx, uint32_t m, double eps2, double G,
                         double* partial, uint32_t splits, double* acc, hipStream_t s);
uint32_t direct_splits(uint32_t n, uint32_t m);
// partial[splits * ceil(n/256)]: per (slice, block) sums of m_i * m_j / r_ij over
// unordered pairs {i, j} (each counted once, from the i of the block)
hipError_t launch_potential(const double4* src, uint32_t n, double* partial, uint32_t splits, hipStream_t s);

}  // namespace bhk
