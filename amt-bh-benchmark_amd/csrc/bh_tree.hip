// Octree topology, multipoles and walk targets from the Morton-sorted keys.
//
// Reference semantics (whole key range owned, one rank):
//   LeafRule                 flat_tree.h:144-156   leaf iff count <= max_leaf || depth >= 21
//   split_octants/split_node flat_tree.h:167-225   children = non-empty octants, ascending, consecutive
//   child_center             flat_tree.h:182-185   c +- hw/2 per octant bit, hw/2
//   finalize_node            flat_tree.h:257-274   raw moments about the cell centre, parallel-axis shift
//   group roots              flat_tree.h:190-192
//   targets                  grouped_ttg.h:253     topmost nodes with leaf || count <= target_size
//
// Construction without a level loop. For sorted keys K and particle i:
//   L_i = 1 + (octal digits K[i-1] and K[i] share)      (L_0 = 0, 22 if equal)
//         i is the first particle of its cell at every depth d >= L_i;
//   P_u = digits shared by K[u] and K[u+max_leaf]       (-1 if u+max_leaf >= n)
//         the depth-d cell of u holds > max_leaf particles iff d <= P_u;
//   W_i = max P_u over u in [i-max_leaf, i]
//         the depth-d cell containing i holds > max_leaf particles iff d <= W_i
//         (a cell of >= max_leaf+1 particles containing i contains a window
//          [u, u+max_leaf] with u <= i <= u+max_leaf, and conversely);
//   H_i = min(W_i, 20) + 1 (0 if n <= max_leaf) = depth of the leaf holding i,
//         since a cell is a node iff its parent is split (> max_leaf, depth < 21).
// So particle i begins exactly the nodes (d, i) for d in [L_i, H_i]. Nodes are
// numbered level by level (depth, then first particle): the level-order BFS
// the reference's build_subtree produces for a tree of <= 4096 particles.
// A node's children are the level d+1 nodes that begin inside it; with
// cs(n) = number of level-(d+1) begins before begin(n), n's children are
// [cs(n), cs(next node of level d)).
//
// A batch of independent systems is built in the same launches: each
// system's particles are contiguous, its first particle begins its root
// (level-0 node s for system s), and the windows stop at its boundaries, so
// every system gets exactly the tree it would get alone.
//
// Compiled with -ffp-contract=off: centres and multipoles round exactly as
// the reference's FP64 code does.
#include "bh_internal.h"

namespace bhk {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;

__device__ inline uint64_t lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// OR of v over the wavefront (DPP inclusive scan, lane 63), wave-uniform.
__device__ inline uint32_t wave_or(uint32_t v) {
  int x = static_cast<int>(v);
  x |= __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x |= __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x |= __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x |= __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x |= __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x |= __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return static_cast<uint32_t>(__builtin_amdgcn_readlane(x, 63));
}

// Levels lj..hj as bits (0 when the particle begins no node).
__device__ inline uint32_t level_bits(int lj, int hj) {
  return lj <= hj ? (2u << hj) - (1u << lj) : 0u;
}

// Number of leading octal digits (of 21) two 63-bit keys share.
__device__ inline int common_digits(uint64_t a, uint64_t b) {
  const uint64_t x = a ^ b;
  return x == 0 ? kDepthCap : (__clzll(static_cast<long long>(x)) - 1) / 3;
}

__device__ inline void domain_from_lohi(const double* lohi, double c[3], double* hw) {
  c[0] = 0.5 * (lohi[0] + lohi[3]);
  c[1] = 0.5 * (lohi[1] + lohi[4]);
  c[2] = 0.5 * (lohi[2] + lohi[5]);
  const double ex = lohi[3] - lohi[0], ey = lohi[4] - lohi[1], ez = lohi[5] - lohi[2];
  double m = ex;
  if (m < ey) m = ey;
  if (m < ez) m = ez;
  double h = 0.5 * m;
  if (!(h > 0.0)) h = 1.0;
  h *= 1.0 + 1e-9;
  *hw = h;
}

// L, H per particle and per-tile counts of level begins and leaf begins.
// With several systems the windows stop at system boundaries: a system's
// first particle begins its root (L = 0) and P, W only look at particles of
// the same system, with that system's max_leaf.
__global__ __launch_bounds__(kBlock) void k_tree_count(const uint64_t* __restrict__ keys, uint32_t n,
                                                       const uint16_t* __restrict__ sid, const SysDev* __restrict__ sys,
                                                       uint32_t mlx, uint8_t* __restrict__ L, uint8_t* __restrict__ H,
                                                       uint32_t* __restrict__ tile_counts) {
  // dynamic LDS sized by the batch's largest max_leaf (mlx): keys of
  // [base - mlx - 1, base + tile + mlx), then P of [base - mlx, base + tile)
  extern __shared__ __align__(8) unsigned char s_dyn[];
  uint64_t* s_k = reinterpret_cast<uint64_t*>(s_dyn);
  int8_t* s_p = reinterpret_cast<int8_t*>(s_k + kTreeTile + 2 * mlx + 1);
  __shared__ uint32_t s_cnt[kTreeCols];
  __shared__ uint32_t s_one;  // 1 + system id when the tile and its halo hold one system, else 0
  const uint32_t tile = blockIdx.x, base = tile * kTreeTile;
  const int64_t g0 = static_cast<int64_t>(base) - mlx - 1;  // global index of s_k[0]
  constexpr int kPer = static_cast<int>(kTreeTile) / kBlock;
  {
    // the tile's keys: every load in flight before the first LDS write
    uint64_t kv[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const uint32_t i = base + threadIdx.x + j * kBlock;
      kv[j] = i < n ? keys[i] : 0ull;
    }
    // the halos [g0, base) and [base + tile, base + tile + mlx)
    for (uint32_t h = threadIdx.x; h < 2 * mlx + 1; h += kBlock) {
      const uint32_t e = h <= mlx ? h : h + kTreeTile;
      const int64_t g = g0 + e;
      s_k[e] = (g >= 0 && g < n) ? keys[g] : 0ull;
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) s_k[mlx + 1 + threadIdx.x + j * kBlock] = kv[j];
  }
  if (threadIdx.x < kTreeCols) s_cnt[threadIdx.x] = 0;
  if (threadIdx.x == 0) {
    const int64_t lo = g0 + 1 > 0 ? g0 + 1 : 0;
    const int64_t hi = static_cast<int64_t>(base) + kTreeTile + mlx < n ? static_cast<int64_t>(base) + kTreeTile + mlx : n - 1;
    s_one = sid[lo] == sid[hi] ? 1u + sid[lo] : 0u;
  }
  __syncthreads();
  const uint32_t one = s_one;
  // the system's off, n and max_leaf: loaded once when the tile holds one
  // system, per particle otherwise
  uint32_t o_off = 0, o_n = 0, o_ml = 0;
  if (one) {
    o_off = sys[one - 1u].off;
    o_n = sys[one - 1u].n;
    o_ml = sys[one - 1u].max_leaf;
  }
  // P_u for u in [base - mlx, base + tile)
  for (uint32_t e = threadIdx.x; e < kTreeTile + mlx; e += kBlock) {
    const int64_t u = static_cast<int64_t>(base) - mlx + e;
    int8_t pv = -1;
    if (u >= 0 && u < n) {
      uint32_t qoff = o_off, qn = o_n, qml = o_ml;
      if (!one) {
        const SysDev& q = sys[sid[u]];
        qoff = q.off;
        qn = q.n;
        qml = q.max_leaf;
      }
      if (u + qml < qoff + qn) pv = static_cast<int8_t>(common_digits(s_k[u - g0], s_k[u + qml - g0]));
    }
    s_p[e] = pv;
  }
  // L_i while the keys are still here (their space holds the window table
  // below)
  int li_r[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t e = threadIdx.x + j * kBlock, i = base + e;
    li_r[j] = 0;
    if (i < n) {
      const uint32_t qo = one ? o_off : sys[sid[i]].off;
      li_r[j] = (i == qo) ? 0 : common_digits(s_k[i - 1 - g0], s_k[i - g0]) + 1;
    }
  }
  __syncthreads();
  // W_i = max P over the window [i - span, i]. With one system in the tile
  // and its halo, full windows (span = max_leaf, w = max_leaf + 1 entries)
  // come from a doubling table: level k holds max P over [e, e + 2^k), and
  // with 2^K <= w < 2^(K+1) a window is the max of two level-K entries.
  // Levels are built in two byte buffers over the keys' space; level 0 is
  // s_p itself, kept for the windows cut short at a system's first
  // particles and for tiles that hold several systems (direct scan).
  const uint32_t np = kTreeTile + mlx;
  int8_t* const tb0 = reinterpret_cast<int8_t*>(s_k);
  int8_t* const tb1 = tb0 + ((np + 15u) & ~15u);
  const int8_t* top = s_p;
  int K = 0;
  if (one) {
    const uint32_t w = o_ml + 1u;
    K = 31 - __clz(static_cast<int>(w));
    for (int k = 1; k <= K; ++k) {
      int8_t* dst = (k & 1) ? tb0 : tb1;
      const uint32_t h = 1u << (k - 1);
      for (uint32_t e = threadIdx.x; e < np; e += kBlock) {
        const int a0 = top[e];
        const int a1 = e + h < np ? top[e + h] : -1;
        dst[e] = static_cast<int8_t>(a0 > a1 ? a0 : a1);
      }
      __syncthreads();
      top = dst;
    }
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const uint32_t e = threadIdx.x + j * kBlock, i = base + e;
    if (i >= n) break;
    uint32_t qoff = o_off, qn = o_n, ml = o_ml;
    if (!one) {
      const SysDev& q = sys[sid[i]];
      qoff = q.off;
      qn = q.n;
      ml = q.max_leaf;
    }
    const int li = li_r[j];
    const uint32_t span = (i - qoff < ml) ? i - qoff : ml;  // u = i - d, d <= span, stays in the system
    int wmax = -1;
    if (one && span == ml) {
      const int a0 = top[e + mlx - ml], a1 = top[e + mlx + 1u - (1u << K)];
      wmax = a0 > a1 ? a0 : a1;
    } else {
      for (uint32_t d = 0; d <= span; ++d) wmax = max(wmax, static_cast<int>(s_p[e + mlx - d]));
    }
    const int hi = (qn > ml) ? min(wmax, kDepthCap - 1) + 1 : 0;
    L[i] = static_cast<uint8_t>(li);
    H[i] = static_cast<uint8_t>(hi);
    for (int d = li; d <= hi; ++d) atomicAdd(&s_cnt[d], 1u);
    if (li <= hi) atomicAdd(&s_cnt[kLevels], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kTreeCols) tile_counts[uint64_t(tile) * kTreeCols + threadIdx.x] = s_cnt[threadIdx.x];
}

// Column-wise exclusive scan over tiles: one block per column.
__global__ __launch_bounds__(1024) void k_col_scan(const uint32_t* __restrict__ in, uint32_t tiles, uint32_t stride,
                                                   uint32_t* __restrict__ prefix, uint32_t* __restrict__ totals) {
  __shared__ uint32_t s_w[16];
  const uint32_t col = blockIdx.x;
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t carry = 0;
  for (uint32_t b = 0; b < tiles; b += 1024) {
    const uint32_t t = b + threadIdx.x;
    const uint32_t v = t < tiles ? in[uint64_t(t) * stride + col] : 0u;
    uint32_t x = v;
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t u = __shfl_up(x, o);
      if (lane >= o) x += u;
    }
    if (lane == 63) s_w[w] = x;
    __syncthreads();
    uint32_t off = 0, tot = 0;
    for (int k = 0; k < 16; ++k) {
      if (k < w) off += s_w[k];
      tot += s_w[k];
    }
    if (t < tiles) prefix[uint64_t(t) * stride + col] = carry + off + x - v;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) totals[col] = carry;
}

// Node numbering and topology. Nothing is written when the tree has more
// nodes or leaves than the arrays hold (kAbortNodes; the host grows them and
// runs the step again); a tree deeper than the upward launches planned is
// numbered but flagged (kAbortDepth), so the kernels after this one skip it.
__global__ __launch_bounds__(kBlock) void k_tree_emit(const uint8_t* __restrict__ L, const uint8_t* __restrict__ H, uint32_t n,
                                                      const uint32_t* __restrict__ tile_prefix, TreeDev t,
                                                      uint32_t depth_plan) {
  __shared__ uint32_t s_loff[kLevels + 2];
  __shared__ uint32_t s_wc[4][kTreeCols];
  __shared__ uint32_t s_over;
  const uint32_t tile = blockIdx.x, base = tile * kTreeTile;
  const int w = threadIdx.x >> 6, lane = __lane_id();
  if (threadIdx.x == 0) {
    const uint32_t* totals = t.misc + kMiscLevels;
    uint32_t off = 0, deepest = 0;
    for (int d = 0; d < kLevels; ++d) {
      s_loff[d] = off;
      off += totals[d];
      if (totals[d]) deepest = static_cast<uint32_t>(d);
    }
    s_loff[kLevels] = off;
    s_loff[kLevels + 1] = off;
    const bool over = off > t.cap || totals[kLevels] > t.cap;
    s_over = over ? 1u : 0u;
    if (tile == 0) {
      t.misc[kMiscMaxDepth] = deepest;
      const uint32_t abort = (over ? kAbortNodes : 0u) | (deepest > depth_plan ? kAbortDepth : 0u);
      if (abort) t.misc[kMiscAbort] = abort;
      if (!over)
        for (int d = 0; d < kLevels + 2; ++d) t.level_off[d] = s_loff[d];
    }
  }
  __syncthreads();
  if (s_over) return;
  // phase 1: per-wave counts of level begins and leaf begins
  uint32_t c[kLevels + 1];
#pragma unroll
  for (int d = 0; d <= kLevels; ++d) c[d] = 0;
  // an item's L and H, packed (255 = no particle); the next item's load is
  // issued before the current item's levels are counted
  auto load_lh = [&](int j) -> uint32_t {
    const uint32_t idx = base + w * (kTreeTile / 4) + j * 64 + lane;
    return idx < n ? static_cast<uint32_t>(L[idx]) | (static_cast<uint32_t>(H[idx]) << 8) : 255u;
  };
  uint32_t nlh = load_lh(0);
#pragma unroll 1
  for (int j = 0; j < kItems; ++j) {
    const uint32_t lh = nlh;
    if (j + 1 < kItems) nlh = load_lh(j + 1);
    const int lj = static_cast<int>(lh & 255u), hj = static_cast<int>(lh >> 8);
    // only the levels where some lane begins a node
    const uint32_t wm = wave_or(level_bits(lj, hj));
#pragma unroll
    for (int d = 0; d < kLevels; ++d)
      if ((wm >> d) & 1u) c[d] += __popcll(__ballot(lj <= d && d <= hj));
    c[kLevels] += __popcll(__ballot(lj <= hj));
  }
  if (lane == 0) {
#pragma unroll
    for (int d = 0; d <= kLevels; ++d) s_wc[w][d] = c[d];
  }
  __syncthreads();
  if (threadIdx.x <= kLevels) {
    uint32_t a = 0;
    for (int k = 0; k < 4; ++k) {
      const uint32_t v = s_wc[k][threadIdx.x];
      s_wc[k][threadIdx.x] = a;
      a += v;
    }
  }
  __syncthreads();
  // phase 2: global ranks
  uint32_t run[kLevels + 1];
#pragma unroll
  for (int d = 0; d < kLevels; ++d) run[d] = s_loff[d] + tile_prefix[uint64_t(tile) * kTreeCols + d] + s_wc[w][d];
  run[kLevels] = tile_prefix[uint64_t(tile) * kTreeCols + kLevels] + s_wc[w][kLevels];
  const uint32_t total_nodes = s_loff[kLevels];
  const uint64_t lt = lanemask_lt();
  nlh = load_lh(0);
#pragma unroll 1
  for (int j = 0; j < kItems; ++j) {
    const uint32_t idx = base + w * (kTreeTile / 4) + j * 64 + lane;
    const uint32_t lh = nlh;
    if (j + 1 < kItems) nlh = load_lh(j + 1);
    const int lj = static_cast<int>(lh & 255u), hj = static_cast<int>(lh >> 8);
    const uint64_t lm = __ballot(lj <= hj);
    const uint32_t lr = run[kLevels] + __popcll(lm & lt);
    run[kLevels] += __popcll(lm);
    // levels with begins in this wave, and the levels just below them (the
    // children's start ranks); elsewhere no lane begins a node and a rank is
    // the running count
    const uint32_t wm = wave_or(level_bits(lj, hj));
    const uint32_t need = wm | (wm << 1);
    uint32_t prev = 0;
#pragma unroll
    for (int d = 0; d <= kLevels; ++d) {
      uint32_t rank;
      if (d < kLevels) {
        rank = run[d];
        if ((need >> d) & 1u) {
          const uint64_t m = __ballot(lj <= d && d <= hj);
          rank += __popcll(m & lt);
          run[d] += __popcll(m);
        }
      } else {
        rank = total_nodes;
      }
      if (d >= 1 && lj <= d - 1 && d - 1 <= hj) {
        const uint32_t node = prev;
        t.begin[node] = idx;
        t.depth[node] = static_cast<uint8_t>(d - 1);
        t.cs[node] = rank;
        const bool leaf = (d - 1 == hj);
        t.leaf[node] = leaf ? 1 : 0;
        if (leaf) {
          t.leaf_node[lr] = node;
          t.leaf_first[lr] = idx;
          t.leaf_idx[node] = lr;
        }
      }
      prev = rank;
    }
  }
}

// Children, parents, leaf counts and cell geometry of every node.
__global__ __launch_bounds__(kBlock) void k_tree_links(const uint64_t* __restrict__ keys, uint32_t n,
                                                       const uint16_t* __restrict__ sid, const double* __restrict__ lohi,
                                                       TreeDev t) {
  if (t.misc[kMiscAbort]) return;
  const uint32_t n_nodes = t.level_off[kLevels], n_leaves = t.misc[kMiscLevels + kLevels];
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n_nodes; i += gridDim.x * kBlock) {
    const int d = t.depth[i];
    const uint32_t lend = t.level_off[d + 1];
    const uint32_t next = (i + 1 < lend) ? t.cs[i + 1] : t.level_off[d + 2];
    const uint32_t cs = t.cs[i];
    if (d == 0) t.parent[i] = 0xffffffffu;  // the root of a system
    double c0[3], hw0;
    domain_from_lohi(lohi + 6 * sid[t.begin[i]], c0, &hw0);
    if (t.leaf[i]) {
      t.first_child[i] = 0;
      t.nchild[i] = 0;
      const uint32_t lr = t.leaf_idx[i];
      const uint32_t end = (lr + 1 < n_leaves) ? t.leaf_first[lr + 1] : n;
      t.count[i] = end - t.leaf_first[lr];
    } else {
      const uint32_t k = next - cs;
      t.first_child[i] = cs;
      t.nchild[i] = static_cast<uint8_t>(k);
      for (uint32_t c = 0; c < k; ++c) t.parent[cs + c] = i;
    }
    // child_center applied digit by digit from the root (flat_tree.h:182-185)
    const uint64_t key = keys[t.begin[i]];
    double cx = c0[0], cy = c0[1], cz = c0[2], h = hw0;
    for (int e = 1; e <= d; ++e) {
      const unsigned oct = static_cast<unsigned>((key >> (63 - 3 * e)) & 7u);
      const double q = 0.5 * h;
      cx = cx + ((oct & 1u) ? q : -q);
      cy = cy + ((oct & 2u) ? q : -q);
      cz = cz + ((oct & 4u) ? q : -q);
      h = 0.5 * h;
    }
    t.ctr[i] = make_double4(cx, cy, cz, h);
  }
}

// finalize_raw_moment (multipole.h:76-87) of node i from its raw sums about
// the cell centre, and the node's outputs: FP64 box, mass, com, second
// moments, the FP64 quadrupole and the walk's node record.
__device__ inline void finalize_write(TreeDev t, uint32_t i, const double4& cc, double rm, const double* m1,
                                      const double* m2, const double* blo, const double* bhi, uint32_t cnt,
                                      bool leaf) {
  const double inv = 1.0 / rm;
  const double dl[3] = {m1[0] * inv, m1[1] * inv, m1[2] * inv};
  const double com[3] = {cc.x + dl[0], cc.y + dl[1], cc.z + dl[2]};
  double ib[6];
  ib[0] = m2[0] - rm * dl[0] * dl[0];
  ib[1] = m2[1] - rm * dl[1] * dl[1];
  ib[2] = m2[2] - rm * dl[2] * dl[2];
  ib[3] = m2[3] - rm * dl[0] * dl[1];
  ib[4] = m2[4] - rm * dl[0] * dl[2];
  ib[5] = m2[5] - rm * dl[1] * dl[2];
  for (int a = 0; a < 3; ++a) {
    t.nbox[6 * uint64_t(i) + a] = blo[a];
    t.nbox[6 * uint64_t(i) + 3 + a] = bhi[a];
  }
  t.mass[i] = rm;
  for (int a = 0; a < 3; ++a) t.com[3 * i + a] = com[a];
  for (int a = 0; a < 6; ++a) t.iab[6 * i + a] = ib[a];
  // walk records: MAC data, links, traceless quadrupole (multipole.h:185-188)
  const double edge = 2.0 * cc.w;
  const double tr = ib[0] + ib[1] + ib[2];
  const double q[6] = {3.0 * ib[0] - tr, 3.0 * ib[1] - tr, 3.0 * ib[2] - tr, 3.0 * ib[3], 3.0 * ib[4], 3.0 * ib[5]};
  double* q64 = &t.q64[8 * i];
  q64[0] = rm;
  for (int a = 0; a < 6; ++a) q64[1 + a] = q[a];
  q64[7] = 0.0;
  NodeRec r;
  r.cx = com[0];
  r.cy = com[1];
  r.cz = com[2];
  r.edge2 = edge * edge;
  r.first = leaf ? t.begin[i] : t.first_child[i];
  r.count = leaf ? (cnt | 0x80000000u) : static_cast<uint32_t>(t.nchild[i]);
  r.m = static_cast<float>(rm);
  r.qxx = static_cast<float>(q[0]);
  r.qyy = static_cast<float>(q[1]);
  r.qxy = static_cast<float>(q[3]);
  r.qxz = static_cast<float>(q[4]);
  r.qyz = static_cast<float>(q[5]);
  t.rec[i] = r;
}

// finalize_node (flat_tree.h:257-274) of every leaf, at any depth, in one
// launch. A block takes 256 consecutive leaves of the leaf list (particle
// order): their particles are one contiguous range, which the block loads
// into LDS in coalesced chunks of kUpChunk; each thread then adds the
// particles of its leaf that lie in the chunk, in order, so its sums run
// over the leaf's particles in the reference's order.
constexpr uint32_t kUpChunk = 1024;

__global__ __launch_bounds__(kBlock) void k_upward_leaves(const double4* __restrict__ posm, uint32_t n, TreeDev t) {
  // one array per coordinate: lanes read scattered particles, and 8-byte
  // elements spread those reads over the banks better than 32-byte ones
  __shared__ double s_x[kUpChunk], s_y[kUpChunk], s_z[kUpChunk], s_m[kUpChunk];
  if (t.misc[kMiscAbort]) return;
  const uint32_t n_leaves = t.misc[kMiscLevels + kLevels];
  const double inf = __builtin_huge_val();
  for (uint32_t l0 = blockIdx.x * kBlock; l0 < n_leaves; l0 += gridDim.x * kBlock) {
    const uint32_t lr = l0 + threadIdx.x;
    const bool mine = lr < n_leaves;
    const uint32_t l1 = l0 + kBlock < n_leaves ? l0 + kBlock : n_leaves;
    const uint32_t span_lo = t.leaf_first[l0];
    const uint32_t span_hi = l1 < n_leaves ? t.leaf_first[l1] : n;
    uint32_t i = 0, b = 0, e = 0;
    double4 cc = make_double4(0.0, 0.0, 0.0, 0.0);
    if (mine) {
      i = t.leaf_node[lr];
      b = t.leaf_first[lr];
      e = b + t.count[i];
      cc = t.ctr[i];
    }
    double rm = 0.0, m1[3] = {0.0, 0.0, 0.0}, m2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // the FP64 box of the node's particles (multipole.h:115-128; exact)
    double blo[3] = {inf, inf, inf}, bhi[3] = {-inf, -inf, -inf};
    for (uint32_t c0 = span_lo; c0 < span_hi; c0 += kUpChunk) {
      const uint32_t cn = span_hi - c0 < kUpChunk ? span_hi - c0 : kUpChunk;
      __syncthreads();
      {
        // every load of the chunk in flight before the first LDS write
        constexpr int kR = kUpChunk / kBlock;
        double4 q[kR];
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const uint32_t k = threadIdx.x + r * kBlock;
          q[r] = posm[c0 + (k < cn ? k : cn - 1)];
        }
#pragma unroll
        for (int r = 0; r < kR; ++r) {
          const uint32_t k = threadIdx.x + r * kBlock;
          if (k < cn) {
            s_x[k] = q[r].x;
            s_y[k] = q[r].y;
            s_z[k] = q[r].z;
            s_m[k] = q[r].w;
          }
        }
      }
      __syncthreads();
      const uint32_t lo = b > c0 ? b : c0;
      const uint32_t hi = e < c0 + cn ? e : c0 + cn;
      for (uint32_t pidx = lo; pidx < hi; ++pidx) {
        const uint32_t o = pidx - c0;
        const double x = s_x[o], y = s_y[o], z = s_z[o], mass = s_m[o];
        blo[0] = fmin(blo[0], x);
        blo[1] = fmin(blo[1], y);
        blo[2] = fmin(blo[2], z);
        bhi[0] = fmax(bhi[0], x);
        bhi[1] = fmax(bhi[1], y);
        bhi[2] = fmax(bhi[2], z);
        const double d[3] = {x - cc.x, y - cc.y, z - cc.z};
        rm += mass;
        for (int a = 0; a < 3; ++a) m1[a] += d[a] * mass;
        m2[0] += mass * d[0] * d[0];
        m2[1] += mass * d[1] * d[1];
        m2[2] += mass * d[2] * d[2];
        m2[3] += mass * d[0] * d[1];
        m2[4] += mass * d[0] * d[2];
        m2[5] += mass * d[1] * d[2];
      }
    }
    if (mine) finalize_write(t, i, cc, rm, m1, m2, blo, bhi, e - b, true);
  }
}

// finalize_node for the interior nodes of one level, children already
// final. A team of 8 lanes takes a node: lane j loads child j (at most 8)
// and forms its terms as the reference does (mass, m d, the child's second
// moment + m d_a d_b, its box); the team then adds the terms in child
// order through shuffles, so the sums are the reference's to the bit, and
// every lane holds them; lane 0 writes the node. (Loading the children one
// after another from one thread left each level waiting on ~8 dependent
// memory round trips.)
constexpr int kKids = 8;

__global__ __launch_bounds__(kBlock) void k_upward(int level, TreeDev t) {
  if (t.misc[kMiscAbort]) return;
  const uint32_t lo = t.level_off[level], hi = t.level_off[level + 1];
  const double inf = __builtin_huge_val();
  const int j = static_cast<int>(threadIdx.x) & (kKids - 1);
  const uint32_t teams = gridDim.x * (kBlock / kKids);
  for (uint32_t i0 = lo + blockIdx.x * (kBlock / kKids); i0 < hi; i0 += teams) {
    const uint32_t i = i0 + threadIdx.x / kKids;
    const bool node_ok = i < hi && !t.leaf[i < hi ? i : lo];
    const uint32_t fc = node_ok ? t.first_child[i] : 0u;
    const uint32_t k = node_ok ? t.nchild[i] : 0u;
    const double4 cc = node_ok ? t.ctr[i] : make_double4(0.0, 0.0, 0.0, 0.0);
    // this lane's child and its terms
    const bool has = static_cast<uint32_t>(j) < k;
    const uint32_t c = fc + static_cast<uint32_t>(j);
    double tm = 0.0, t1[3] = {0.0, 0.0, 0.0}, t2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double clo[3] = {inf, inf, inf}, chi[3] = {-inf, -inf, -inf};
    uint32_t ccount = 0;
    if (has) {
      const double mass = t.mass[c];
      const double d[3] = {t.com[3 * c + 0] - cc.x, t.com[3 * c + 1] - cc.y, t.com[3 * c + 2] - cc.z};
      tm = mass;
      for (int a = 0; a < 3; ++a) t1[a] = d[a] * mass;
      const double* ib = &t.iab[6 * c];
      t2[0] = ib[0] + mass * d[0] * d[0];
      t2[1] = ib[1] + mass * d[1] * d[1];
      t2[2] = ib[2] + mass * d[2] * d[2];
      t2[3] = ib[3] + mass * d[0] * d[1];
      t2[4] = ib[4] + mass * d[0] * d[2];
      t2[5] = ib[5] + mass * d[1] * d[2];
      ccount = t.count[c];
      const double* cb = &t.nbox[6 * uint64_t(c)];
      for (int a = 0; a < 3; ++a) {
        clo[a] = cb[a];
        chi[a] = cb[3 + a];
      }
    }
    // sums in child order; the team's lanes share k, so every shuffle reads
    // a lane of the team that is in the loop
    double rm = 0.0, m1[3] = {0.0, 0.0, 0.0}, m2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double blo[3] = {inf, inf, inf}, bhi[3] = {-inf, -inf, -inf};
    uint32_t cnt = 0;
    for (uint32_t q = 0; q < k; ++q) {
      rm += __shfl(tm, static_cast<int>(q), kKids);
      for (int a = 0; a < 3; ++a) m1[a] += __shfl(t1[a], static_cast<int>(q), kKids);
      for (int a = 0; a < 6; ++a) m2[a] += __shfl(t2[a], static_cast<int>(q), kKids);
      for (int a = 0; a < 3; ++a) {
        blo[a] = fmin(blo[a], __shfl(clo[a], static_cast<int>(q), kKids));
        bhi[a] = fmax(bhi[a], __shfl(chi[a], static_cast<int>(q), kKids));
      }
      cnt += static_cast<uint32_t>(__shfl(static_cast<int>(ccount), static_cast<int>(q), kKids));
    }
    if (node_ok && j == 0) {
      t.count[i] = cnt;
      finalize_write(t, i, cc, rm, m1, m2, blo, bhi, cnt, false);
    }
  }
}

// Marks each walk target at its first particle: topmost nodes with
// leaf || count <= tsm (tsm = min(target_size, group_size), which makes the
// group-root / target nesting of grouped_ttg.h:251-288 a single predicate).
// With `groups`, the walk groups instead: topmost nodes with
// leaf || count <= max(tsm, tile).
__global__ __launch_bounds__(kBlock) void k_target_mark(TreeDev t, const uint16_t* __restrict__ sid,
                                                        const SysDev* __restrict__ sys, bool groups,
                                                        uint32_t* __restrict__ mark) {
  if (t.misc[kMiscAbort]) return;
  const uint32_t n_nodes = t.level_off[kLevels];
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n_nodes; i += gridDim.x * kBlock) {
    const SysDev& q = sys[sid[t.begin[i]]];
    const uint32_t tsm = groups ? max(q.tsm, q.tile) : q.tsm;
    const bool pred = t.leaf[i] || t.count[i] <= tsm;
    if (!pred) continue;
    const bool ppred = (t.depth[i] != 0) && (t.count[t.parent[i]] <= tsm);
    if (!ppred) mark[t.begin[i]] = i + 1;
  }
}

// Walk units from the walk groups. Below a node with more than
// max(tsm, tile) particles (tile: the system's walk tile), a run of
// consecutive children that are single targets (leaf || count <= tsm) is
// packed greedily into units of at most tile particles; a unit of two or
// more walks as one group (one
// shared walk, then its targets), the others keep their own walks. The
// unit's first child keeps its mark and gets unit_n = particles | 1 << 31,
// the other members lose theirs; every other walk group gets unit_n = 0.
// Any box holding the targets' boxes gives the same lists (near2 and far2
// are monotone in the box), so a unit is exact whatever it groups.
__global__ __launch_bounds__(kBlock) void k_group_units(TreeDev t, const uint16_t* __restrict__ sid,
                                                        const SysDev* __restrict__ sys, uint32_t* __restrict__ mark,
                                                        uint32_t* __restrict__ unit_n) {
  if (t.misc[kMiscAbort]) return;
  const uint32_t n_nodes = t.level_off[kLevels];
  for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n_nodes; i += gridDim.x * kBlock) {
    const SysDev& q = sys[sid[t.begin[i]]];
    const uint32_t tsm = q.tsm, tile = q.tile;
    const uint32_t thr = max(tsm, tile);
    const bool is_group = t.leaf[i] || t.count[i] <= thr;
    if (t.depth[i] == 0 && is_group) unit_n[i] = 0u;  // a system that is one walk group
    if (is_group) continue;
    const uint32_t fc = t.first_child[i], k = t.nchild[i];
    uint32_t run = 0, run_len = 0, run_n = 0;
    for (uint32_t c = fc; c <= fc + k; ++c) {
      const bool in = c < fc + k;
      const uint32_t cn = in ? t.count[c] : 0u;
      const bool single = in && (t.leaf[c] || cn <= tsm) && cn <= tile;
      if (in && (t.leaf[c] || cn <= thr)) unit_n[c] = 0u;
      if (single && run_len > 0 && run_n + cn <= tile) {
        mark[t.begin[c]] = 0u;
        ++run_len;
        run_n += cn;
        continue;
      }
      if (run_len > 1) unit_n[run] = run_n | 0x80000000u;
      run_len = single ? 1u : 0u;
      run = c;
      run_n = cn;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_compact_count(const uint32_t* __restrict__ mark, uint32_t n,
                                                          uint32_t* __restrict__ tile_counts) {
  __shared__ uint32_t s;
  if (threadIdx.x == 0) s = 0;
  __syncthreads();
  const uint32_t base = blockIdx.x * kTreeTile;
  uint32_t c = 0;
  for (uint32_t e = threadIdx.x; e < kTreeTile; e += kBlock) {
    const uint32_t i = base + e;
    if (i < n && mark[i]) ++c;
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if (__lane_id() == 0) atomicAdd(&s, c);
  __syncthreads();
  if (threadIdx.x == 0) tile_counts[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_compact_write(const uint32_t* __restrict__ mark, uint32_t n,
                                                          const uint32_t* __restrict__ tile_prefix, uint32_t* __restrict__ out) {
  __shared__ uint32_t s_wc[4];
  const uint32_t base = blockIdx.x * kTreeTile;
  const int w = threadIdx.x >> 6, lane = __lane_id();
  uint32_t vals[kItems];
  uint32_t cnt = 0;
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t i = base + w * (kTreeTile / 4) + j * 64 + lane;
    vals[j] = i < n ? mark[i] : 0u;
    cnt += __popcll(__ballot(vals[j] != 0));
  }
  if (lane == 0) s_wc[w] = cnt;
  __syncthreads();
  uint32_t pos = tile_prefix[blockIdx.x];
  for (int k = 0; k < w; ++k) pos += s_wc[k];
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint64_t m = __ballot(vals[j] != 0);
    if (vals[j]) out[pos + __popcll(m & lt)] = vals[j] - 1;
    pos += __popcll(m);
  }
}

inline uint32_t grid_for(uint64_t n, uint32_t cap) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  return static_cast<uint32_t>(g < cap ? g : cap);
}

}  // namespace

hipError_t launch_tree_count(const uint64_t* keys, uint32_t n, const uint16_t* sid, const SysDev* sys,
                             uint32_t max_leaf_max, uint8_t* L, uint8_t* H, uint32_t* tile_counts, hipStream_t s) {
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  const size_t lds = (kTreeTile + 2 * size_t(max_leaf_max) + 1) * sizeof(uint64_t) + kTreeTile + max_leaf_max;
  hipLaunchKernelGGL(k_tree_count, dim3(tiles), dim3(kBlock), lds, s, keys, n, sid, sys, max_leaf_max, L, H, tile_counts);
  return hipGetLastError();
}

hipError_t launch_tree_scan(const uint32_t* tile_counts, uint32_t tiles, uint32_t* tile_prefix, uint32_t* totals,
                            hipStream_t s) {
  hipLaunchKernelGGL(k_col_scan, dim3(kLevels + 1), dim3(1024), 0, s, tile_counts, tiles, static_cast<uint32_t>(kTreeCols),
                     tile_prefix, totals);
  return hipGetLastError();
}

hipError_t launch_tree_emit(const uint8_t* L, const uint8_t* H, uint32_t n, const uint32_t* tile_prefix, TreeDev t,
                            uint32_t depth_plan, hipStream_t s) {
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  hipLaunchKernelGGL(k_tree_emit, dim3(tiles), dim3(kBlock), 0, s, L, H, n, tile_prefix, t, depth_plan);
  return hipGetLastError();
}

hipError_t launch_tree_finish(const uint64_t* keys, const double4* posm, uint32_t n, const uint16_t* sid, uint32_t n_sys,
                              const double* lohi, uint32_t depth_plan, TreeDev t, hipStream_t s) {
  // grid-stride loops over the node count read on the device
  hipLaunchKernelGGL(k_tree_links, dim3(grid_for(t.cap, 4096)), dim3(kBlock), 0, s, keys, n, sid, lohi, t);
  // every leaf, then the interior nodes bottom-up: one launch per level
  // below the planned depth (the deepest level holds leaves only), each
  // reading its level's range on the device and grid-striding over it.
  // Level d holds at most min(n_sys * 8^d, n) nodes, which sizes its grid.
  hipLaunchKernelGGL(k_upward_leaves, dim3(grid_for(n, 4096)), dim3(kBlock), 0, s, posm, n, t);
  const int top = static_cast<int>(depth_plan < kDepthCap ? depth_plan : kDepthCap);
  for (int d = top - 1; d >= 0; --d) {
    uint64_t bound = n_sys;
    for (int e = 0; e < d && bound < n; ++e) bound *= 8;
    if (bound > n) bound = n;
    hipLaunchKernelGGL(k_upward, dim3(grid_for(bound * kKids, 4096)), dim3(kBlock), 0, s, d, t);
  }
  return hipGetLastError();
}

hipError_t launch_targets(TreeDev t, uint32_t n, const uint16_t* sid, const SysDev* sys, bool groups, uint32_t* mark,
                          uint32_t* tile_counts, uint32_t* tile_prefix, uint32_t* targets, uint32_t* n_targets,
                          uint32_t* unit_n, hipStream_t s) {
  const uint32_t tiles = (n + kTreeTile - 1) / kTreeTile;
  hipError_t e = launch_fill_u32(mark, 0u, n, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_target_mark, dim3(grid_for(t.cap, 4096)), dim3(kBlock), 0, s, t, sid, sys, groups, mark);
  if (groups && unit_n)
    hipLaunchKernelGGL(k_group_units, dim3(grid_for(t.cap, 4096)), dim3(kBlock), 0, s, t, sid, sys, mark, unit_n);
  hipLaunchKernelGGL(k_compact_count, dim3(tiles), dim3(kBlock), 0, s, mark, n, tile_counts);
  hipLaunchKernelGGL(k_col_scan, dim3(1), dim3(1024), 0, s, tile_counts, tiles, 1u, tile_prefix, n_targets);
  hipLaunchKernelGGL(k_compact_write, dim3(tiles), dim3(kBlock), 0, s, mark, n, tile_prefix, targets);
  return hipGetLastError();
}

}  // namespace bhk
