// Barnes-Hut oracle: an FP64 CPU restatement of the reference's grouped
// force path (tree build + grouped walk + leapfrog kick/drift).
//
// TEST INFRASTRUCTURE ONLY. This library is the checker for the GPU path and
// the timed CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product path never calls it.
//
// Every function names the reference file:line it restates (paths relative
// to the reference repository root, e.g. ttg-parsec/src/flat_tree.h).
// Floating-point expressions keep the reference's operation order and the
// library is built with -ffp-contract=off, so keys, topology and multipoles
// are bit-identical to the reference built with GCC in ISO C++20 mode (its
// CMakeLists sets CMAKE_CXX_EXTENSIONS OFF, under which GCC does not
// contract a*b+c into FMAs).
//
// Pinning (what checks each part, besides the GPU):
//   domain/morton keys  : compiled reference functions (oracle/_ref), KATs
//   sort order          : brute-force std::sort on (key, index)
//   topology + layout   : compiled reference split/count/build_subtree
//   multipoles          : compiled reference finalize_node; closed form
//   grouped walk        : independent per-target root walk (bho_target_walk)
//                         and the compiled reference MAC + force functions;
//                         theta=0 equals brute force
//   force kernel        : two-body KAT (test_force_kernel.cc:12-25)
//   leapfrog            : schedule of main.cc:199-246; energy drift bound
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <thread>
#include <vector>

namespace {

// particle.h:7-24 layout: id@0, pad, mass@8, pos@16, vel@40 (64 B).
struct Particle {
  uint32_t id;
  uint32_t pad;
  double mass;
  double pos[3];
  double vel[3];
};
static_assert(sizeof(Particle) == 64, "Particle layout");

// flat_tree.h:61-70 WalkNode layout (56 B).
struct WalkNode {
  double com[3];
  double mass;
  double half_width;
  uint32_t first_child;
  uint32_t particle_begin;
  uint32_t particle_count;
  uint8_t num_children;
  uint8_t leaf;
  uint8_t depth;
  uint8_t pad;
};
static_assert(sizeof(WalkNode) == 56, "WalkNode layout");

struct ForceParams {  // traversal_ttg.h:136-141
  double theta;
  double eps2;
  double G;
  int32_t use_quadrupole;
  int32_t pad;
};

using Sym3 = std::array<double, 6>;  // xx, yy, zz, xy, xz, yz (multipole.h:12)

constexpr int kDepthCap = 21;                    // flat_tree.h:102
constexpr uint32_t kSubtreeCutoff = 4096;        // flat_build_ttg.h:52

int hw_threads(int requested) {
  if (requested > 0) return requested;
  unsigned h = std::thread::hardware_concurrency();
  return h == 0 ? 1 : static_cast<int>(h);
}

// Runs f(i) for i in [0, n) on `threads` threads, dynamic chunks.
template <typename F>
void parallel_for(uint64_t n, int threads, F&& f, uint64_t chunk = 1) {
  if (threads <= 1 || n <= chunk) {
    for (uint64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::atomic<uint64_t> next{0};
  auto body = [&]() {
    for (;;) {
      uint64_t b = next.fetch_add(chunk);
      if (b >= n) break;
      uint64_t e = std::min(n, b + chunk);
      for (uint64_t i = b; i < e; ++i) f(i);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < threads; ++t) pool.emplace_back(body);
  body();
  for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------- domain/key
// domain.h:23-30
void domain_from_bounds(const double lo[3], const double hi[3], double c[3], double* hw) {
  c[0] = 0.5 * (lo[0] + hi[0]);
  c[1] = 0.5 * (lo[1] + hi[1]);
  c[2] = 0.5 * (lo[2] + hi[2]);
  double ext = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  double h = 0.5 * ext;
  if (!(h > 0.0)) h = 1.0;
  h *= 1.0 + 1e-9;
  *hw = h;
}

// flat_build_ttg.h:93-113 (min/max is exact, so chunking does not matter)
void particle_bounds(const Particle* p, uint64_t n, double lo[3], double hi[3]) {
  const double inf = std::numeric_limits<double>::infinity();
  lo[0] = lo[1] = lo[2] = inf;
  hi[0] = hi[1] = hi[2] = -inf;
  for (uint64_t i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], p[i].pos[a]);
      hi[a] = std::max(hi[a], p[i].pos[a]);
    }
  }
}

// flat_tree.h:104-112
inline uint64_t spread_bits3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// flat_tree.h:114-125
inline uint64_t morton_key(const double c[3], double hw, const double* pos) {
  const double cells = 2097152.0;  // 2^21
  double scale = cells / (2.0 * hw);
  uint32_t q[3];
  for (int a = 0; a < 3; ++a) {
    double lo = c[a] - hw;
    double t = std::floor((pos[a] - lo) * scale);
    t = std::clamp(t, 0.0, cells - 1.0);
    q[a] = static_cast<uint32_t>(t);
  }
  return spread_bits3(q[0]) | (spread_bits3(q[1]) << 1) | (spread_bits3(q[2]) << 2);
}

// flat_tree.h:127-129
inline unsigned octant_digit(uint64_t key, int depth) { return static_cast<unsigned>((key >> (63 - 3 * depth)) & 7u); }

// ------------------------------------------------------------------- moments
// multipole.h:34-102 (RawMoment / finalize / parallel-axis shift), with the
// reference's left-to-right evaluation order.
struct RawMoment {
  double mass = 0.0;
  double m1[3] = {0.0, 0.0, 0.0};
  Sym3 m2{};
  void add(const RawMoment& o) {
    mass += o.mass;
    for (int a = 0; a < 3; ++a) m1[a] += o.m1[a];
    for (int i = 0; i < 6; ++i) m2[i] += o.m2[i];
  }
};

inline Sym3 outer_sym(const double d[3], double s) {
  return {s * d[0] * d[0], s * d[1] * d[1], s * d[2] * d[2], s * d[0] * d[1], s * d[0] * d[2], s * d[1] * d[2]};
}

inline RawMoment raw_moment_of_particle(double mass, const double pos[3], const double ref[3]) {
  RawMoment r;
  r.mass = mass;
  double d[3] = {pos[0] - ref[0], pos[1] - ref[1], pos[2] - ref[2]};
  for (int a = 0; a < 3; ++a) r.m1[a] = d[a] * mass;
  r.m2 = outer_sym(d, mass);
  return r;
}

inline void finalize_raw_moment(const RawMoment& rm, const double ref[3], double* mass, double com[3], Sym3& iab) {
  *mass = rm.mass;
  double inv = 1.0 / rm.mass;
  double delta[3] = {rm.m1[0] * inv, rm.m1[1] * inv, rm.m1[2] * inv};
  for (int a = 0; a < 3; ++a) com[a] = ref[a] + delta[a];
  Sym3 shift = outer_sym(delta, rm.mass);
  for (int i = 0; i < 6; ++i) iab[i] = rm.m2[i] - shift[i];
}

inline RawMoment raw_moment_from_multipole(double mass, const double com[3], const Sym3& iab, const double ref[3]) {
  RawMoment r;
  r.mass = mass;
  double delta[3] = {com[0] - ref[0], com[1] - ref[1], com[2] - ref[2]};
  for (int a = 0; a < 3; ++a) r.m1[a] = delta[a] * mass;
  Sym3 shift = outer_sym(delta, mass);
  r.m2 = iab;
  for (int i = 0; i < 6; ++i) r.m2[i] += shift[i];
  return r;
}

// ---------------------------------------------------------------------- tree
struct Tree {
  uint64_t n = 0;
  double dom_c[3] = {0, 0, 0};
  double dom_hw = 0;
  std::vector<uint64_t> keys;
  std::vector<uint32_t> perm;  // sorted position -> input index
  std::vector<Particle> particles;
  std::vector<double> x, y, z, m;
  std::vector<WalkNode> nodes;
  std::vector<std::array<double, 3>> center;
  std::vector<Sym3> iab;
  std::vector<uint32_t> groups;
  unsigned max_leaf = 1;
  uint32_t group_size = 0;
};

struct OctantRange {
  unsigned oct;
  uint32_t lo, hi;
};

// flat_tree.h:167-180: non-empty child octants of sorted keys [lo, hi)
int split_octants(const uint64_t* keys, uint32_t lo, uint32_t hi, int cdepth, OctantRange out[8]) {
  int k = 0;
  uint32_t j = lo;
  while (j < hi) {
    unsigned oct = octant_digit(keys[j], cdepth);
    const uint64_t* e = std::partition_point(keys + j, keys + hi, [&](uint64_t c) { return octant_digit(c, cdepth) <= oct; });
    uint32_t ei = static_cast<uint32_t>(e - keys);
    out[k++] = {oct, j, ei};
    j = ei;
  }
  return k;
}

// flat_tree.h:144-156 with the whole key range owned
inline bool leaf_stops(uint32_t count, int depth, unsigned max_leaf) { return count <= max_leaf || depth >= kDepthCap; }

// flat_tree.h:190-192
inline bool is_group_root(uint32_t count, bool leaf, uint32_t parent_count, uint32_t group_size) {
  return (leaf || count <= group_size) && parent_count > group_size;
}

// flat_tree.h:196-206
void init_node(Tree& t, uint32_t i, uint32_t begin, uint32_t count, const double c[3], double hw, int depth) {
  WalkNode& w = t.nodes[i];
  w = WalkNode{};
  w.particle_begin = begin;
  w.particle_count = count;
  w.half_width = hw;
  w.depth = static_cast<uint8_t>(depth);
  w.leaf = leaf_stops(count, depth, t.max_leaf) ? 1 : 0;
  t.center[i] = {c[0], c[1], c[2]};
}

// flat_tree.h:182-185 + 211-225
int split_node(Tree& t, uint32_t i, uint32_t first, std::vector<uint32_t>& groups) {
  OctantRange r[8];
  WalkNode& nd = t.nodes[i];
  int k = split_octants(t.keys.data(), nd.particle_begin, nd.particle_begin + nd.particle_count, nd.depth + 1, r);
  nd.first_child = first;
  nd.num_children = static_cast<uint8_t>(k);
  const double q = 0.5 * nd.half_width;
  const std::array<double, 3> pc = t.center[i];
  for (int c = 0; c < k; ++c) {
    uint32_t ci = first + static_cast<uint32_t>(c);
    double cc[3] = {pc[0] + ((r[c].oct & 1u) ? q : -q), pc[1] + ((r[c].oct & 2u) ? q : -q),
                    pc[2] + ((r[c].oct & 4u) ? q : -q)};
    init_node(t, ci, r[c].lo, r[c].hi - r[c].lo, cc, 0.5 * t.nodes[i].half_width, t.nodes[i].depth + 1);
    if (is_group_root(r[c].hi - r[c].lo, t.nodes[ci].leaf, t.nodes[i].particle_count, t.group_size)) groups.push_back(ci);
  }
  return k;
}

// flat_tree.h:236-251
uint32_t count_subtree(const Tree& t, uint32_t lo, uint32_t hi, int depth) {
  struct KR {
    uint32_t lo, hi;
    int depth;
  };
  std::vector<KR> stack{{lo, hi, depth}};
  uint32_t count = 0;
  while (!stack.empty()) {
    KR kr = stack.back();
    stack.pop_back();
    ++count;
    if (leaf_stops(kr.hi - kr.lo, kr.depth, t.max_leaf)) continue;
    OctantRange r[8];
    int k = split_octants(t.keys.data(), kr.lo, kr.hi, kr.depth + 1, r);
    for (int c = 0; c < k; ++c) stack.push_back({r[c].lo, r[c].hi, kr.depth + 1});
  }
  return count;
}

// flat_tree.h:257-274
void finalize_node(Tree& t, uint32_t i) {
  WalkNode& nd = t.nodes[i];
  const double ref[3] = {t.center[i][0], t.center[i][1], t.center[i][2]};
  RawMoment rm;
  if (nd.leaf) {
    for (uint32_t p = nd.particle_begin; p < nd.particle_begin + nd.particle_count; ++p) {
      double pos[3] = {t.x[p], t.y[p], t.z[p]};
      rm.add(raw_moment_of_particle(t.m[p], pos, ref));
    }
  } else {
    for (uint8_t c = 0; c < nd.num_children; ++c) {
      const uint32_t ci = nd.first_child + c;
      const WalkNode& ch = t.nodes[ci];
      rm.add(raw_moment_from_multipole(ch.mass, ch.com, t.iab[ci], ref));
    }
  }
  finalize_raw_moment(rm, ref, &nd.mass, nd.com, t.iab[i]);
}

// flat_tree.h:280-289
void build_subtree(Tree& t, uint32_t root, uint32_t base, std::vector<uint32_t>& groups) {
  uint32_t next = base;
  if (!t.nodes[root].leaf) next += split_node(t, root, next, groups);
  for (uint32_t i = base; i < next; ++i) {
    if (!t.nodes[i].leaf) next += split_node(t, i, next, groups);
  }
  for (uint32_t i = next; i-- > base;) finalize_node(t, i);
  finalize_node(t, root);
}

// flat_build_ttg.h:117-190 (result only: ascending (key, input index)). The
// parallel form mirrors the reference's phases: per-chunk key+sort, regular
// samples, splitters, per-bucket merge; the order is unique either way.
void sort_particles(Tree& t, const Particle* in, int threads) {
  const uint64_t n = t.n;
  struct KI {
    uint64_t key;
    uint32_t idx;
    bool operator<(const KI& o) const { return key < o.key || (key == o.key && idx < o.idx); }
  };
  std::vector<KI> ki(n);
  const uint32_t P = static_cast<uint32_t>(std::clamp<uint64_t>(n / 4096, 1, 1024));
  auto chunk_begin = [&](uint64_t c) { return n * c / P; };
  parallel_for(P, threads, [&](uint64_t c) {
    for (uint64_t i = chunk_begin(c); i < chunk_begin(c + 1); ++i) ki[i] = {morton_key(t.dom_c, t.dom_hw, in[i].pos), static_cast<uint32_t>(i)};
    std::sort(ki.begin() + chunk_begin(c), ki.begin() + chunk_begin(c + 1));
  });
  if (P > 1) {
    std::vector<KI> samples;
    for (uint32_t c = 0; c < P; ++c) {
      uint64_t b = chunk_begin(c), len = chunk_begin(c + 1) - b;
      for (uint32_t s = 0; s < 16 && len > 0; ++s) samples.push_back(ki[b + len * s / 16]);
    }
    std::sort(samples.begin(), samples.end());
    std::vector<KI> split(P - 1);
    for (uint32_t b = 1; b < P; ++b) split[b - 1] = samples[samples.size() * b / P];
    std::vector<uint64_t> bounds(uint64_t(P) * (P + 1));
    parallel_for(P, threads, [&](uint64_t c) {
      uint64_t* bc = &bounds[c * (P + 1)];
      auto first = ki.begin() + chunk_begin(c), last = ki.begin() + chunk_begin(c + 1);
      bc[0] = chunk_begin(c);
      for (uint32_t b = 1; b < P; ++b) bc[b] = std::lower_bound(first, last, split[b - 1]) - ki.begin();
      bc[P] = chunk_begin(c + 1);
    });
    std::vector<uint64_t> bucket_begin(P + 1, 0);
    for (uint32_t b = 0; b < P; ++b) {
      uint64_t size = 0;
      for (uint32_t c = 0; c < P; ++c) size += bounds[c * (P + 1) + b + 1] - bounds[c * (P + 1) + b];
      bucket_begin[b + 1] = bucket_begin[b] + size;
    }
    std::vector<KI> sorted(n);
    parallel_for(P, threads, [&](uint64_t b) {
      uint64_t out = bucket_begin[b];
      for (uint32_t c = 0; c < P; ++c) {
        const uint64_t* bc = &bounds[c * (P + 1)];
        for (uint64_t i = bc[b]; i < bc[b + 1]; ++i) sorted[out++] = ki[i];
      }
      std::sort(sorted.begin() + bucket_begin[b], sorted.begin() + bucket_begin[b + 1]);
    });
    ki.swap(sorted);
  }
  t.keys.resize(n);
  t.perm.resize(n);
  t.particles.resize(n);
  t.x.resize(n);
  t.y.resize(n);
  t.z.resize(n);
  t.m.resize(n);
  parallel_for(
      n, threads,
      [&](uint64_t i) {
        const Particle& p = in[ki[i].idx];
        t.keys[i] = ki[i].key;
        t.perm[i] = ki[i].idx;
        t.particles[i] = p;
        t.x[i] = p.pos[0];
        t.y[i] = p.pos[1];
        t.z[i] = p.pos[2];
        t.m[i] = p.mass;
      },
      65536);
}

// flat_build_ttg.h:194-273: backbone BFS, subtree counts and layout,
// in-place subtree builds, backbone multipoles, group concatenation.
void build_nodes(Tree& t, int threads) {
  t.groups.clear();
  t.nodes.clear();
  t.center.clear();
  t.iab.clear();
  const uint64_t n = t.n;
  if (n == 0) return;
  auto grow = [&](size_t s) {
    t.nodes.resize(s);
    t.center.resize(s);
    t.iab.resize(s);
  };
  grow(1);
  init_node(t, 0, 0, static_cast<uint32_t>(n), t.dom_c, t.dom_hw, 0);
  if (is_group_root(static_cast<uint32_t>(n), t.nodes[0].leaf, std::numeric_limits<uint32_t>::max(), t.group_size)) t.groups.push_back(0);
  auto is_subtree_root = [&](uint32_t i) { return t.nodes[i].particle_count <= kSubtreeCutoff || t.nodes[i].leaf; };
  std::vector<uint32_t> sroot;
  for (uint32_t i = 0; i < t.nodes.size(); ++i) {
    if (is_subtree_root(i)) {
      sroot.push_back(i);
      continue;
    }
    uint32_t first = static_cast<uint32_t>(t.nodes.size());
    grow(first + 8);
    grow(first + split_node(t, i, first, t.groups));
  }
  const uint32_t backbone = static_cast<uint32_t>(t.nodes.size());
  const uint32_t ns = static_cast<uint32_t>(sroot.size());
  std::vector<uint32_t> cnt(ns);
  parallel_for(ns, threads, [&](uint64_t s) {
    const WalkNode& r = t.nodes[sroot[s]];
    cnt[s] = count_subtree(t, r.particle_begin, r.particle_begin + r.particle_count, r.depth);
  });
  std::vector<uint32_t> base(ns + 1);
  base[0] = backbone;
  for (uint32_t s = 0; s < ns; ++s) base[s + 1] = base[s] + cnt[s] - 1;
  grow(base[ns]);
  std::vector<std::vector<uint32_t>> sg(ns);
  parallel_for(ns, threads, [&](uint64_t s) { build_subtree(t, sroot[s], base[s], sg[s]); });
  for (uint32_t i = backbone; i-- > 0;) {
    if (!is_subtree_root(i)) finalize_node(t, i);
  }
  for (uint32_t s = 0; s < ns; ++s) t.groups.insert(t.groups.end(), sg[s].begin(), sg[s].end());
}

// flat_build_ttg.h:276-281
void build_flat_tree(Tree& t, const Particle* in, uint64_t n, unsigned max_leaf, uint32_t group_size, int threads) {
  t.n = n;
  t.max_leaf = max_leaf;
  t.group_size = group_size;
  double lo[3], hi[3];
  particle_bounds(in, n, lo, hi);
  domain_from_bounds(lo, hi, t.dom_c, &t.dom_hw);
  sort_particles(t, in, threads);
  build_nodes(t, threads);
}

// ---------------------------------------------------------------------- walk
struct BBox {  // multipole.h:115-142
  double lo[3], hi[3];
  BBox() {
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::numeric_limits<double>::max();
      hi[a] = std::numeric_limits<double>::lowest();
    }
  }
  void add(double x, double y, double z) {
    lo[0] = std::min(lo[0], x);
    lo[1] = std::min(lo[1], y);
    lo[2] = std::min(lo[2], z);
    hi[0] = std::max(hi[0], x);
    hi[1] = std::max(hi[1], y);
    hi[2] = std::max(hi[2], z);
  }
  double near2(const double p[3]) const {
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = std::max(std::max(lo[a] - p[a], p[a] - hi[a]), 0.0);
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  }
  double far2(const double p[3]) const {
    double d[3];
    for (int a = 0; a < 3; ++a) d[a] = std::max(p[a] - lo[a], hi[a] - p[a]);
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  }
};

// multipole.h:151-160
inline bool mac_accept(const BBox& b, const double com[3], double edge, double theta) { return theta * theta * b.near2(com) > edge * edge; }
inline bool mac_reject_all(const BBox& b, const double com[3], double edge, double theta) { return theta * theta * b.far2(com) <= edge * edge; }

// multipole.h:171-200 (G applied by the caller when G == 1.0 is passed)
inline void accel_from_source(const double tp[3], double mass, const double com[3], const Sym3& iab, double eps2, bool quad, double G, double out[3]) {
  double d[3] = {com[0] - tp[0], com[1] - tp[1], com[2] - tp[2]};
  double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  double ds2 = d2 + eps2;
  double ds = std::sqrt(ds2);
  double inv3 = 1.0 / (ds2 * ds);
  double s = mass * inv3;
  double a[3] = {d[0] * s, d[1] * s, d[2] * s};
  if (quad) {
    double tr = iab[0] + iab[1] + iab[2];
    double q[6];
    for (int i = 0; i < 3; ++i) q[i] = 3.0 * iab[i] - tr;
    for (int i = 3; i < 6; ++i) q[i] = 3.0 * iab[i];
    double qd[3] = {q[0] * d[0] + q[3] * d[1] + q[4] * d[2], q[3] * d[0] + q[1] * d[1] + q[5] * d[2],
                    q[4] * d[0] + q[5] * d[1] + q[2] * d[2]};
    double dqd = d[0] * qd[0] + d[1] * qd[1] + d[2] * qd[2];
    double inv5 = inv3 / ds2;
    double inv7 = inv5 / ds2;
    double s5 = (-1.0) * inv5;
    for (int a2 = 0; a2 < 3; ++a2) a[a2] += qd[a2] * s5;
    double s7 = 2.5 * dqd * inv7;
    for (int a2 = 0; a2 < 3; ++a2) a[a2] += d[a2] * s7;
  }
  for (int a2 = 0; a2 < 3; ++a2) out[a2] = a[a2] * G;
}

struct WalkOut {
  double* acc;          // 3 per particle, Morton order (nullable)
  uint32_t* n_pp;       // per particle: P2P list length (nullable)
  uint32_t* n_pc;       // per particle: quadrupole cell count (nullable)
  std::atomic<uint64_t> tot_pp{0}, tot_pc{0};
};

// One target's evaluation: P2P list (with Mask when eps2 == 0,
// grouped_ttg.h:124-163) then quadrupole cells (grouped_ttg.h:191-196).
void evaluate_target(const Tree& t, uint32_t target, const ForceParams& fp, const std::vector<double>& lx,
                     const std::vector<double>& ly, const std::vector<double>& lz, const std::vector<double>& lm, size_t ln,
                     const std::vector<uint32_t>& cells, WalkOut& out) {
  const WalkNode& tn = t.nodes[target];
  for (uint32_t p = tn.particle_begin; p < tn.particle_begin + tn.particle_count; ++p) {
    const double px = t.x[p], py = t.y[p], pz = t.z[p];
    double ax = 0, ay = 0, az = 0;
    for (size_t j = 0; j < ln; ++j) {
      double dx = lx[j] - px, dy = ly[j] - py, dz = lz[j] - pz;
      double r2 = dx * dx + dy * dy + dz * dz + fp.eps2;
      double inv = (r2 > 0.0) ? 1.0 / std::sqrt(r2) : 0.0;
      double s = lm[j] * inv * inv * inv;
      ax += s * dx;
      ay += s * dy;
      az += s * dz;
    }
    const double tp[3] = {px, py, pz};
    for (uint32_t c : cells) {
      const WalkNode& sn = t.nodes[c];
      double a[3];
      accel_from_source(tp, sn.mass, sn.com, t.iab[c], fp.eps2, true, 1.0, a);
      ax += a[0];
      ay += a[1];
      az += a[2];
    }
    if (out.acc) {
      out.acc[3 * p + 0] = fp.G * ax;
      out.acc[3 * p + 1] = fp.G * ay;
      out.acc[3 * p + 2] = fp.G * az;
    }
    if (out.n_pp) out.n_pp[p] = static_cast<uint32_t>(ln);
    if (out.n_pc) out.n_pc[p] = static_cast<uint32_t>(cells.size());
  }
  out.tot_pp += uint64_t(ln) * tn.particle_count;
  out.tot_pc += uint64_t(cells.size()) * tn.particle_count;
}

// grouped_ttg.h:204-289: the shared walk of one group.
struct GroupWalker {
  const Tree& t;
  const ForceParams& fp;
  uint32_t target_size;
  WalkOut& out;
  std::vector<double> lx, ly, lz, lm;
  size_t ln = 0;
  std::vector<uint32_t> cells;
  std::vector<uint32_t> work;
  std::vector<std::vector<uint32_t>> open;

  GroupWalker(const Tree& t, const ForceParams& fp, uint32_t ts, WalkOut& o)
      : t(t), fp(fp), target_size(ts), out(o), open(kDepthCap + 2) {}

  void push(double x, double y, double z, double m) {
    if (ln == lx.size()) {
      size_t cap = std::max<size_t>(64, 2 * lx.size());
      lx.resize(cap);
      ly.resize(cap);
      lz.resize(cap);
      lm.resize(cap);
    }
    lx[ln] = x;
    ly[ln] = y;
    lz[ln] = z;
    lm[ln] = m;
    ++ln;
  }

  void push_source(uint32_t s) {  // grouped_ttg.h:223-231
    const WalkNode& sn = t.nodes[s];
    if (fp.use_quadrupole) {
      cells.push_back(s);
    } else {
      push(sn.com[0], sn.com[1], sn.com[2], sn.mass);
    }
  }

  void open_source(uint32_t s) {  // grouped_ttg.h:234-249
    const WalkNode& sn = t.nodes[s];
    if (sn.leaf) {
      for (uint32_t p = sn.particle_begin; p < sn.particle_begin + sn.particle_count; ++p) push(t.x[p], t.y[p], t.z[p], t.m[p]);
    } else {
      for (uint8_t c = 0; c < sn.num_children; ++c) work.push_back(sn.first_child + c);
    }
  }

  void visit(uint32_t i, int depth) {  // grouped_ttg.h:251-288
    const WalkNode& tn = t.nodes[i];
    const bool target = tn.leaf || tn.particle_count <= target_size;
    BBox box;
    for (uint32_t p = tn.particle_begin; p < tn.particle_begin + tn.particle_count; ++p) box.add(t.x[p], t.y[p], t.z[p]);
    const size_t list_mark = ln, cells_mark = cells.size();
    std::vector<uint32_t>& rest = open[depth + 1];
    rest.clear();
    work.assign(open[depth].begin(), open[depth].end());
    while (!work.empty()) {
      uint32_t s = work.back();
      work.pop_back();
      const WalkNode& sn = t.nodes[s];
      double edge = 2.0 * sn.half_width;
      if (mac_accept(box, sn.com, edge, fp.theta)) {
        push_source(s);
      } else if (target || mac_reject_all(box, sn.com, edge, fp.theta)) {
        open_source(s);
      } else {
        rest.push_back(s);
      }
    }
    if (target) {
      evaluate_target(t, i, fp, lx, ly, lz, lm, ln, cells, out);
    } else {
      for (uint8_t c = 0; c < tn.num_children; ++c) visit(tn.first_child + c, depth + 1);
    }
    ln = list_mark;
    cells.resize(cells_mark);
  }

  void walk_group(uint32_t g) {
    open[0].assign(1, 0u);
    visit(g, 0);
  }
};

// The targets reachable from the group roots (grouped_ttg.h:253): topmost
// nodes below a group root with leaf || count <= target_size.
void collect_targets(const Tree& t, uint32_t target_size, std::vector<uint32_t>& out) {
  std::vector<uint32_t> st;
  for (uint32_t g : t.groups) {
    st.assign(1, g);
    while (!st.empty()) {
      uint32_t i = st.back();
      st.pop_back();
      const WalkNode& nd = t.nodes[i];
      if (nd.leaf || nd.particle_count <= target_size) {
        out.push_back(i);
      } else {
        for (int c = nd.num_children - 1; c >= 0; --c) st.push_back(nd.first_child + c);
      }
    }
  }
}

// Independent per-target walk from the root (the equivalence stated at
// grouped_ttg.h:42-51): accept by the target's bbox MAC, else direct-sum a
// leaf, else open its children.
void target_walk(const Tree& t, uint32_t target, const ForceParams& fp, WalkOut& out, std::vector<double>& lx,
                 std::vector<double>& ly, std::vector<double>& lz, std::vector<double>& lm, std::vector<uint32_t>& cells) {
  const WalkNode& tn = t.nodes[target];
  BBox box;
  for (uint32_t p = tn.particle_begin; p < tn.particle_begin + tn.particle_count; ++p) box.add(t.x[p], t.y[p], t.z[p]);
  size_t ln = 0;
  cells.clear();
  auto push = [&](double x, double y, double z, double m) {
    if (ln == lx.size()) {
      size_t cap = std::max<size_t>(64, 2 * lx.size());
      lx.resize(cap);
      ly.resize(cap);
      lz.resize(cap);
      lm.resize(cap);
    }
    lx[ln] = x;
    ly[ln] = y;
    lz[ln] = z;
    lm[ln] = m;
    ++ln;
  };
  std::vector<uint32_t> st{0u};
  while (!st.empty()) {
    uint32_t s = st.back();
    st.pop_back();
    const WalkNode& sn = t.nodes[s];
    double edge = 2.0 * sn.half_width;
    if (mac_accept(box, sn.com, edge, fp.theta)) {
      if (fp.use_quadrupole)
        cells.push_back(s);
      else
        push(sn.com[0], sn.com[1], sn.com[2], sn.mass);
    } else if (sn.leaf) {
      for (uint32_t p = sn.particle_begin; p < sn.particle_begin + sn.particle_count; ++p) push(t.x[p], t.y[p], t.z[p], t.m[p]);
    } else {
      for (uint8_t c = 0; c < sn.num_children; ++c) st.push_back(sn.first_child + c);
    }
  }
  evaluate_target(t, target, fp, lx, ly, lz, lm, ln, cells, out);
}

// grouped_ttg.h:297-323: one unit of work per group.
void grouped_pass(const Tree& t, const ForceParams& fp, uint32_t target_size, int threads, WalkOut& out) {
  parallel_for(t.groups.size(), threads, [&](uint64_t g) {
    GroupWalker w(t, fp, target_size, out);
    w.walk_group(t.groups[g]);
  });
}

// integrate.h:15-16 + integrate_ttg.h:74-77
inline void kick_drift(Particle& p, const double a[3], double kick_dt, double drift_dt) {
  for (int c = 0; c < 3; ++c) p.vel[c] += a[c] * kick_dt;
  if (drift_dt != 0.0) {
    for (int c = 0; c < 3; ++c) p.pos[c] += p.vel[c] * drift_dt;
  }
}

bool finite_input(const Particle* p, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    if (!std::isfinite(p[i].mass)) return false;
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(p[i].pos[a]) || !std::isfinite(p[i].vel[a])) return false;
  }
  return true;
}

}  // namespace

extern "C" {

struct bho_tree;

int bho_domain(const Particle* p, uint64_t n, double* out4) {
  if (n == 0) return 1;
  double lo[3], hi[3], c[3], hw;
  particle_bounds(p, n, lo, hi);
  domain_from_bounds(lo, hi, c, &hw);
  out4[0] = c[0];
  out4[1] = c[1];
  out4[2] = c[2];
  out4[3] = hw;
  return 0;
}

int bho_morton_keys(const Particle* p, uint64_t n, const double* dom4, uint64_t* keys) {
  for (uint64_t i = 0; i < n; ++i) keys[i] = morton_key(dom4, dom4[3], p[i].pos);
  return 0;
}

uint64_t bho_spread_bits3(uint32_t v) { return spread_bits3(v); }

int bho_build(const Particle* p, uint64_t n, uint32_t max_leaf, uint32_t group_size, int threads, bho_tree** out) {
  if (n == 0 || n >= (1ull << 31) || !finite_input(p, n)) return 1;
  Tree* t = new Tree();
  build_flat_tree(*t, p, n, max_leaf, group_size, hw_threads(threads));
  *out = reinterpret_cast<bho_tree*>(t);
  return 0;
}

void bho_tree_free(bho_tree* h) { delete reinterpret_cast<Tree*>(h); }

int bho_tree_sizes(const bho_tree* h, uint64_t* n_nodes, uint64_t* n_groups, double* dom4) {
  const Tree* t = reinterpret_cast<const Tree*>(h);
  *n_nodes = t->nodes.size();
  *n_groups = t->groups.size();
  if (dom4) {
    dom4[0] = t->dom_c[0];
    dom4[1] = t->dom_c[1];
    dom4[2] = t->dom_c[2];
    dom4[3] = t->dom_hw;
  }
  return 0;
}

int bho_tree_export(const bho_tree* h, uint64_t* keys, uint32_t* perm, Particle* sorted, WalkNode* nodes, double* center,
                    double* iab, uint32_t* groups) {
  const Tree* t = reinterpret_cast<const Tree*>(h);
  if (keys) std::memcpy(keys, t->keys.data(), t->n * 8);
  if (perm) std::memcpy(perm, t->perm.data(), t->n * 4);
  if (sorted) std::memcpy(sorted, t->particles.data(), t->n * 64);
  if (nodes) std::memcpy(nodes, t->nodes.data(), t->nodes.size() * sizeof(WalkNode));
  if (center)
    for (size_t i = 0; i < t->center.size(); ++i)
      for (int a = 0; a < 3; ++a) center[3 * i + a] = t->center[i][a];
  if (iab)
    for (size_t i = 0; i < t->iab.size(); ++i)
      for (int a = 0; a < 6; ++a) iab[6 * i + a] = t->iab[i][a];
  if (groups) std::memcpy(groups, t->groups.data(), t->groups.size() * 4);
  return 0;
}

// Grouped pass (FP64). acc/n_pp/n_pc are indexed by Morton position.
int bho_grouped_pass(const bho_tree* h, const ForceParams* fp, uint32_t target_size, int threads, double* acc, uint32_t* n_pp,
                     uint32_t* n_pc, uint64_t* totals2) {
  const Tree* t = reinterpret_cast<const Tree*>(h);
  WalkOut out;
  out.acc = acc;
  out.n_pp = n_pp;
  out.n_pc = n_pc;
  grouped_pass(*t, *fp, target_size, hw_threads(threads), out);
  if (totals2) {
    totals2[0] = out.tot_pp;
    totals2[1] = out.tot_pc;
  }
  return 0;
}

// Independent per-target walks over the same targets as the grouped pass.
int bho_target_walk(const bho_tree* h, const ForceParams* fp, uint32_t target_size, int threads, double* acc, uint32_t* n_pp,
                    uint32_t* n_pc, uint64_t* totals2) {
  const Tree* t = reinterpret_cast<const Tree*>(h);
  std::vector<uint32_t> targets;
  collect_targets(*t, target_size, targets);
  WalkOut out;
  out.acc = acc;
  out.n_pp = n_pp;
  out.n_pc = n_pc;
  parallel_for(targets.size(), hw_threads(threads), [&](uint64_t k) {
    std::vector<double> lx, ly, lz, lm;
    std::vector<uint32_t> cells;
    target_walk(*t, targets[k], *fp, out, lx, ly, lz, lm, cells);
  });
  if (totals2) {
    totals2[0] = out.tot_pp;
    totals2[1] = out.tot_pc;
  }
  return 0;
}

// The per-node TTG path (traversal_ttg.h:156-196) restated serially: every
// leaf is a target; for each (target leaf, source node), starting at the
// root: the MAC on the target leaf's particle box (never for the target
// itself) accepts the source's multipole; a leaf source not accepted is
// summed directly, skipping the particle itself by id; any other source is
// descended into its children. Contributions are summed per source subtree
// in child order (the reference's aggregator order is unspecified).
// Per-particle counts: direct sources (n_pp) and accepted cells (n_pc).
int bho_per_node_pass(const bho_tree* h, const ForceParams* fp, int threads, double* acc, uint32_t* n_pp,
                      uint32_t* n_pc) {
  const Tree* t = reinterpret_cast<const Tree*>(h);
  std::vector<uint32_t> leaves;
  for (uint32_t i = 0; i < t->nodes.size(); ++i)
    if (t->nodes[i].leaf) leaves.push_back(i);
  parallel_for(leaves.size(), hw_threads(threads), [&](uint64_t k) {
    const uint32_t tl = leaves[k];
    const WalkNode& tn = t->nodes[tl];
    BBox box;
    for (uint32_t i = tn.particle_begin; i < tn.particle_begin + tn.particle_count; ++i)
      box.add(t->x[i], t->y[i], t->z[i]);
    for (uint32_t i = tn.particle_begin; i < tn.particle_begin + tn.particle_count; ++i) {
      const Particle& tp = t->particles[i];
      uint32_t pp = 0, pc = 0;
      // recursive descent, children in order (source-subtree sums)
      auto visit = [&](auto&& self, uint32_t s) -> std::array<double, 3> {
        const WalkNode& sn = t->nodes[s];
        std::array<double, 3> a{0.0, 0.0, 0.0};
        if (s != tl && mac_accept(box, sn.com, 2.0 * sn.half_width, fp->theta)) {
          double o[3];
          accel_from_source(tp.pos, sn.mass, sn.com, t->iab[s], fp->eps2, fp->use_quadrupole != 0, fp->G, o);
          ++pc;
          return {o[0], o[1], o[2]};
        }
        if (sn.leaf) {
          static const Sym3 zero{};
          for (uint32_t j = sn.particle_begin; j < sn.particle_begin + sn.particle_count; ++j) {
            const Particle& sp = t->particles[j];
            ++pp;
            if (sp.id == tp.id) continue;
            double o[3];
            accel_from_source(tp.pos, sp.mass, sp.pos, zero, fp->eps2, false, fp->G, o);
            a[0] += o[0];
            a[1] += o[1];
            a[2] += o[2];
          }
          return a;
        }
        for (uint32_t c = 0; c < sn.num_children; ++c) {
          const auto b = self(self, sn.first_child + c);
          a[0] += b[0];
          a[1] += b[1];
          a[2] += b[2];
        }
        return a;
      };
      const auto a = visit(visit, 0u);
      acc[3 * i + 0] = a[0];
      acc[3 * i + 1] = a[1];
      acc[3 * i + 2] = a[2];
      if (n_pp) n_pp[i] = pp;
      if (n_pc) n_pc[i] = pc;
    }
  });
  return 0;
}

// Target node ids in Morton order (topmost with leaf || count <= target_size
// below the group roots); returns the count, fills up to cap entries.
uint64_t bho_targets(const bho_tree* h, uint32_t target_size, uint32_t* out, uint64_t cap) {
  const Tree* t = reinterpret_cast<const Tree*>(h);
  std::vector<uint32_t> targets;
  collect_targets(*t, target_size, targets);
  std::sort(targets.begin(), targets.end(), [&](uint32_t a, uint32_t b) { return t->nodes[a].particle_begin < t->nodes[b].particle_begin; });
  for (uint64_t i = 0; i < targets.size() && i < cap; ++i) out[i] = targets[i];
  return targets.size();
}

// integrate_ttg.h:88-107, grouped branch on one rank: build, walk with
// fused kick/drift, return the particles in Morton order (in place).
// timing3 (nullable): build_ms, walk_ms, 0 (StepTiming, integrate_ttg.h:24-28).
int bho_step(Particle* p, uint64_t n, const ForceParams* fp, uint32_t max_leaf, uint32_t group_size, uint32_t target_size,
             double kick_dt, double drift_dt, int threads, double* timing3) {
  if (n == 0 || n >= (1ull << 31) || group_size == 0 || !finite_input(p, n)) return 1;
  const int th = hw_threads(threads);
  using clock = std::chrono::steady_clock;
  auto t0 = clock::now();
  Tree t;
  build_flat_tree(t, p, n, max_leaf, group_size, th);
  auto t1 = clock::now();
  std::vector<double> acc(3 * n);
  WalkOut out;
  out.acc = acc.data();
  out.n_pp = nullptr;
  out.n_pc = nullptr;
  grouped_pass(t, *fp, target_size, th, out);
  parallel_for(
      n, th, [&](uint64_t i) { kick_drift(t.particles[i], &acc[3 * i], kick_dt, drift_dt); }, 65536);
  std::memcpy(p, t.particles.data(), n * sizeof(Particle));
  auto t2 = clock::now();
  if (timing3) {
    timing3[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    timing3[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
    timing3[2] = 0.0;
  }
  return 0;
}

// Timed CPU baseline sample: full build of all n particles, then the grouped
// walk over every `stride`-th group only (no kick/drift). out3: build_ms,
// walk_ms over the sampled groups, fraction of particles those groups hold.
int bho_baseline_sample(const Particle* p, uint64_t n, const ForceParams* fp, uint32_t max_leaf, uint32_t group_size,
                        uint32_t target_size, uint32_t stride, int threads, double* out3, double* acc_out,
                        uint8_t* sampled) {
  if (n == 0 || n >= (1ull << 31) || group_size == 0 || stride == 0) return 1;
  const int th = hw_threads(threads);
  using clock = std::chrono::steady_clock;
  auto t0 = clock::now();
  Tree t;
  build_flat_tree(t, p, n, max_leaf, group_size, th);
  auto t1 = clock::now();
  std::vector<uint32_t> sample;
  uint64_t covered = 0;
  for (size_t g = 0; g < t.groups.size(); g += stride) {
    sample.push_back(t.groups[g]);
    covered += t.nodes[t.groups[g]].particle_count;
  }
  std::vector<double> acc(3 * n);
  WalkOut out;
  out.acc = acc.data();
  out.n_pp = nullptr;
  out.n_pc = nullptr;
  parallel_for(sample.size(), th, [&](uint64_t k) {
    GroupWalker w(t, *fp, target_size, out);
    w.walk_group(sample[k]);
  });
  auto t2 = clock::now();
  out3[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
  out3[1] = std::chrono::duration<double, std::milli>(t2 - t1).count();
  out3[2] = static_cast<double>(covered) / static_cast<double>(n);
  // the sampled groups' accelerations, Morton order (optional)
  if (acc_out) std::copy(acc.begin(), acc.end(), acc_out);
  if (sampled) {
    std::fill(sampled, sampled + n, uint8_t(0));
    for (uint32_t g : sample) {
      const auto& nd = t.nodes[g];
      std::fill(sampled + nd.particle_begin, sampled + nd.particle_begin + nd.particle_count, uint8_t(1));
    }
  }
  return 0;
}

// brute_force.h:15-28: O(N^2) FP64 accelerations, i == j skipped, G applied
// per pair as accel_from_source does.
int bho_brute_force(const Particle* p, uint64_t n, double eps2, double G, int threads, double* acc) {
  static const Sym3 zero{};
  parallel_for(
      n, hw_threads(threads),
      [&](uint64_t i) {
        double a[3] = {0, 0, 0};
        for (uint64_t j = 0; j < n; ++j) {
          if (i == j) continue;
          double t[3];
          accel_from_source(p[i].pos, p[j].mass, p[j].pos, zero, eps2, false, G, t);
          a[0] += t[0];
          a[1] += t[1];
          a[2] += t[2];
        }
        acc[3 * i + 0] = a[0];
        acc[3 * i + 1] = a[1];
        acc[3 * i + 2] = a[2];
      },
      16);
  return 0;
}

// Brute force on the targets idx[0..m) only (all n sources).
int bho_brute_subset(const Particle* p, uint64_t n, const uint32_t* idx, uint64_t m, double eps2, double G, int threads,
                     double* acc) {
  static const Sym3 zero{};
  parallel_for(
      m, hw_threads(threads),
      [&](uint64_t k) {
        const uint64_t i = idx[k];
        double a[3] = {0, 0, 0};
        for (uint64_t j = 0; j < n; ++j) {
          if (i == j) continue;
          double t[3];
          accel_from_source(p[i].pos, p[j].mass, p[j].pos, zero, eps2, false, G, t);
          a[0] += t[0];
          a[1] += t[1];
          a[2] += t[2];
        }
        acc[3 * k + 0] = a[0];
        acc[3 * k + 1] = a[1];
        acc[3 * k + 2] = a[2];
      },
      4);
  return 0;
}

// accel_from_source (multipole.h:171-200) exposed for the force-kernel KAT.
void bho_accel_from_source(const double* tp, double mass, const double* com, const double* iab6, double eps2, int quad,
                           double G, double* out3) {
  Sym3 iab{};
  for (int i = 0; i < 6; ++i) iab[i] = iab6[i];
  accel_from_source(tp, mass, com, iab, eps2, quad != 0, G, out3);
}

// integrate.h:27-31
double bho_kinetic_energy(const Particle* p, uint64_t n) {
  double e = 0.0;
  for (uint64_t i = 0; i < n; ++i) {
    const double v2 = p[i].vel[0] * p[i].vel[0] + p[i].vel[1] * p[i].vel[1] + p[i].vel[2] * p[i].vel[2];
    e += 0.5 * p[i].mass * v2;
  }
  return e;
}

// integrate.h:35-43 (unsoftened, pair order i < j); per-i partial sums are
// combined in i order so the result is independent of the thread count.
double bho_potential_energy(const Particle* p, uint64_t n, double G, int threads) {
  std::vector<double> part(n, 0.0);
  parallel_for(
      n, hw_threads(threads),
      [&](uint64_t i) {
        double e = 0.0;
        for (uint64_t j = i + 1; j < n; ++j) {
          const double d[3] = {p[i].pos[0] - p[j].pos[0], p[i].pos[1] - p[j].pos[1], p[i].pos[2] - p[j].pos[2]};
          e -= G * p[i].mass * p[j].mass / std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        }
        part[i] = e;
      },
      16);
  double e = 0.0;
  for (uint64_t i = 0; i < n; ++i) e += part[i];
  return e;
}

}  // extern "C"
