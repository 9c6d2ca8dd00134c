// Driver that links the reference's own TTG-free headers into a small shared
// library (oracle/_ref/libbhref.so), used only to pin the oracle.
//
// TEST INFRASTRUCTURE ONLY: built from the reference sources where they lie
// (see oracle/Makefile, target `ref`); the sources are never copied into this
// repository and nothing here runs on the product path.
//
// What runs reference code (compiled from ttg-parsec/src):
//   bh::compute_domain / domain_from_bounds   (domain.h:23-52)
//   bh::morton_key / spread_bits3             (flat_tree.h:104-125)
//   bh::detail::init_node / split_node / count_subtree / build_subtree /
//     finalize_node / is_group_root           (flat_tree.h:182-289)
//   bh::BBox / mac_accept_bbox                (multipole.h:115-153)
//   bh::accel_from_source                     (multipole.h:171-200)
//   bh::brute_force_accel                     (brute_force.h:15-28)
//   bh::kick_one / drift_one / energies       (integrate.h:15-43)
//   bh::Tree (the per-node path's octree)     (tree.h:76-132, tree.cc:30-100)
//   bh::Key / box_geometry                    (key.h, domain.h:59-67)
// What this driver supplies itself (the reference's versions live in headers
// that need the TTG runtime): the (key, index) sort, the serial loop that
// orders backbone split / subtree count / subtree build / backbone finalize
// as flat_build_ttg.h:194-273 does, a per-target walk loop, and for the
// per-node path the upward pass (upward_ttg.h:75-118: leaf raw moments about
// the leaf's own box centre, interior sums of the children's shifted
// moments, here in octant order where TTG sums them in arrival order) and
// the dispatch recursion (traversal_ttg.h:156-196) as one serial walk per
// target leaf.
#include "brute_force.h"
#include "flat_tree.h"
#include "integrate.h"
#include "multipole.h"
#include "tree.h"

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <unordered_map>
#include <vector>

static_assert(sizeof(bh::Particle) == 64, "Particle layout");
static_assert(offsetof(bh::Particle, mass) == 8 && offsetof(bh::Particle, pos) == 16 && offsetof(bh::Particle, vel) == 40,
              "Particle offsets");
static_assert(sizeof(bh::WalkNode) == 56, "WalkNode layout");

namespace {

struct RefTree {
  bh::FlatTree f;
  std::vector<std::uint64_t> keys;
  std::vector<std::uint32_t> perm;
};

std::vector<bh::Particle> to_vec(const bh::Particle* p, std::uint64_t n) { return std::vector<bh::Particle>(p, p + n); }

// The per-node path's tree and its multipoles (interior nodes and leaves
// other than a lone root leaf).
struct PerNodeTree {
  bh::Tree tree;
  std::unordered_map<bh::Key, bh::Multipole> mp;
  PerNodeTree(std::vector<bh::Particle> p, unsigned max_leaf) : tree(std::move(p), max_leaf) {}
};

// upward_ttg.h:75-118 for the subtree at `key`; returns its multipole.
bh::Multipole upward(PerNodeTree& t, const bh::Key& key) {
  const bh::Node& nd = t.tree.node(key);
  const bh::Vec3 ref = t.tree.box_geometry(key).first;
  bh::RawMoment sum;
  if (nd.is_leaf()) {
    const auto& ps = t.tree.particles();
    for (std::uint32_t i = nd.particle_begin; i < nd.particle_begin + nd.particle_count; ++i)
      sum += bh::raw_moment_of_particle(ps[i].mass, ps[i].pos, ref);
  } else {
    t.tree.for_each_child(key, [&](const bh::Key& ch) { sum += bh::raw_moment_from_multipole(upward(t, ch), ref); });
  }
  const bh::Multipole m = bh::finalize_raw_moment(sum, ref);
  if (!(nd.is_leaf() && key == t.tree.root())) t.mp[key] = m;
  return m;
}

void leaves_in_order(const bh::Tree& tree, const bh::Key& key, std::vector<bh::Key>& out) {
  if (tree.is_leaf(key)) {
    out.push_back(key);
    return;
  }
  tree.for_each_child(key, [&](const bh::Key& ch) { leaves_in_order(tree, ch, out); });
}

}  // namespace

extern "C" {

int ref_domain(const bh::Particle* p, std::uint64_t n, double* out4) {
  bh::Domain d = bh::compute_domain(to_vec(p, n));
  out4[0] = d.center.x;
  out4[1] = d.center.y;
  out4[2] = d.center.z;
  out4[3] = d.half_width;
  return 0;
}

int ref_morton_keys(const bh::Particle* p, std::uint64_t n, const double* dom4, std::uint64_t* keys) {
  bh::Domain d;
  d.center = bh::Vec3{dom4[0], dom4[1], dom4[2]};
  d.half_width = dom4[3];
  for (std::uint64_t i = 0; i < n; ++i) keys[i] = bh::morton_key(d, p[i].pos);
  return 0;
}

std::uint64_t ref_spread_bits3(std::uint32_t v) { return bh::spread_bits3(v); }

int ref_build(const bh::Particle* p, std::uint64_t n, std::uint32_t max_leaf, std::uint32_t group_size, void** out) {
  if (n == 0) return 1;
  auto* t = new RefTree();
  bh::FlatTree& f = t->f;
  f.domain = bh::compute_domain(to_vec(p, n));
  // ascending (key, input index)
  std::vector<std::pair<std::uint64_t, std::uint32_t>> ki(n);
  for (std::uint64_t i = 0; i < n; ++i) ki[i] = {bh::morton_key(f.domain, p[i].pos), static_cast<std::uint32_t>(i)};
  std::sort(ki.begin(), ki.end());
  t->keys.resize(n);
  t->perm.resize(n);
  f.particles.resize(n);
  f.x.resize(n);
  f.y.resize(n);
  f.z.resize(n);
  f.m.resize(n);
  for (std::uint64_t i = 0; i < n; ++i) {
    t->keys[i] = ki[i].first;
    t->perm[i] = ki[i].second;
    const bh::Particle& q = p[ki[i].second];
    f.particles[i] = q;
    f.x[i] = q.pos.x;
    f.y[i] = q.pos.y;
    f.z[i] = q.pos.z;
    f.m[i] = q.mass;
  }
  const std::uint64_t* keys = t->keys.data();
  bh::LeafRule rule{max_leaf};
  constexpr std::uint32_t cutoff = 4096;
  auto grow = [&](std::size_t s) {
    f.nodes.resize(s);
    f.center.resize(s);
    f.iab.resize(s);
  };
  grow(1);
  bh::detail::init_node(f, 0, 0, static_cast<std::uint32_t>(n), f.domain.center, f.domain.half_width, 0, keys, rule);
  if (bh::detail::is_group_root(static_cast<std::uint32_t>(n), f.nodes[0].leaf, std::numeric_limits<std::uint32_t>::max(),
                                group_size))
    f.groups.push_back(0);
  auto is_sroot = [&](bh::NodeIndex i) { return f.nodes[i].particle_count <= cutoff || f.nodes[i].leaf; };
  std::vector<bh::NodeIndex> sroot;
  for (bh::NodeIndex i = 0; i < f.nodes.size(); ++i) {
    if (is_sroot(i)) {
      sroot.push_back(i);
      continue;
    }
    auto first = static_cast<bh::NodeIndex>(f.nodes.size());
    grow(first + 8);
    grow(first + bh::detail::split_node(f, i, first, keys, rule, group_size, f.groups));
  }
  const auto backbone = static_cast<bh::NodeIndex>(f.nodes.size());
  std::vector<bh::NodeIndex> base(sroot.size() + 1);
  base[0] = backbone;
  std::vector<bh::detail::KeyRange> stack;
  for (std::size_t s = 0; s < sroot.size(); ++s) {
    const bh::WalkNode& r = f.nodes[sroot[s]];
    std::uint32_t c = bh::detail::count_subtree(keys, r.particle_begin, r.particle_begin + r.particle_count, r.depth, rule, stack);
    base[s + 1] = base[s] + c - 1;
  }
  grow(base[sroot.size()]);
  std::vector<std::vector<bh::NodeIndex>> sg(sroot.size());
  for (std::size_t s = 0; s < sroot.size(); ++s) bh::detail::build_subtree(f, sroot[s], base[s], keys, rule, group_size, sg[s]);
  for (bh::NodeIndex i = backbone; i-- > 0;) {
    if (!is_sroot(i)) bh::detail::finalize_node(f, i);
  }
  for (auto& g : sg) f.groups.insert(f.groups.end(), g.begin(), g.end());
  *out = t;
  return 0;
}

void ref_tree_free(void* h) { delete static_cast<RefTree*>(h); }

int ref_tree_sizes(const void* h, std::uint64_t* n_nodes, std::uint64_t* n_groups) {
  const RefTree* t = static_cast<const RefTree*>(h);
  *n_nodes = t->f.nodes.size();
  *n_groups = t->f.groups.size();
  return 0;
}

int ref_tree_export(const void* h, std::uint64_t* keys, std::uint32_t* perm, bh::WalkNode* nodes, double* center, double* iab,
                    std::uint32_t* groups) {
  const RefTree* t = static_cast<const RefTree*>(h);
  const bh::FlatTree& f = t->f;
  std::memcpy(keys, t->keys.data(), t->keys.size() * 8);
  std::memcpy(perm, t->perm.data(), t->perm.size() * 4);
  std::memcpy(nodes, f.nodes.data(), f.nodes.size() * sizeof(bh::WalkNode));
  for (std::size_t i = 0; i < f.center.size(); ++i) {
    center[3 * i + 0] = f.center[i].x;
    center[3 * i + 1] = f.center[i].y;
    center[3 * i + 2] = f.center[i].z;
    for (int a = 0; a < 6; ++a) iab[6 * i + a] = f.iab[i][a];
  }
  std::memcpy(groups, f.groups.data(), f.groups.size() * 4);
  return 0;
}

// Per-target root walk with the reference's BBox / mac_accept_bbox /
// accel_from_source: particles of opened leaves go through accel_from_source
// with a zero second moment (brute_force.h's use), accepted nodes with their
// iab when quadrupole is on. acc / n_pp / n_pc are indexed by Morton position.
int ref_target_walk(const void* h, double theta, double eps2, double G, int use_quadrupole, std::uint32_t target_size,
                    double* acc, std::uint32_t* n_pp, std::uint32_t* n_pc) {
  const RefTree* t = static_cast<const RefTree*>(h);
  const bh::FlatTree& f = t->f;
  std::vector<bh::NodeIndex> targets, st;
  for (bh::NodeIndex g : f.groups) {
    st.assign(1, g);
    while (!st.empty()) {
      bh::NodeIndex i = st.back();
      st.pop_back();
      const bh::WalkNode& nd = f.nodes[i];
      if (nd.leaf || nd.particle_count <= target_size) {
        targets.push_back(i);
      } else {
        for (int c = nd.num_children - 1; c >= 0; --c) st.push_back(nd.first_child + c);
      }
    }
  }
  const bh::Sym3 zero{};
  for (bh::NodeIndex ti : targets) {
    const bh::WalkNode& tn = f.nodes[ti];
    bh::BBox box;
    for (std::uint32_t p = tn.particle_begin; p < tn.particle_begin + tn.particle_count; ++p) box.add(f.x[p], f.y[p], f.z[p]);
    std::vector<bh::NodeIndex> mono, quad, leaves;
    st.assign(1, 0);
    std::uint32_t npp = 0;
    while (!st.empty()) {
      bh::NodeIndex s = st.back();
      st.pop_back();
      const bh::WalkNode& sn = f.nodes[s];
      double edge = 2.0 * sn.half_width;
      if (bh::mac_accept_bbox(box, sn.com, edge, theta)) {
        if (use_quadrupole) {
          quad.push_back(s);
        } else {
          mono.push_back(s);
          ++npp;
        }
      } else if (sn.leaf) {
        leaves.push_back(s);
        npp += sn.particle_count;
      } else {
        for (std::uint8_t c = 0; c < sn.num_children; ++c) st.push_back(sn.first_child + c);
      }
    }
    for (std::uint32_t p = tn.particle_begin; p < tn.particle_begin + tn.particle_count; ++p) {
      bh::Vec3 tp{f.x[p], f.y[p], f.z[p]};
      bh::Vec3 a{};
      for (bh::NodeIndex s : leaves) {
        const bh::WalkNode& sn = f.nodes[s];
        for (std::uint32_t q = sn.particle_begin; q < sn.particle_begin + sn.particle_count; ++q) {
          if (q == p) continue;  // the self term contributes 0 (eps2 > 0) or is masked (eps2 == 0)
          bh::Vec3 sp{f.x[q], f.y[q], f.z[q]};
          if (eps2 == 0.0 && sp.x == tp.x && sp.y == tp.y && sp.z == tp.z) continue;
          a += bh::accel_from_source(tp, f.m[q], sp, zero, eps2, false, 1.0);
        }
      }
      for (bh::NodeIndex s : mono) a += bh::accel_from_source(tp, f.nodes[s].mass, f.nodes[s].com, zero, eps2, false, 1.0);
      for (bh::NodeIndex s : quad) a += bh::accel_from_source(tp, f.nodes[s].mass, f.nodes[s].com, f.iab[s], eps2, true, 1.0);
      acc[3 * p + 0] = G * a.x;
      acc[3 * p + 1] = G * a.y;
      acc[3 * p + 2] = G * a.z;
      if (n_pp) n_pp[p] = npp;
      if (n_pc) n_pc[p] = static_cast<std::uint32_t>(quad.size());
    }
  }
  return 0;
}

int ref_brute_force(const bh::Particle* p, std::uint64_t n, double eps2, double G, double* acc) {
  std::vector<bh::Vec3> a = bh::brute_force_accel(to_vec(p, n), eps2, G);
  for (std::uint64_t i = 0; i < n; ++i) {
    acc[3 * i + 0] = a[i].x;
    acc[3 * i + 1] = a[i].y;
    acc[3 * i + 2] = a[i].z;
  }
  return 0;
}

void ref_accel_from_source(const double* tp, double mass, const double* com, const double* iab6, double eps2, int quad, double G,
                           double* out3) {
  bh::Sym3 iab{};
  for (int i = 0; i < 6; ++i) iab[i] = iab6[i];
  bh::Vec3 a = bh::accel_from_source(bh::Vec3{tp[0], tp[1], tp[2]}, mass, bh::Vec3{com[0], com[1], com[2]}, iab, eps2, quad != 0, G);
  out3[0] = a.x;
  out3[1] = a.y;
  out3[2] = a.z;
}

void ref_kick_drift(bh::Particle* p, std::uint64_t n, const double* acc, double kick_dt, double drift_dt) {
  for (std::uint64_t i = 0; i < n; ++i) {
    bh::kick_one(p[i], bh::Vec3{acc[3 * i], acc[3 * i + 1], acc[3 * i + 2]}, kick_dt);
    if (drift_dt != 0.0) bh::drift_one(p[i], drift_dt);
  }
}

// ---- the per-node path (tree.h / tree.cc, traversal_ttg.h:156-196)

int ref_pn_build(const bh::Particle* p, std::uint64_t n, unsigned max_leaf, void** out) {
  if (n == 0) return 1;
  auto* t = new PerNodeTree(to_vec(p, n), max_leaf);
  upward(*t, t->tree.root());
  *out = t;
  return 0;
}

void ref_pn_free(void* h) { delete static_cast<PerNodeTree*>(h); }

std::uint64_t ref_pn_num_leaves(const void* h) {
  const PerNodeTree* t = static_cast<const PerNodeTree*>(h);
  std::uint64_t k = 0;
  for (const auto& kv : t->tree.nodes()) k += kv.second.is_leaf() ? 1 : 0;
  return k;
}

// Leaves in octant (Morton) order: level, translation x/y/z, particle count,
// and the ids of each leaf's particles, ascending, concatenated.
int ref_pn_leaves(const void* h, std::uint32_t* level, std::int32_t* txyz, std::uint32_t* count, std::uint32_t* ids) {
  const PerNodeTree* t = static_cast<const PerNodeTree*>(h);
  std::vector<bh::Key> lv;
  leaves_in_order(t->tree, t->tree.root(), lv);
  const auto& ps = t->tree.particles();
  std::uint64_t pos = 0;
  for (std::size_t k = 0; k < lv.size(); ++k) {
    const bh::Node& nd = t->tree.node(lv[k]);
    level[k] = lv[k].level();
    for (int a = 0; a < 3; ++a) txyz[3 * k + a] = lv[k].translation()[a];
    count[k] = nd.particle_count;
    std::vector<std::uint32_t> v;
    for (std::uint32_t i = nd.particle_begin; i < nd.particle_begin + nd.particle_count; ++i) v.push_back(ps[i].id);
    std::sort(v.begin(), v.end());
    std::memcpy(ids + pos, v.data(), v.size() * 4);
    pos += v.size();
  }
  return 0;
}

// Every leaf a target: the dispatch recursion from the root with the
// target's particle box, a source other than the target tested by the MAC
// (accepted: its multipole), else a leaf source summed directly skipping
// the particle itself, else its children. acc (G applied by
// accel_from_source), n_pp (particles of the leaves summed, the target's
// own included) and n_pc (accepted sources) are indexed by particle id
// (ids must be 0..n-1).
int ref_pn_pass(const void* h, double theta, double eps2, double G, int use_quadrupole, double* acc, std::uint32_t* n_pp,
                std::uint32_t* n_pc) {
  const PerNodeTree* t = static_cast<const PerNodeTree*>(h);
  const bh::Tree& tree = t->tree;
  const auto& ps = tree.particles();
  const std::uint64_t n = ps.size();
  for (const auto& q : ps)
    if (q.id >= n) return 1;
  std::vector<bh::Key> lv;
  leaves_in_order(tree, tree.root(), lv);
  for (const bh::Key& tk : lv) {
    const bh::Node& tn = tree.node(tk);
    bh::BBox box;
    for (std::uint32_t i = tn.particle_begin; i < tn.particle_begin + tn.particle_count; ++i)
      box.add(ps[i].pos.x, ps[i].pos.y, ps[i].pos.z);
    for (std::uint32_t i = tn.particle_begin; i < tn.particle_begin + tn.particle_count; ++i) {
      const bh::Particle& tp = ps[i];
      std::uint32_t pp = 0, pc = 0;
      auto visit = [&](auto&& self, const bh::Key& s) -> bh::Vec3 {
        if (!(s == tk)) {
          const bh::Multipole& m = t->mp.at(s);
          if (bh::mac_accept_bbox(box, m.com, 2.0 * tree.box_geometry(s).second, theta)) {
            ++pc;
            return bh::accel_from_source(tp.pos, m.mass, m.com, m.iab, eps2, use_quadrupole != 0, G);
          }
        }
        bh::Vec3 a{};
        if (tree.is_leaf(s)) {
          const bh::Node& sn = tree.node(s);
          for (std::uint32_t j = sn.particle_begin; j < sn.particle_begin + sn.particle_count; ++j) {
            ++pp;
            if (ps[j].id == tp.id) continue;
            a += bh::accel_from_source(tp.pos, ps[j].mass, ps[j].pos, bh::Sym3{}, eps2, false, G);
          }
          return a;
        }
        tree.for_each_child(s, [&](const bh::Key& ch) { a += self(self, ch); });
        return a;
      };
      const bh::Vec3 a = visit(visit, tree.root());
      acc[3 * tp.id + 0] = a.x;
      acc[3 * tp.id + 1] = a.y;
      acc[3 * tp.id + 2] = a.z;
      n_pp[tp.id] = pp;
      n_pc[tp.id] = pc;
    }
  }
  return 0;
}

double ref_kinetic_energy(const bh::Particle* p, std::uint64_t n) { return bh::kinetic_energy(to_vec(p, n)); }
double ref_potential_energy(const bh::Particle* p, std::uint64_t n, double G) { return bh::potential_energy(to_vec(p, n), G); }

}  // extern "C"
