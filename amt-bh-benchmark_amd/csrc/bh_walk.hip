// The grouped force walk on CDNA4: one wavefront per walk group.
//
// Reference semantics:
//   targets           grouped_ttg.h:253 (topmost nodes with leaf || count <= target_size)
//   MAC               multipole.h:130-153: accept iff theta^2 * near2(com) > (2*hw)^2,
//                     near2 against the FP64 bounding box of the target's particles
//   lists             grouped_ttg.h:223-249: accepted nodes -> quadrupole cells (or
//                     monopoles without quadrupole), opened leaves -> their particles
//                     (the target's own included), opened interior nodes -> children
//   P2P               grouped_ttg.h:124-163: a += m d (r^2 + eps^2)^(-3/2), r = 0 masked
//                     when eps2 == 0
//   P2C               multipole.h:171-200: monopole + traceless quadrupole
//   kick/drift        integrate.h:15-16 via integrate_ttg.h:74-77, fused
//
// Work is split into walk units, one wavefront each: walk groups (topmost
// nodes with leaf || count <= max(tsm, 128)), with runs of sibling groups
// that are single targets packed into units of up to 128 particles. A
// group that is itself a target (and not packed) walks as one target.
// Otherwise the wavefront first walks the tree for the unit's particle box
// with the reference's three-way decision at a group node
// (grouped_ttg.h:259-276): accepted (theta^2 near2 > edge^2) -> the shared
// list, rejected for every box inside (theta^2 far2 <= edge^2) -> opened
// for all, otherwise deferred. The shared list is evaluated for all the
// unit's particles at once. Then each target of the unit walks only the
// deferred nodes. This gives every
// target exactly the interaction list of its own walk from the root: near2
// and far2 are monotone in the box (also in floating point) and near2 <=
// far2, so a group's accept / reject-for-all is the decision of every
// target box inside, and a target opens whatever it does not accept
// (grouped_ttg.h:42-51).
//
// Per iteration a wavefront pops up to 64 nodes of the top frame of its
// LDS stack (frame d holds child ranges d levels below the walk's start),
// tests them with 64 lanes in FP64 - bit-identical to the reference's
// decisions - and compacts the outcomes with ballots: children ranges into
// frame d+1, accepted cells into the P2C ring, opened leaves' particles
// into the P2P ring. Ring entries are staged as (source - box centre) in
// FP64, then rounded to the walk precision, so FP32 arithmetic never sees
// absolute coordinates. Whenever a ring holds a batch it is evaluated. For a
// target with T particles, lane = (particle p, slice s): each of the
// S = 64 / T slices takes every S-th entry, so small leaves still use the
// whole wavefront. FP32 batch sums are accumulated per lane in FP64;
// slices, the group's share and the target's own list are combined in
// FP64.
//
// No MFMA: the only GEMM-shaped part, the quadrupole's Q.d, is a K = 4
// product; on gfx950 the FP32 MFMA does not issue beside the VALU, and a
// bf16-split form would need more registers than four waves per SIMD leave.
#include "bh_internal.h"

namespace bhk {

namespace {

constexpr int kBatch = 48;  // entries evaluated per batch (at most)
constexpr int kBuf = 112;   // staging capacity: kBatch - 1 left over + 64 staged at once
constexpr int kGroupChunks = kGroupCap / kWave;  // a group's particle slots per lane
static_assert(kBatch == 48, "the group layouts assume 48-entry batches");

template <typename Real>
struct Vec4T;
template <>
struct Vec4T<float> {
  using T = float4;
};
template <>
struct Vec4T<double> {
  using T = double4;
};
template <typename Real>
struct Vec2T;
template <>
struct Vec2T<float> {
  using T = float2;
};
template <>
struct Vec2T<double> {
  using T = double2;
};

// The P2C ring: cell i is {x, y, z, m} a[i], {Qxx, Qyy, Qzz, Qxy} b[i],
// {Qxz, Qyz} c[i] (40 bytes in FP32).
template <typename Real>
struct PcRing {
  typename Vec4T<Real>::T a[kBuf];
  typename Vec4T<Real>::T b[kBuf];
  typename Vec2T<Real>::T c[kBuf];
};

// The walk's stack: frame d holds the child ranges d levels below the
// walk's start. Frames [0, F) live in LDS; with kDeep, frames [F, kLevels)
// live in a per-group global scratch area (trees deeper than F levels).
template <typename Real, int F, bool kDeep>
struct WaveLds {
  static constexpr int kFrames = F;
  static constexpr bool kDeepFrames = kDeep;
  using V4 = typename Vec4T<Real>::T;
  uint32_t fr[F][kWave];  // child ranges: (first << 3) | (count - 1)
  uint32_t fr_n[kLevels];
  uint32_t fr_cur[kLevels];
  uint32_t own[kWave];          // lane -> owning range, via prefix max
  V4 pp[kBuf];                  // x, y, z (relative to the target centre), m
  PcRing<Real> pc;
};

__device__ inline float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline double rsq(double x) { return rsqrt(x); }
__device__ inline float fmaR(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline double fmaR(double a, double b, double c) { return fma(a, b, c); }

// Per-lane acceleration accumulator. The FP32 walk sums each batch in plain
// FP32 and adds the batch sums into FP64 accumulators: on CDNA4 an FP64 add
// issues at the FP32 rate, so this takes 6 instructions per batch sum where
// Kahan-compensated FP32 took 12, in the same 6 registers, with a smaller
// error. The FP64 walk adds in FP64 throughout.
template <typename Real>
struct Acc {
  double x = 0.0, y = 0.0, z = 0.0;
  __device__ inline void add(Real bx, Real by, Real bz) {
    x += static_cast<double>(bx);
    y += static_cast<double>(by);
    z += static_cast<double>(bz);
  }
  __device__ inline double sx() const { return x; }
  __device__ inline double sy() const { return y; }
  __device__ inline double sz() const { return z; }
};

// Wave64 inclusive scans on DPP (row shifts within 16-lane rows, then the
// row broadcasts of lanes 15 and 31), no LDS round trips.
__device__ inline uint32_t wave_incl_add(uint32_t v) {
  int x = static_cast<int>(v);
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false);  // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false);  // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return static_cast<uint32_t>(x);
}
// Inclusive max-scan of non-negative values.
__device__ inline uint32_t wave_incl_max(uint32_t v) {
  int x = static_cast<int>(v);
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false));
  x = max(x, __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false));
  return static_cast<uint32_t>(x);
}

__device__ inline uint64_t lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// A wave-uniform double, held in scalar registers.
__device__ inline double uniform_d(double v) {
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
  return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// A walk box (FP64): lo x, y, z, hi x, y, z.
struct Box6 {
  double v[6];
};

template <typename Real>
__device__ inline typename Vec4T<Real>::T make_entry(double x, double y, double z, double m) {
  using V4 = typename Vec4T<Real>::T;
  V4 v;
  v.x = static_cast<Real>(x);
  v.y = static_cast<Real>(y);
  v.z = static_cast<Real>(z);
  v.w = static_cast<Real>(m);
  return v;
}

// Padding entry: zero mass far away, contributes exactly 0 without NaNs.
template <typename Real>
__device__ inline typename Vec4T<Real>::T pad_entry() {
  return make_entry<Real>(1e18, 1e18, 1e18, 0.0);
}

// P2P over this lane's C consecutive entries (grouped_ttg.h:137-153),
// added to the plain partial sums bx, by, bz.
template <typename Real, bool kMask>
__device__ inline void eval_pp_part(const typename Vec4T<Real>::T* e, int C, Real tx, Real ty, Real tz, Real eps2,
                                    Real& bx, Real& by, Real& bz) {
  using V4 = typename Vec4T<Real>::T;
#pragma unroll 4
  for (int k = 0; k < C; ++k) {
    const V4 q = e[k];
    const Real dx = q.x - tx, dy = q.y - ty, dz = q.z - tz;
    const Real r2 = fmaR(dx, dx, fmaR(dy, dy, fmaR(dz, dz, eps2)));
    Real inv = rsq(r2);
    if (kMask) inv = r2 > Real(0) ? inv : Real(0);
    const Real s3 = q.w * inv * inv * inv;
    bx = fmaR(s3, dx, bx);
    by = fmaR(s3, dy, by);
    bz = fmaR(s3, dz, bz);
  }
}

// The same, added to the lane's accumulator.
template <typename Real, bool kMask>
__device__ inline void eval_pp(const typename Vec4T<Real>::T* e, int C, Real tx, Real ty, Real tz, Real eps2,
                               Acc<Real>& acc) {
  Real bx = 0, by = 0, bz = 0;
  eval_pp_part<Real, kMask>(e, C, tx, ty, tz, eps2, bx, by, bz);
  acc.add(bx, by, bz);
}

// Monopole + traceless quadrupole of one cell (multipole.h:171-200):
// a = M d / r^3 - Q.d / r^5 + 2.5 (d.Q.d) d / r^7, evaluated as
// r^-5 (d (M r^2 + 2.5 (d.Q.d) / r^2) - Q.d) with the softened r^2.
template <typename V4, typename V2, typename Real>
__device__ inline void p2c_term(const V4& e0, const V4& e1, const V2& e2, Real tx, Real ty, Real tz, Real eps2,
                                Real& bx, Real& by, Real& bz) {
  const Real dx = e0.x - tx, dy = e0.y - ty, dz = e0.z - tz;
  const Real r2 = fmaR(dx, dx, fmaR(dy, dy, fmaR(dz, dz, eps2)));
  const Real inv = rsq(r2);
  const Real inv2 = inv * inv;
  const Real inv5 = (inv * inv2) * inv2;
  // Q.d with Q = {xx, yy, zz, xy, xz, yz} = {e1.x, e1.y, e1.z, e1.w, e2.x, e2.y}
  const Real qx = fmaR(e1.x, dx, fmaR(e1.w, dy, e2.x * dz));
  const Real qy = fmaR(e1.w, dx, fmaR(e1.y, dy, e2.y * dz));
  const Real qz = fmaR(e2.x, dx, fmaR(e2.y, dy, e1.z * dz));
  const Real dqd = fmaR(dx, qx, fmaR(dy, qy, dz * qz));
  const Real g = fmaR(Real(2.5) * dqd, inv2, e0.w * r2);
  bx = fmaR(inv5, fmaR(g, dx, -qx), bx);
  by = fmaR(inv5, fmaR(g, dy, -qy), by);
  bz = fmaR(inv5, fmaR(g, dz, -qz), bz);
}

// P2C over this lane's C consecutive cells from j0, added to the plain
// partial sums bx, by, bz.
template <typename Real>
__device__ inline void eval_pc_part(const PcRing<Real>& r, int j0, int C, Real tx, Real ty, Real tz, Real eps2, Real& bx,
                                    Real& by, Real& bz) {
#pragma unroll 2
  for (int k = 0; k < C; ++k) p2c_term(r.a[j0 + k], r.b[j0 + k], r.c[j0 + k], tx, ty, tz, eps2, bx, by, bz);
}

// The same, added to the lane's accumulator.
template <typename Real>
__device__ inline void eval_pc(const PcRing<Real>& r, int j0, int C, Real tx, Real ty, Real tz, Real eps2,
                               Acc<Real>& acc) {
  Real bx = 0, by = 0, bz = 0;
  eval_pc_part<Real>(r, j0, C, tx, ty, tz, eps2, bx, by, bz);
  acc.add(bx, by, bz);
}

// Group-phase P2P / P2C: every lane holds NC (<= 2) particles of the group
// and takes the C entries at e (the same entries for both).
template <typename Real, bool kMask, int NC>
__device__ inline void eval_pp_group(const typename Vec4T<Real>::T* e, int C, const Real* gx, const Real* gy,
                                     const Real* gz, Real eps2, Acc<Real>* acc) {
  using V4 = typename Vec4T<Real>::T;
  Real bx[NC], by[NC], bz[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) bx[k] = by[k] = bz[k] = 0;
#pragma unroll 2
  for (int j = 0; j < C; ++j) {
    const V4 q = e[j];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const Real dx = q.x - gx[k], dy = q.y - gy[k], dz = q.z - gz[k];
      const Real r2 = fmaR(dx, dx, fmaR(dy, dy, fmaR(dz, dz, eps2)));
      Real inv = rsq(r2);
      if (kMask) inv = r2 > Real(0) ? inv : Real(0);
      const Real s3 = q.w * inv * inv * inv;
      bx[k] = fmaR(s3, dx, bx[k]);
      by[k] = fmaR(s3, dy, by[k]);
      bz[k] = fmaR(s3, dz, bz[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k].add(bx[k], by[k], bz[k]);
}

template <typename Real, int NC>
__device__ inline void eval_pc_group(const PcRing<Real>& r, int j0, int C, const Real* gx, const Real* gy,
                                     const Real* gz, Real eps2, Acc<Real>* acc) {
  using V4 = typename Vec4T<Real>::T;
  using V2 = typename Vec2T<Real>::T;
  Real bx[NC], by[NC], bz[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) bx[k] = by[k] = bz[k] = 0;
  for (int k2 = 0; k2 < C; ++k2) {
    const V4 e0 = r.a[j0 + k2];
    const V4 e1 = r.b[j0 + k2];
    const V2 e2 = r.c[j0 + k2];
#pragma unroll
    for (int k = 0; k < NC; ++k) p2c_term(e0, e1, e2, gx[k], gy[k], gz[k], eps2, bx[k], by[k], bz[k]);
  }
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k].add(bx[k], by[k], bz[k]);
}

// Evaluator of a target's particles: lane = (particle p, slice s), batches
// of B = S * C entries, slice s takes entries [s*C, (s+1)*C). Batches are
// summed in plain FP32 and folded into the lane's FP64 accumulator once it
// has summed kFoldEntries entries since the last fold (two full
// single-slice batches; small targets have many slices and short per-lane
// batches, so they fold after more batches).
template <typename Real, bool kMask>
struct TargetEval {
  using V4 = typename Vec4T<Real>::T;
  static constexpr int kFoldEntries = 2 * kBatch;
  Real tx, ty, tz, eps2;
  int S, C, s;  // slices, entries per slice and batch, this lane's slice
  bool active;
  uint32_t B;
  int fold_every = 1;  // batches between folds: max(1, kFoldEntries / C)
  int pending = 0;
  Real px = 0, py = 0, pz = 0;
  Acc<Real> acc;
  __device__ inline void fold() {
    acc.add(px, py, pz);
    px = py = pz = 0;
    pending = 0;
  }
  // The last, partial batch of n entries (padded to B): each slice takes
  // ceil(n / S) entries; restore() returns to full batches.
  __device__ inline void last_batch(uint32_t n) { C = (static_cast<int>(n) + S - 1) / S; }
  __device__ inline void restore() { C = static_cast<int>(B) / S; }
  __device__ inline void pp(const V4* ring) {
    if (active) eval_pp_part<Real, kMask>(&ring[s * C], C, tx, ty, tz, eps2, px, py, pz);
    if (++pending >= fold_every) fold();
  }
  __device__ inline void pc(const PcRing<Real>& ring) {
    if (active) eval_pc_part<Real>(ring, s * C, C, tx, ty, tz, eps2, px, py, pz);
    if (++pending >= fold_every) fold();
  }
};

// Evaluator of a whole group (G <= 128 particles, batches of kBatch
// entries). Each lane has two particle slots; slot k of lane l holds
// particle base_k + l % P_k and takes slice l / P_k of S_k slices (C_k =
// kBatch / S_k consecutive entries of each batch). Layouts, by G:
//   G <= 42      two slots, P = ceil(G / 2), S = the largest divisor of
//                kBatch with P S <= 64 (one fused loop)
//   43 .. 53     slot 0: particles 0..31 in 2 slices; slot 1: the r = G - 32
//                others, S_1 = the largest divisor of kBatch with r S_1 <= 64
//   54 .. 64     one slot, P = 64, S = 1
//   65 .. 96     slot 0: particles 0..63, every entry; slot 1: the r = G - 64
//                others, S_1 = the largest divisor of kBatch with r S_1 <= 64
//   97 .. 128    two slots, P = 64, S = 1 (one fused loop)
// Slices are summed in FP64 at the end.
template <typename Real, bool kMask>
struct GroupEval {
  using V4 = typename Vec4T<Real>::T;
  Real gx[kGroupChunks], gy[kGroupChunks], gz[kGroupChunks], eps2;
  int nc;          // slots in the fused loop (1 or 2)
  int c0, o0;      // fused loop: entries [o0, o0 + c0) of each batch
  int c1, o1;      // tail slot 1 (c1 > 0): entries [o1, o1 + c1)
  uint32_t sl0 = 0, sl1 = 0;  // this lane's slices
  bool tail_on;    // this lane's tail slot holds a particle
  uint32_t base[kGroupChunks], rows[kGroupChunks], slices[kGroupChunks];
  uint32_t B;
  Acc<Real> acc[kGroupChunks];

  __device__ inline void setup(uint32_t G, int lane) {
    const uint32_t l = static_cast<uint32_t>(lane);
    c1 = 0;
    o1 = 0;
    tail_on = false;
    constexpr uint32_t kDiv[9] = {48u, 24u, 16u, 12u, 8u, 6u, 4u, 3u, 2u};  // divisors of kBatch
    if (G <= 42) {
      const uint32_t P = (G + 1) / 2;
      uint32_t S = 1;
#pragma unroll
      for (int q = 0; q < 9; ++q)
        if (S == 1 && kDiv[q] * P <= static_cast<uint32_t>(kWave)) S = kDiv[q];
      nc = 2;
      base[0] = 0;
      base[1] = P;
      rows[0] = rows[1] = P;
      slices[0] = slices[1] = S;
    } else if (G <= 53) {
      // particles 0..31 in two slices, the other r = G - 32 in S1 slices
      const uint32_t r = G - 32;
      uint32_t S1 = 1;
#pragma unroll
      for (int q = 0; q < 9; ++q)
        if (S1 == 1 && kDiv[q] * r <= static_cast<uint32_t>(kWave)) S1 = kDiv[q];
      nc = 1;
      base[0] = 0;
      base[1] = 32;
      rows[0] = 32;
      rows[1] = r;
      slices[0] = 2;
      slices[1] = S1;
      c1 = kBatch / static_cast<int>(S1);
      sl1 = l / r < S1 ? l / r : S1 - 1;
      o1 = static_cast<int>(sl1) * c1;
      tail_on = l < r * S1;
    } else if (G <= 64) {
      nc = 1;
      base[0] = base[1] = 0;
      rows[0] = rows[1] = kWave;
      slices[0] = slices[1] = 1;
    } else if (G <= 96) {
      const uint32_t r = G - kWave;
      uint32_t S1 = 1;
#pragma unroll
      for (int q = 0; q < 9; ++q)
        if (S1 == 1 && kDiv[q] * r <= static_cast<uint32_t>(kWave)) S1 = kDiv[q];
      nc = 1;
      base[0] = 0;
      base[1] = kWave;
      rows[0] = kWave;
      rows[1] = r;
      slices[0] = 1;
      slices[1] = S1;
      c1 = kBatch / static_cast<int>(S1);
      sl1 = l / r < S1 ? l / r : S1 - 1;
      o1 = static_cast<int>(sl1) * c1;
      tail_on = l < r * S1;
    } else {
      nc = 2;
      base[0] = 0;
      base[1] = kWave;
      rows[0] = rows[1] = kWave;
      slices[0] = slices[1] = 1;
    }
    c0 = kBatch / static_cast<int>(slices[0]);
    sl0 = l / rows[0] < slices[0] ? l / rows[0] : slices[0] - 1;
    o0 = static_cast<int>(sl0) * c0;
    B = kBatch;
  }
  // The last, partial batch of n entries (padded to B): each slice takes
  // ceil(n / S) entries; restore() returns to full batches.
  __device__ inline void last_batch(uint32_t n) {
    c0 = static_cast<int>((n + slices[0] - 1) / slices[0]);
    o0 = static_cast<int>(sl0) * c0;
    if (c1 > 0) {
      c1 = static_cast<int>((n + slices[1] - 1) / slices[1]);
      o1 = static_cast<int>(sl1) * c1;
    }
  }
  __device__ inline void restore() {
    c0 = kBatch / static_cast<int>(slices[0]);
    o0 = static_cast<int>(sl0) * c0;
    if (c1 > 0) {
      c1 = kBatch / static_cast<int>(slices[1]);
      o1 = static_cast<int>(sl1) * c1;
    }
  }
  // the particle (index within the group, clamped to it) of slot k
  __device__ inline uint32_t particle(int k, int lane, uint32_t G) const {
    const uint32_t i = base[k] + static_cast<uint32_t>(lane) % rows[k];
    return i < G ? i : G - 1;
  }
  __device__ inline void pp(const V4* ring) {
    if (nc == 2) {
      eval_pp_group<Real, kMask, 2>(ring + o0, c0, gx, gy, gz, eps2, acc);
    } else {
      eval_pp_group<Real, kMask, 1>(ring + o0, c0, gx, gy, gz, eps2, acc);
      if (c1 > 0 && tail_on) eval_pp<Real, kMask>(ring + o1, c1, gx[1], gy[1], gz[1], eps2, acc[1]);
    }
  }
  __device__ inline void pc(const PcRing<Real>& ring) {
    if (nc == 2) {
      eval_pc_group<Real, 2>(ring, o0, c0, gx, gy, gz, eps2, acc);
    } else {
      eval_pc_group<Real, 1>(ring, o0, c0, gx, gy, gz, eps2, acc);
      if (c1 > 0 && tail_on) eval_pc<Real>(ring, o1, c1, gx[1], gy[1], gz[1], eps2, acc[1]);
    }
  }
};

// The depth-first walk of one box (a target's or a group's particles) over
// the node ranges `start` (global, packed (first << 3) | (count - 1)), or
// from the system's root when start is null. Accepted nodes and the
// particles of opened leaves are staged in the rings relative to (cx, cy,
// cz) and handed to ev in batches of ev.B; the last partial batch is padded.
//
// kGroup: a group's walk (grouped_ttg.h:259-276 at the group node): a node
// that is neither accepted nor rejected for every box inside
// (theta^2 * far2 > edge^2) is deferred: appended to dl, for its targets.
// Returns false when more than kDefer nodes would be deferred.
template <typename Real, bool kMask, bool kGroup, typename Eval, typename Lds>
__device__ bool traverse(const WalkArgs& a, Lds& L, uint32_t* gfr, const Box6& box, double cx, double cy, double cz,
                         uint32_t sys_id, const uint32_t* start, uint32_t n_start, bool use_quad, double theta2,
                         Eval& ev, uint32_t& n_pp, uint32_t& n_pc, uint32_t* dl, uint32_t& n_def) {
  constexpr int F = Lds::kFrames;
  constexpr int kStack = Lds::kDeepFrames ? kLevels : F;
  using V4 = typename Vec4T<Real>::T;
  const int lane = __lane_id();
  const uint64_t lt = lanemask_lt();
  const uint32_t last_node = a.n_nodes - 1, last_part = a.n_parts - 1;
  const uint32_t B = ev.B;
  const double blo0 = uniform_d(box.v[0]), blo1 = uniform_d(box.v[1]), blo2 = uniform_d(box.v[2]);
  const double bhi0 = uniform_d(box.v[3]), bhi1 = uniform_d(box.v[4]), bhi2 = uniform_d(box.v[5]);
  cx = uniform_d(cx);
  cy = uniform_d(cy);
  cz = uniform_d(cz);
  theta2 = uniform_d(theta2);
  uint32_t pp_cnt = 0, pc_cnt = 0;
  bool ok = true;

  // Evaluate the first B staged entries, then move the rest (< 64: at most
  // B - 1 + 64 were staged) to the front; callers repeat while >= B remain.
  auto flush_pp = [&]() {
    __builtin_amdgcn_wave_barrier();
    ev.pp(L.pp);
    __builtin_amdgcn_wave_barrier();
    const uint32_t rest = pp_cnt - B;
    V4 v;
    if (static_cast<uint32_t>(lane) < rest) v = L.pp[B + lane];
    __builtin_amdgcn_wave_barrier();
    if (static_cast<uint32_t>(lane) < rest) L.pp[lane] = v;
    pp_cnt = rest;
    __builtin_amdgcn_wave_barrier();
  };
  auto flush_pc = [&]() {
    __builtin_amdgcn_wave_barrier();
    ev.pc(L.pc);
    __builtin_amdgcn_wave_barrier();
    const uint32_t rest = pc_cnt - B;
    V4 v0, v1;
    typename Vec2T<Real>::T v2;
    if (static_cast<uint32_t>(lane) < rest) {
      v0 = L.pc.a[B + lane];
      v1 = L.pc.b[B + lane];
      v2 = L.pc.c[B + lane];
    }
    __builtin_amdgcn_wave_barrier();
    if (static_cast<uint32_t>(lane) < rest) {
      L.pc.a[lane] = v0;
      L.pc.b[lane] = v1;
      L.pc.c[lane] = v2;
    }
    pc_cnt = rest;
    __builtin_amdgcn_wave_barrier();
  };

  if (lane == 0) {
    L.fr[0][0] = sys_id << 3;  // the system's root: node sys_id, one node
    L.fr_n[0] = start ? n_start : 1u;
    L.fr_cur[0] = 0;
  }
  __builtin_amdgcn_wave_barrier();
  int depth = 0;
  while (depth >= 0) {
    const uint32_t nr = uniform(L.fr_n[depth]);
    const uint32_t cur = uniform(L.fr_cur[depth]);
    if (cur >= nr) {
      --depth;
      continue;
    }
    // Pop whole child ranges, up to 64 nodes. Range j covers slots
    // [excl_j, excl_j + count_j); its index j, placed at its first slot,
    // is non-decreasing over the slots, so a prefix max gives every slot its
    // range: node = first_j + (slot - excl_j). (The ranges themselves need
    // not be ascending: children of deferred nodes come in walk order.)
    const uint32_t rr = cur + lane;
    const bool rv = rr < nr;
    uint32_t node, total;
    if (depth == 0 && start) {
      // a deferred list: single nodes in walk order, one per lane
      node = rv ? (start[rr] >> 3) : 0u;
      total = nr - cur < kWave ? nr - cur : kWave;
      if (lane == 0) L.fr_cur[0] = cur + total;
      __builtin_amdgcn_wave_barrier();
    } else {
      uint32_t packed = 0u;
      if constexpr (Lds::kDeepFrames) {
        if (rv) packed = depth < F ? L.fr[depth][rr] : gfr[(depth - F) * kWave + rr];
      } else {
        if (rv) packed = L.fr[depth][rr];
      }
      const uint32_t rcnt = rv ? (packed & 7u) + 1u : 0u;
      const uint32_t rfirst = packed >> 3;
      const uint32_t incl = wave_incl_add(rcnt);
      const bool taken = rv && incl <= kWave;
      const int k = __popcll(__ballot(taken));
      total = uniform(__builtin_amdgcn_readlane(incl, k - 1));
      const uint32_t excl = incl - rcnt;
      L.own[lane] = 0u;
      __builtin_amdgcn_wave_barrier();
      if (taken) L.own[excl] = static_cast<uint32_t>(lane);
      if (lane == 0) L.fr_cur[depth] = cur + k;
      __builtin_amdgcn_wave_barrier();
      const int j = static_cast<int>(wave_incl_max(L.own[lane]));
      node = static_cast<uint32_t>(__shfl(static_cast<int>(rfirst), j)) +
             (static_cast<uint32_t>(lane) - static_cast<uint32_t>(__shfl(static_cast<int>(excl), j)));
    }
    const bool nv = static_cast<uint32_t>(lane) < total;
    // MAC in FP64, reference operation order (multipole.h:130-160). Every
    // lane loads and tests a record (lanes past `total` lane 0's, a line
    // already being fetched) and the flags below drop the extra lanes: no
    // divergent branch.
    // The record's second half, {first, count, m, Qxx} and {Qyy, Qxy, Qxz,
    // Qyz}, comes with the MAC data, so accepted cells are staged from
    // registers without a second round trip to memory.
    const uint32_t node0 = uniform(node);
    const uint32_t nd = nv ? (node < last_node ? node : last_node) : (node0 < last_node ? node0 : last_node);
    const NodeRec* rp = &a.rec[nd];
    const uint4* rw = reinterpret_cast<const uint4*>(rp);
    const uint4 rw2 = rw[2];
    uint4 rw3 = make_uint4(0, 0, 0, 0);
    if constexpr (sizeof(Real) == 4) rw3 = rw[3];
    const double4 mac = make_double4(rp->cx, rp->cy, rp->cz, rp->edge2);
    bool defer = false;
    const double dx0 = fmax(fmax(blo0 - mac.x, mac.x - bhi0), 0.0);
    const double dx1 = fmax(fmax(blo1 - mac.y, mac.y - bhi1), 0.0);
    const double dx2 = fmax(fmax(blo2 - mac.z, mac.z - bhi2), 0.0);
    const double near2 = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
    const bool accept = theta2 * near2 > mac.w;
    if constexpr (kGroup) {
      // no deferred list (a target's walk through this evaluator): accept or open
      if (!accept && dl != nullptr) {
        const double fx0 = fmax(mac.x - blo0, bhi0 - mac.x);
        const double fx1 = fmax(mac.y - blo1, bhi1 - mac.y);
        const double fx2 = fmax(mac.z - blo2, bhi2 - mac.z);
        const double far2 = fx0 * fx0 + fx1 * fx1 + fx2 * fx2;
        defer = !(theta2 * far2 <= mac.w);
      }
    }
    const bool isleaf = (rw2.y >> 31) != 0;
    const uint32_t lfirst = rw2.x;
    const uint32_t lcnt = rw2.y & 0x7fffffffu;
    const bool cellf = nv && accept;
    const bool deff = nv && defer;
    const bool leaff = nv && !accept && !defer && isleaf;
    // An interior node to open needs a child and a free stack frame. The
    // launch sizes the stack for the tree's depth, so neither check fails on
    // a well-formed tree; if one does, the node is not opened and the step
    // reports an incomplete walk (the reference's missing > 0,
    // integrate_ttg.h:104-106).
    const bool openable = !isleaf && lcnt != 0u && depth + 1 < kStack && static_cast<uint32_t>(depth) + 1 < a.stack_limit;
    const bool openf = nv && !accept && !defer && !isleaf && openable;
    if (__ballot(nv && !accept && !defer && !isleaf && !openable)) {
      if (lane == 0) atomicOr(&a.misc[kMiscIncomplete], 1u);
    }
    if constexpr (kGroup) {
      const uint64_t dm = __ballot(deff);
      if (dm) {
        const uint32_t nd_new = __popcll(dm);
        if (n_def + nd_new > kDefer) {
          ok = false;
          break;
        }
        if (deff) dl[n_def + __popcll(dm & lt)] = nd << 3;
        n_def += nd_new;
      }
    }
    // children of opened interior nodes -> frame depth + 1
    const uint64_t om = __ballot(openf);
    if (om) {
      const uint32_t pos = __popcll(om & lt);
      if constexpr (Lds::kDeepFrames) {
        if (depth + 1 < F) {
          if (openf) L.fr[depth + 1][pos] = (lfirst << 3) | (lcnt - 1u);
        } else {
          if (openf) gfr[(depth + 1 - F) * kWave + pos] = (lfirst << 3) | (lcnt - 1u);
          __threadfence_block();  // read back by other lanes
        }
      } else {
        if (openf) L.fr[depth + 1][pos] = (lfirst << 3) | (lcnt - 1u);
      }
      if (lane == 0) {
        L.fr_n[depth + 1] = __popcll(om);
        L.fr_cur[depth + 1] = 0;
      }
    }
    // accepted nodes
    const uint64_t cm = __ballot(cellf);
    if (cm) {
      const uint32_t nc = __popcll(cm);
      const uint32_t pos = __popcll(cm & lt);
      if (use_quad) {
        if (cellf) {
          const uint32_t slot = pc_cnt + pos;
          if constexpr (sizeof(Real) == 4) {
            const float qxx = __uint_as_float(rw2.w), qyy = __uint_as_float(rw3.x);
            L.pc.a[slot] = make_entry<Real>(mac.x - cx, mac.y - cy, mac.z - cz, __uint_as_float(rw2.z));
            L.pc.b[slot] = make_float4(qxx, qyy, -(qxx + qyy), __uint_as_float(rw3.y));
            L.pc.c[slot] = make_float2(__uint_as_float(rw3.z), __uint_as_float(rw3.w));
          } else {
            const double* q = &a.q64[8 * nd];
            L.pc.a[slot] = make_entry<Real>(mac.x - cx, mac.y - cy, mac.z - cz, q[0]);
            L.pc.b[slot] = make_double4(q[1], q[2], q[3], q[4]);
            L.pc.c[slot] = make_double2(q[5], q[6]);
          }
        }
        pc_cnt += nc;
        n_pc += nc;
        while (pc_cnt >= B) flush_pc();
      } else {
        if (cellf) {
          const double m = sizeof(Real) == 4 ? static_cast<double>(__uint_as_float(rw2.z)) : a.q64[8 * nd];
          L.pp[pp_cnt + pos] = make_entry<Real>(mac.x - cx, mac.y - cy, mac.z - cz, m);
        }
        pp_cnt += nc;
        n_pp += nc;
        while (pp_cnt >= B) flush_pp();
      }
    }
    // particles of opened leaves, mapped to slots like the nodes: the lane
    // of the leaf covering each slot by a prefix max of lane indices
    const uint64_t lm = __ballot(leaff);
    if (lm) {
      const uint32_t lc = leaff ? lcnt : 0u;
      const uint32_t lincl = wave_incl_add(lc);
      const uint32_t lexcl = lincl - lc;
      const uint32_t ltot = uniform(__builtin_amdgcn_readlane(lincl, 63));
      for (uint32_t b = 0; b < ltot; b += kWave) {
        // the leaf holding slot b: the last leaf lane starting at or before it
        const uint64_t before = __ballot(leaff && lexcl <= b);
        const int j0 = 63 - __clzll(static_cast<long long>(before));
        L.own[lane] = static_cast<uint32_t>(j0);
        __builtin_amdgcn_wave_barrier();
        if (leaff && lexcl > b && lexcl < b + kWave) L.own[lexcl - b] = static_cast<uint32_t>(lane);
        __builtin_amdgcn_wave_barrier();
        const uint32_t e = b + lane;
        const int jj = static_cast<int>(wave_incl_max(L.own[lane]));
        uint32_t q = static_cast<uint32_t>(__shfl(static_cast<int>(lfirst), jj)) +
                     (e - static_cast<uint32_t>(__shfl(static_cast<int>(lexcl), jj)));
        q = q < last_part ? q : last_part;
        const uint32_t chunk = ltot - b < kWave ? ltot - b : kWave;
        if (e < ltot) {
          const double4 sp = a.posm[q];
          L.pp[pp_cnt + lane] = make_entry<Real>(sp.x - cx, sp.y - cy, sp.z - cz, sp.w);
        }
        pp_cnt += chunk;
        n_pp += chunk;
        while (pp_cnt >= B) flush_pp();
      }
    }
    __builtin_amdgcn_wave_barrier();
    if (om) ++depth;
  }
  if (!ok) return false;
  // final partial batches (fewer than B entries), padded to B; each slice
  // evaluates only its share of the entries actually staged
  if (pp_cnt) {
    if (static_cast<uint32_t>(lane) >= pp_cnt && static_cast<uint32_t>(lane) < B) L.pp[lane] = pad_entry<Real>();
    ev.last_batch(pp_cnt);
    pp_cnt = B;
    flush_pp();
    ev.restore();
  }
  if (pc_cnt) {
    if (static_cast<uint32_t>(lane) >= pc_cnt && static_cast<uint32_t>(lane) < B) {
      L.pc.a[lane] = pad_entry<Real>();
      L.pc.b[lane] = make_entry<Real>(0.0, 0.0, 0.0, 0.0);
      L.pc.c[lane].x = Real(0);
      L.pc.c[lane].y = Real(0);
    }
    ev.last_batch(pc_cnt);
    pc_cnt = B;
    flush_pc();
    ev.restore();
  }
  __builtin_amdgcn_wave_barrier();
  return true;
}

// The unit's interaction totals, one atomic pair per wavefront.
__device__ inline void add_totals(const WalkArgs& a, unsigned long long pp, unsigned long long pc) {
  if (__lane_id() == 0) {
    atomicAdd(&a.totals[0], pp);
    atomicAdd(&a.totals[1], pc);
  }
}

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// FP64 bounding box of particles [b, b + n) (multipole.h:115-128).
__device__ inline Box6 particle_box(const WalkArgs& a, uint32_t b, uint32_t n) {
  const double inf = __builtin_huge_val();
  double lo0 = inf, lo1 = inf, lo2 = inf, hi0 = -inf, hi1 = -inf, hi2 = -inf;
  for (uint32_t p = __lane_id(); p < n; p += kWave) {
    const double4 q = a.posm[b + p];
    lo0 = fmin(lo0, q.x);
    lo1 = fmin(lo1, q.y);
    lo2 = fmin(lo2, q.z);
    hi0 = fmax(hi0, q.x);
    hi1 = fmax(hi1, q.y);
    hi2 = fmax(hi2, q.z);
  }
  Box6 box;
  box.v[0] = uniform_d(wave_min(lo0));
  box.v[1] = uniform_d(wave_min(lo1));
  box.v[2] = uniform_d(wave_min(lo2));
  box.v[3] = uniform_d(wave_max(hi0));
  box.v[4] = uniform_d(wave_max(hi1));
  box.v[5] = uniform_d(wave_max(hi2));
  return box;
}

// A node's particle box as computed by the upward pass.
__device__ inline Box6 node_box(const WalkArgs& a, uint32_t node) {
  const double* p = a.nbox + 6 * uint64_t(node);
  Box6 box;
#pragma unroll
  for (int k = 0; k < 6; ++k) box.v[k] = p[k];
  return box;
}

// The end of one target particle's walk: its FP64 sums plus the group's
// share (gacc, when given), times G, to the outputs (grouped_ttg.h:63-66,
// the update callback) and the fused kick/drift.
__device__ inline void finish_particle(const WalkArgs& a, uint32_t k3, double ax, double ay, double az,
                                       const double* gacc, double sys_G, double kick_dt, double drift_dt,
                                       uint32_t n_pp, uint32_t n_pc) {
  if (gacc) {
    ax += gacc[3 * uint64_t(k3) + 0];
    ay += gacc[3 * uint64_t(k3) + 1];
    az += gacc[3 * uint64_t(k3) + 2];
  }
  const double gx = sys_G * ax, gy = sys_G * ay, gz = sys_G * az;
  if (a.acc) {
    a.acc[3 * uint64_t(k3) + 0] = gx;
    a.acc[3 * uint64_t(k3) + 1] = gy;
    a.acc[3 * uint64_t(k3) + 2] = gz;
  }
  if (a.n_pp) a.n_pp[k3] = n_pp;
  if (a.n_pc) a.n_pc[k3] = n_pc;
  if (a.parts) {
    // kick_one / drift_one (integrate.h:15-16): v += a*dt; x += v*dt. The
    // particle comes from the unsorted input (src at perm) when given: the
    // whole Morton-ordered record is written here, not by the sort's gather.
    bh_particle pt = a.src ? a.src[a.perm[k3]] : a.parts[k3];
    const double vx = pt.vel[0] + gx * kick_dt;
    const double vy = pt.vel[1] + gy * kick_dt;
    const double vz = pt.vel[2] + gz * kick_dt;
    pt.vel[0] = vx;
    pt.vel[1] = vy;
    pt.vel[2] = vz;
    if (drift_dt != 0.0) {
      pt.pos[0] = pt.pos[0] + vx * drift_dt;
      pt.pos[1] = pt.pos[1] + vy * drift_dt;
      pt.pos[2] = pt.pos[2] + vz * drift_dt;
    }
    a.parts[k3] = pt;
    if (a.carry) {
      // the next step's build reads these instead of the records (x, y, z
      // and m in blocks of n_parts)
      const uint64_t np = a.n_parts;
      a.carry[k3] = pt.pos[0];
      a.carry[np + k3] = pt.pos[1];
      a.carry[2 * np + k3] = pt.pos[2];
      a.carry[3 * np + k3] = pt.mass;
      if (!(isfinite(pt.pos[0]) && isfinite(pt.pos[1]) && isfinite(pt.pos[2]) && isfinite(vx) && isfinite(vy) &&
            isfinite(vz) && isfinite(pt.mass)))
        atomicOr(a.carry_bad, 1u);
    }
  }
}

// Targets of at least kWideMin particles use the group evaluator (its
// layouts cover 33 to kGroupCap particles, two slots per lane; larger
// targets walk in chunks of kGroupCap); smaller ones use lane = (particle,
// slice).
constexpr uint32_t kWideMin = 33;

// One target (node tnode) with the target evaluator (the unit's loop gives
// it targets of at most 32 particles): the walk over `start` (the group's
// deferred nodes) or from the root, for each chunk of <= 64 of its
// particles; then the slices' sums, plus the group's share (gacc, counts
// g_pp / g_pc) when given.
template <typename Real, bool kMask, typename Lds>
__device__ void walk_target(const WalkArgs& a, Lds& L, uint32_t* gfr, uint32_t tnode, const uint32_t* start,
                            uint32_t n_start, const double* gacc, uint32_t g_pp, uint32_t g_pc,
                            unsigned long long& tot_pp, unsigned long long& tot_pc) {
  const int lane = __lane_id();
  const uint32_t tb = uniform(a.begin[tnode]);
  const uint32_t T = uniform(a.count[tnode]);
  const uint32_t sys_id = uniform(a.sid[tb]);
  const SysDev& sp = a.sys[sys_id];
  const bool use_quad = sp.use_quad != 0;
  const double4 tcv = a.ctr[tnode];
  const double tcx = tcv.x, tcy = tcv.y, tcz = tcv.z;
  const Box6 box = node_box(a, tnode);
  const double theta2 = sp.theta2;

  for (uint32_t c0 = 0; c0 < T; c0 += kWave) {
    const uint32_t tc = T - c0 < kWave ? T - c0 : kWave;
    const int S = kWave / static_cast<int>(tc) < kBatch ? kWave / static_cast<int>(tc) : kBatch;
    TargetEval<Real, kMask> ev;
    ev.S = S;
    ev.C = kBatch / S;
    ev.B = static_cast<uint32_t>(S * ev.C);
    ev.fold_every = ev.C >= TargetEval<Real, kMask>::kFoldEntries ? 1 : TargetEval<Real, kMask>::kFoldEntries / ev.C;
    ev.active = lane < static_cast<int>(tc) * S;
    const int p = lane % static_cast<int>(tc);
    ev.s = lane / static_cast<int>(tc);
    ev.eps2 = static_cast<Real>(sp.eps2);
    {
      const double4 q = a.posm[tb + c0 + p];
      ev.tx = static_cast<Real>(q.x - tcx);
      ev.ty = static_cast<Real>(q.y - tcy);
      ev.tz = static_cast<Real>(q.z - tcz);
    }
    uint32_t n_pp = 0, n_pc = 0, n_def = 0;
    traverse<Real, kMask, false>(a, L, gfr, box, tcx, tcy, tcz, sys_id, start, n_start, use_quad, theta2, ev, n_pp,
                                 n_pc, nullptr, n_def);
    n_pp += g_pp;
    n_pc += g_pc;
    ev.fold();
    // combine slices in FP64 (the P2C ring is free now)
    double* red = reinterpret_cast<double*>(L.pc.a);  // 3 x 64 doubles
    if (ev.active) {
      red[3 * lane + 0] = ev.acc.sx();
      red[3 * lane + 1] = ev.acc.sy();
      red[3 * lane + 2] = ev.acc.sz();
    }
    __builtin_amdgcn_wave_barrier();
    if (lane < static_cast<int>(tc)) {
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int k2 = 0; k2 < S; ++k2) {
        const int src = k2 * static_cast<int>(tc) + lane;
        ax += red[3 * src + 0];
        ay += red[3 * src + 1];
        az += red[3 * src + 2];
      }
      finish_particle(a, tb + c0 + lane, ax, ay, az, gacc, sp.G, sp.kick_dt, sp.drift_dt, n_pp, n_pc);
    }
    tot_pp += static_cast<unsigned long long>(n_pp) * tc;
    tot_pc += static_cast<unsigned long long>(n_pc) * tc;
    __builtin_amdgcn_wave_barrier();
  }
}

// One walk unit starting at node g: a walk group (topmost node with leaf ||
// count <= max(tsm, 128)) or a run of sibling single-target groups packed
// into one unit (unit_n). A group that is itself a target walks as one.
// Otherwise the wave walks the tree once for the unit's box - the decisions
// every target inside shares - evaluating the shared list for all the
// unit's particles and keeping the deferred nodes; then each target of the
// unit walks only those.
//
// The unit's work is a sequence of jobs: the group's shared walk (if any),
// then its targets in particle order (a target of more than 128 particles
// as jobs of 128 of them, each walking the target's list with its box). The
// group's walk and the targets of 33 particles or more use the group
// evaluator (its layouts keep the lanes busy for those sizes, and a target
// of 65-128 particles walks once, not once per 64) through one call of its
// traversal; smaller targets use the target evaluator. Deferral is off for
// target jobs (no deferred list): their decisions are accept or open.
template <typename Real, bool kMask, typename Lds>
__device__ void walk_group(const WalkArgs& a, Lds& L, uint32_t slot, uint32_t g) {
  const int lane = __lane_id();
  const uint32_t gb = uniform(a.begin[g]);
  // a unit packed from sibling single-target groups spans several nodes
  const uint32_t un = a.unit_n ? uniform(a.unit_n[g]) : 0u;
  const bool packed = (un >> 31) != 0;
  const uint32_t G = packed ? un & 0x7fffffffu : uniform(a.count[g]);
  const uint32_t sys_id = uniform(a.sid[gb]);
  const SysDev& sp = a.sys[sys_id];
  const bool gleaf = (a.rec[g].count >> 31) != 0;
  uint32_t* gfr = Lds::kDeepFrames ? a.deep_frames + uint64_t(slot) * (kLevels - Lds::kFrames) * kWave : nullptr;
  const bool single = !packed && (gleaf || G <= sp.tsm || G > kGroupCap);
  uint32_t* dl = a.defer + uint64_t(slot) * kDefer;
  uint32_t g_pp = 0, g_pc = 0, n_def = 0;
  bool ok = false;  // the group's shared walk kept all its deferred nodes
  unsigned long long tot_pp = 0, tot_pc = 0;
  bool group_job = !single, single_job = single;
  uint32_t c = 0;     // next 64-particle chunk of the unit to scan for targets
  uint64_t bits = 0;  // targets of the scanned chunk not yet walked
  uint32_t m = 0;     // this lane's mark in the scanned chunk
  uint32_t big = 0, big_T = 0, big_off = 0;  // a target of > kGroupCap particles, walked in chunks
  for (;;) {
    // the next job: the group's walk, a target node, or the next chunk of a
    // large target
    bool is_group = false;
    uint32_t tnode = g, off = 0;
    if (group_job) {
      is_group = true;
      group_job = false;
    } else if (big_off < big_T) {
      tnode = big;
      off = big_off;
    } else if (single_job) {
      single_job = false;
    } else {
      if (single) break;
      while (!bits && c < G) {
        const uint32_t i = c + lane;
        m = i < G ? a.tmark[gb + i] : 0u;
        bits = __ballot(m != 0);
        c += kWave;
      }
      if (!bits) break;
      const int j = __ffsll(static_cast<long long>(bits)) - 1;
      bits &= bits - 1;
      tnode = uniform(__builtin_amdgcn_readlane(m, j)) - 1u;
    }
    const uint32_t* start = ok ? dl : nullptr;
    const uint32_t n_start = ok ? n_def : 0u;
    const double* gacc = ok ? a.gacc : nullptr;
    const uint32_t jpp = ok ? g_pp : 0u, jpc = ok ? g_pc : 0u;
    const uint32_t T = is_group ? 0u : uniform(a.count[tnode]);
    if (!is_group && T < kWideMin) {
      walk_target<Real, kMask>(a, L, gfr, tnode, start, n_start, gacc, jpp, jpc, tot_pp, tot_pc);
      continue;
    }
    if (!is_group && T > kGroupCap && off == 0 && big_T == 0) {
      big = tnode;  // its chunks follow, this one first
      big_T = T;
    }
    if (!is_group && T > kGroupCap) big_off = off + kGroupCap;
    if (!is_group && big_T != 0 && big_off >= big_T) big_T = big_off = 0;
    // group evaluator: the group's shared walk, a target of 33-128 particles
    // or a chunk of up to 128 of a larger target (its box and list, its
    // chunk's particles)
    const uint32_t b0 = is_group ? gb : uniform(a.begin[tnode]) + off;
    const uint32_t size = is_group ? G : (T - off < kGroupCap ? T - off : kGroupCap);
    // a packed unit's box and the origin of its relative coordinates (the
    // box centre); a group's or target's: its node box and centre
    const bool pbox = is_group && packed;
    const Box6 box = pbox ? particle_box(a, gb, G) : node_box(a, tnode);
    const double4 ccv = pbox ? make_double4(0.5 * (box.v[0] + box.v[3]), 0.5 * (box.v[1] + box.v[4]),
                                            0.5 * (box.v[2] + box.v[5]), 0.0)
                             : a.ctr[tnode];
    GroupEval<Real, kMask> ev;
    ev.setup(size, lane);
    ev.eps2 = static_cast<Real>(sp.eps2);
#pragma unroll
    for (int k = 0; k < kGroupChunks; ++k) {
      const double4 q = a.posm[b0 + ev.particle(k, lane, size)];
      ev.gx[k] = static_cast<Real>(q.x - ccv.x);
      ev.gy[k] = static_cast<Real>(q.y - ccv.y);
      ev.gz[k] = static_cast<Real>(q.z - ccv.z);
    }
    uint32_t n_pp = 0, n_pc = 0, j_def = 0;
    const bool walked = traverse<Real, kMask, true>(a, L, gfr, box, ccv.x, ccv.y, ccv.z, sys_id,
                                                    is_group ? nullptr : start, is_group ? 0u : n_start,
                                                    sp.use_quad != 0, sp.theta2, ev, n_pp, n_pc,
                                                    is_group ? dl : nullptr, j_def);
    if (is_group) {
      ok = walked;
      g_pp = n_pp;
      g_pc = n_pc;
      n_def = j_def;
      if (!ok) continue;  // the targets walk from the root
    } else {
      n_pp += jpp;
      n_pc += jpc;
    }
    // per-particle sums: each slot's slices added in FP64 through LDS (the
    // P2C ring is free after the walk); the group's are stored for its
    // targets, a target's finish its particles
    double* red = reinterpret_cast<double*>(L.pc.a);  // 3 x 64 doubles
    const int used = ev.nc == 2 || ev.c1 > 0 ? 2 : 1;
#pragma unroll
    for (int k = 0; k < kGroupChunks; ++k) {
      if (k < used) {
        const uint32_t P = ev.rows[k], S = ev.slices[k];
        red[3 * lane + 0] = ev.acc[k].sx();
        red[3 * lane + 1] = ev.acc[k].sy();
        red[3 * lane + 2] = ev.acc[k].sz();
        __builtin_amdgcn_wave_barrier();
        const uint32_t i = ev.base[k] + static_cast<uint32_t>(lane);
        if (static_cast<uint32_t>(lane) < P && i < size) {
          double ax = 0.0, ay = 0.0, az = 0.0;
          for (uint32_t s2 = 0; s2 < S; ++s2) {
            const uint32_t src = s2 * P + static_cast<uint32_t>(lane);
            ax += red[3 * src + 0];
            ay += red[3 * src + 1];
            az += red[3 * src + 2];
          }
          if (is_group) {
            a.gacc[3 * uint64_t(gb + i) + 0] = ax;
            a.gacc[3 * uint64_t(gb + i) + 1] = ay;
            a.gacc[3 * uint64_t(gb + i) + 2] = az;
          } else {
            finish_particle(a, b0 + i, ax, ay, az, gacc, sp.G, sp.kick_dt, sp.drift_dt, n_pp, n_pc);
          }
        }
        __builtin_amdgcn_wave_barrier();
      }
    }
    if (is_group) {
      // the deferred list and the group sums are read back by other lanes
      __threadfence_block();
    } else {
      tot_pp += static_cast<unsigned long long>(n_pp) * size;
      tot_pc += static_cast<unsigned long long>(n_pc) * size;
    }
  }
  add_totals(a, tot_pp, tot_pc);
}

// A fixed grid of workgroups (one wavefront each) takes the walk units from
// eight queues. Workgroup b runs on XCD b % 8 and takes units from queue
// b % 8, which holds the eighth [x U / 8, (x + 1) U / 8) of the U units
// (Morton order): each XCD walks a contiguous run, so neighbouring units
// share its L2, and within an XCD the units go to whichever wavefront is
// free. A wavefront leaves once its queue is past its end, so every
// wavefront exits. The unit count and the queue heads are on the device
// (counted and cleared by the step), so the launch needs no host read.
template <typename Real, int kWaves, bool kMask, bool kDeep>
__global__ __launch_bounds__(kWaves * 64) __attribute__((amdgpu_waves_per_eu(4))) void k_walk(WalkArgs a) {
  __shared__ WaveLds<Real, kShallowFrames, kDeep> lds[kWaves];
  static_assert(kWaves == 1, "one wavefront per workgroup: the queues hand out units per workgroup");
  if (a.misc[kMiscNonfinite] | a.misc[kMiscAbort]) return;
  const uint32_t units = a.misc[kMiscGroups];
  const uint32_t x = blockIdx.x & 7u;
  const uint32_t lo = static_cast<uint32_t>(uint64_t(units) * x / 8);
  const uint32_t hi = static_cast<uint32_t>(uint64_t(units) * (x + 1) / 8);
  for (;;) {
    // Every lane takes part in the claim (adding 1 for lane 0, 0 for the
    // others) and the unit is read from lane 0's result: no branch on the
    // lane id inside the loop, so all 64 lanes stay together in it (a
    // lane-0-only claim lets the compiler keep the other lanes looping on
    // their own, with a stale claim).
    const uint32_t mine = atomicAdd(&a.misc[kMiscQueue + x], __lane_id() == 0 ? 1u : 0u);
    const uint32_t k = uniform(__builtin_amdgcn_readlane(mine, 0));
    if (k >= hi - lo) break;
    walk_group<Real, kMask>(a, lds[0], blockIdx.x, uniform(a.groups[lo + k]));
  }
}

// brute_force_accel (brute_force.h:15-28) for targets tidx[0..m) over a
// slice of the sources; partial sums per slice, reduced in slice order.
__global__ __launch_bounds__(256) void k_direct(const double4* __restrict__ src, uint32_t n, const uint32_t* __restrict__ tidx,
                                                uint32_t m, double eps2, double* __restrict__ partial, uint32_t splits) {
  __shared__ double4 s_src[256];
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  const uint32_t split = blockIdx.y;
  const uint32_t lo = static_cast<uint32_t>(uint64_t(n) * split / splits);
  const uint32_t hi = static_cast<uint32_t>(uint64_t(n) * (split + 1) / splits);
  const bool ok = k < m;
  const uint32_t i = ok ? (tidx ? tidx[k] : k) : 0xffffffffu;
  double4 t = make_double4(0, 0, 0, 0);
  if (ok) t = src[i];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (uint32_t b = lo; b < hi; b += 256) {
    __syncthreads();
    if (b + threadIdx.x < hi) s_src[threadIdx.x] = src[b + threadIdx.x];
    __syncthreads();
    const uint32_t e = hi - b < 256 ? hi - b : 256;
    for (uint32_t jj = 0; jj < e; ++jj) {
      if (b + jj == i) continue;
      const double4 q = s_src[jj];
      const double dx = q.x - t.x, dy = q.y - t.y, dz = q.z - t.z;
      const double d2 = dx * dx + dy * dy + dz * dz;
      const double ds2 = d2 + eps2;
      const double ds = sqrt(ds2);
      const double sc = q.w * (1.0 / (ds2 * ds));
      ax += dx * sc;
      ay += dy * sc;
      az += dz * sc;
    }
  }
  if (ok) {
    double* o = &partial[(uint64_t(split) * m + k) * 3];
    o[0] = ax;
    o[1] = ay;
    o[2] = az;
  }
}

__global__ __launch_bounds__(256) void k_direct_reduce(const double* __restrict__ partial, uint32_t m, uint32_t splits,
                                                       double G, double* __restrict__ acc) {
  const uint32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= m) return;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (uint32_t s = 0; s < splits; ++s) {
    const double* o = &partial[(uint64_t(s) * m + k) * 3];
    a0 += o[0];
    a1 += o[1];
    a2 += o[2];
  }
  acc[3 * uint64_t(k) + 0] = G * a0;
  acc[3 * uint64_t(k) + 1] = G * a1;
  acc[3 * uint64_t(k) + 2] = G * a2;
}

// e + q.w / |t - q| in FP64. The hardware reciprocal square root is refined
// by the second-order step y (1 + r/2 + 3r^2/8), r = 1 - x y^2 (the Taylor
// series of (1 - r)^(-1/2)), which leaves a relative error of order r^3: far
// below FP64 rounding for the hardware estimate's r.
__device__ __forceinline__ double pot_term(const double4& t, const double4& q, double e) {
  const double dx = t.x - q.x, dy = t.y - q.y, dz = t.z - q.z;
  const double x = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
  const double y = __builtin_amdgcn_rsq(x);
  const double r = __builtin_fma(-(x * y), y, 1.0);
  const double p = __builtin_fma(r, __builtin_fma(r, 0.375, 0.5), 1.0);
  return __builtin_fma(q.w * y, p, e);
}

// potential_energy (integrate.h:35-43) over unordered pairs: particle i
// adds m_j / r_ij for the partners j = i + d (mod n), d = 1 .. floor(n / 2),
// which counts every pair once (in an even-sized set the pair n / 2 apart is
// counted from its lower index only) and gives every thread the same work.
// Split s of the grid's second dimension takes the offsets
// [1 + h s / splits, 1 + h (s + 1) / splits), h = floor(n / 2); a block
// walks the window of partners its 256 particles share in tiles of 256.
// partial[s * blocks + b] = sum over the block's i of m_i sum_j m_j / r_ij.
__global__ __launch_bounds__(256) void k_potential(const double4* __restrict__ src, uint32_t n,
                                                   double* __restrict__ partial, uint32_t splits) {
  __shared__ double4 s_src[256];
  __shared__ double s_red[4];
  const uint32_t i0 = blockIdx.x * 256;
  const uint32_t i = i0 + threadIdx.x;
  const uint32_t split = blockIdx.y;
  const uint64_t h = n / 2;
  const uint64_t d_lo = 1 + h * split / splits, d_hi = 1 + h * (split + 1) / splits;
  const bool ok = i < n;
  // this thread's partner offsets [d_lo, my_hi)
  uint64_t my_hi = d_hi;
  if (n % 2 == 0 && d_hi == h + 1 && i >= h) my_hi = h;
  double4 t = make_double4(0, 0, 0, 0);
  if (ok) t = src[i];
  double e = 0.0;
  // partners of the block: positions i0 + d_lo .. i0 + 255 + d_hi - 1
  // (unwrapped). Every lane reads the same partner at the same time (an LDS
  // broadcast); a tile where some lane's window starts or ends is masked
  // per lane, the tiles in between (nearly all) are not.
  const uint64_t w_lo = uint64_t(i0) + d_lo, w_hi = uint64_t(i0) + 255 + d_hi;
  for (uint64_t jb = w_lo; jb < w_hi; jb += 256) {
    __syncthreads();
    {
      const uint64_t j = jb + threadIdx.x;
      if (j < w_hi) s_src[threadIdx.x] = src[j % n];
    }
    __syncthreads();
    const int cnt = static_cast<int>(w_hi - jb < 256 ? w_hi - jb : 256);
    // this lane's partners in the tile: [lo, hi) (unwrapped partner position
    // of offset d is i + d)
    const int64_t k_lo = static_cast<int64_t>(uint64_t(i) + d_lo) - static_cast<int64_t>(jb);
    const int64_t k_hi = static_cast<int64_t>(uint64_t(i) + my_hi) - static_cast<int64_t>(jb);
    const int lo = ok ? static_cast<int>(k_lo < 0 ? 0 : (k_lo > 256 ? 256 : k_lo)) : 0;
    const int hi = ok ? static_cast<int>(k_hi < 0 ? 0 : (k_hi > 256 ? 256 : k_hi)) : 0;
    const bool whole = __syncthreads_and(lo == 0 && hi >= cnt);
    if (whole) {
#pragma unroll 4
      for (int k = 0; k < cnt; ++k) e = pot_term(t, s_src[k], e);
    } else {
      for (int k = 0; k < cnt; ++k) {
        const double with = pot_term(t, s_src[k], e);
        e = (k >= lo && k < hi) ? with : e;
      }
    }
  }
  double v = ok ? t.w * e : 0.0;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (__lane_id() == 0) s_red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) partial[uint64_t(split) * gridDim.x + blockIdx.x] = s_red[0] + s_red[1] + s_red[2] + s_red[3];
}

// Masking r = 0 is needed only without softening (grouped_ttg.h:119-123); it
// never triggers when eps2 > 0, so one masked launch serves mixed batches.
template <typename Real, int kWaves>
hipError_t launch_walk_t(const WalkArgs& a, uint32_t slots, bool deep, bool any_unsoftened, hipStream_t s) {
  // every XCD gets the same number of workgroups
  const uint32_t blocks = (slots + 7) / 8 * 8;
  // The stack needs one frame per level below a walk's start. The first
  // kShallowFrames frames are in LDS (4 waves per SIMD); trees with more
  // levels (any system of a batch) also use the global deep frames.
  if (!any_unsoftened) {
    if (!deep)
      hipLaunchKernelGGL((k_walk<Real, kWaves, false, false>), dim3(blocks), dim3(kWaves * 64), 0, s, a);
    else
      hipLaunchKernelGGL((k_walk<Real, kWaves, false, true>), dim3(blocks), dim3(kWaves * 64), 0, s, a);
  } else {
    if (!deep)
      hipLaunchKernelGGL((k_walk<Real, kWaves, true, false>), dim3(blocks), dim3(kWaves * 64), 0, s, a);
    else
      hipLaunchKernelGGL((k_walk<Real, kWaves, true, true>), dim3(blocks), dim3(kWaves * 64), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_walk(const WalkArgs& a, uint32_t slots, bool deep, int fp64, bool any_unsoftened, hipStream_t s) {
  if (slots == 0 || slots % 8 != 0 || slots > kWalkSlots) return hipErrorInvalidValue;
  // One wavefront per workgroup: walk units differ widely in work, and a
  // workgroup's LDS is released only when all its waves are done (2 waves
  // per workgroup: 49.7 ms at 16M, 4: 53.9, 8: 64.8; 1: 47.0).
  return fp64 ? launch_walk_t<double, 1>(a, slots, deep, any_unsoftened, s)
              : launch_walk_t<float, 1>(a, slots, deep, any_unsoftened, s);
}

hipError_t launch_potential(const double4* src, uint32_t n, double* partial, uint32_t splits, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_potential, dim3((n + 255) / 256, splits), dim3(256), 0, s, src, n, partial, splits);
  return hipGetLastError();
}

uint32_t direct_splits(uint32_t n, uint32_t m) {
  const uint32_t tblocks = (m + 255) / 256;
  uint32_t splits = 2048 / (tblocks ? tblocks : 1);
  if (splits < 1) splits = 1;
  if (splits > 512) splits = 512;
  while (splits > 1 && n / splits < 256) splits >>= 1;
  return splits;
}

hipError_t launch_direct(const double4* src, uint32_t n, const uint32_t* tidx, uint32_t m, double eps2, double G,
                         double* partial, uint32_t splits, double* acc, hipStream_t s) {
  if (m == 0) return hipSuccess;
  hipLaunchKernelGGL(k_direct, dim3((m + 255) / 256, splits), dim3(256), 0, s, src, n, tidx, m, eps2, partial, splits);
  hipLaunchKernelGGL(k_direct_reduce, dim3((m + 255) / 256), dim3(256), 0, s, partial, m, splits, G, acc);
  return hipGetLastError();
}

}  // namespace bhk
