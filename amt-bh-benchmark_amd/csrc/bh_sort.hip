// Bounding box, Morton keys and the stable (key, index) sort on CDNA4.
//
// Reference semantics:
//   particle_bounds          flat_build_ttg.h:93-113
//   domain_from_bounds       domain.h:23-30
//   morton_key/spread_bits3  flat_tree.h:104-125
//   sort order               flat_build_ttg.h:62-66 (ascending key, ties by input index)
//   gather to Morton order   flat_build_ttg.h:180-188
//
// This file is compiled with -ffp-contract=off: the domain and key
// arithmetic must round exactly as the reference's FP64 code does.
//
// The sort is an LSD radix sort over the 63-bit keys with the input index as
// payload, 8 passes of 8-bit digits. LSD radix sorting is stable, so the
// result is exactly the (key, index) order. Each pass is a single "onesweep"
// kernel: a workgroup ranks a 4096-key tile in LDS with wavefront ballots,
// obtains its global digit offsets by decoupled look-back over the preceding
// tiles, and writes the tile out digit-sorted so stores coalesce.
#include "bh_internal.h"

namespace bhk {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = static_cast<int>(kSortTile) / kBlock;

__device__ inline uint64_t lanemask_lt() {
  const int lane = __lane_id();
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// domain.h:23-30 from the reduced bounds lohi = {lo.x, lo.y, lo.z, hi.x, hi.y, hi.z}.
__device__ inline void domain_from_lohi(const double* lohi, double c[3], double* hw) {
  c[0] = 0.5 * (lohi[0] + lohi[3]);
  c[1] = 0.5 * (lohi[1] + lohi[4]);
  c[2] = 0.5 * (lohi[2] + lohi[5]);
  const double ex = lohi[3] - lohi[0], ey = lohi[4] - lohi[1], ez = lohi[5] - lohi[2];
  // std::max({ex, ey, ez}): the first largest; equal values are identical
  double m = ex;
  if (m < ey) m = ey;
  if (m < ez) m = ez;
  double h = 0.5 * m;
  if (!(h > 0.0)) h = 1.0;
  h *= 1.0 + 1e-9;
  *hw = h;
}

__device__ inline uint64_t spread_bits3(uint32_t v) {
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

__device__ inline uint32_t quantize(double v, double lo, double scale) {
  const double cells = 2097152.0;
  double t = floor((v - lo) * scale);
  t = t < 0.0 ? 0.0 : (cells - 1.0 < t ? cells - 1.0 : t);  // std::clamp(t, 0, cells - 1)
  return static_cast<uint32_t>(t);
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) p[i] = v;
}

// Zeroes several ranges in one launch (a step's flag words, totals, digit
// histograms and look-back words).
__global__ __launch_bounds__(kBlock) void k_zero_ranges(ZeroRanges r) {
  const uint64_t stride = uint64_t(gridDim.x) * kBlock;
  const uint64_t t = blockIdx.x * uint64_t(kBlock) + threadIdx.x;
#pragma unroll
  for (int k = 0; k < ZeroRanges::kMax; ++k)
    for (uint64_t i = t; i < r.n[k]; i += stride) r.p[k][i] = 0u;
}

__device__ inline double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// System id of every particle, one block per chunk.
__global__ __launch_bounds__(kBlock) void k_fill_sid(const uint4* __restrict__ chunks, uint16_t* __restrict__ sid) {
  const uint4 c = chunks[blockIdx.x];
  for (uint32_t i = c.y + threadIdx.x; i < c.z; i += kBlock) sid[i] = static_cast<uint16_t>(c.x);
}

// Where the build reads a particle's position: its record, or the compact
// copy a resident step's walk wrote for its output, four blocks of n
// doubles (x, y, z, m), of which the build reads x, y, z. `finite` checks
// what the source holds (the walk that wrote the compact copy flagged any
// non-finite position, velocity or mass itself).
struct RecordSrc {
  const bh_particle* p;
};
struct CompactSrc {
  const double* p;
  uint64_t n;
};
__device__ inline void pos_of(const RecordSrc& s, uint32_t i, double* x, double* y, double* z, bool* finite) {
  const bh_particle& q = s.p[i];
  *x = q.pos[0];
  *y = q.pos[1];
  *z = q.pos[2];
  *finite = isfinite(*x) && isfinite(*y) && isfinite(*z) && isfinite(q.mass) && isfinite(q.vel[0]) &&
            isfinite(q.vel[1]) && isfinite(q.vel[2]);
}
__device__ inline void pos_of(const CompactSrc& s, uint32_t i, double* x, double* y, double* z, bool* finite) {
  *x = s.p[i];
  *y = s.p[s.n + i];
  *z = s.p[2 * s.n + i];
  *finite = isfinite(*x) && isfinite(*y) && isfinite(*z);
}

// Min/max of the positions of one chunk (flat_build_ttg.h:93-113; min/max
// is exact, so the chunking does not matter); flags any non-finite
// position, velocity or mass (and, from a compact copy, its writer's flag).
template <typename Src>
__global__ __launch_bounds__(kBlock) void k_bounds_partial(Src src, const uint32_t* __restrict__ carried_bad,
                                                           const uint4* __restrict__ chunks, double* __restrict__ partial,
                                                           uint32_t* __restrict__ nonfinite) {
  const double inf = __builtin_huge_val();
  double lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
  bool bad = false;
  if (carried_bad && blockIdx.x == 0 && threadIdx.x == 0 && *carried_bad) atomicOr(nonfinite, 1u);
  const uint4 c = chunks[blockIdx.x];
  // kB elements per lane and round, all loads issued before the first use
  // (indices clamped into the chunk, extra elements dropped after the load)
  constexpr int kB = 4;
  for (uint32_t i0 = c.y + threadIdx.x; i0 < c.z; i0 += kB * kBlock) {
    double x[kB], y[kB], z[kB];
    bool fin[kB];
#pragma unroll
    for (int r = 0; r < kB; ++r) {
      const uint32_t i = i0 + r * kBlock < c.z ? i0 + r * kBlock : c.z - 1;
      pos_of(src, i, &x[r], &y[r], &z[r], &fin[r]);
    }
#pragma unroll
    for (int r = 0; r < kB; ++r) {
      if (i0 + r * kBlock >= c.z) break;
      bad |= !fin[r];
      lo[0] = fmin(lo[0], x[r]);
      lo[1] = fmin(lo[1], y[r]);
      lo[2] = fmin(lo[2], z[r]);
      hi[0] = fmax(hi[0], x[r]);
      hi[1] = fmax(hi[1], y[r]);
      hi[2] = fmax(hi[2], z[r]);
    }
  }
  __shared__ double s[4][6];
  const int w = threadIdx.x >> 6, lane = __lane_id();
  double r[6];
  for (int a = 0; a < 3; ++a) {
    r[a] = wave_min(lo[a]);
    r[3 + a] = wave_max(hi[a]);
  }
  if (lane == 0)
    for (int a = 0; a < 6; ++a) s[w][a] = r[a];
  if (__ballot(bad) != 0 && lane == 0) atomicOr(nonfinite, 1u);
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    double v = s[0][a];
    for (int k = 1; k < 4; ++k) v = a < 3 ? fmin(v, s[k][a]) : fmax(v, s[k][a]);
    partial[blockIdx.x * 6 + a] = v;
  }
}

// Per system: reduce its chunks' partials to lohi[6 s .. 6 s + 6).
__global__ __launch_bounds__(kBlock) void k_bounds_final(const double* __restrict__ partial,
                                                         const uint32_t* __restrict__ chunk_first, double* __restrict__ lohi) {
  const double inf = __builtin_huge_val();
  __shared__ double s[4][6];
  double r[6] = {inf, inf, inf, -inf, -inf, -inf};
  const uint32_t b0 = chunk_first[blockIdx.x], b1 = chunk_first[blockIdx.x + 1];
  for (uint32_t b = b0 + threadIdx.x; b < b1; b += kBlock) {
    for (int a = 0; a < 3; ++a) {
      r[a] = fmin(r[a], partial[b * 6 + a]);
      r[3 + a] = fmax(r[3 + a], partial[b * 6 + 3 + a]);
    }
  }
  const int w = threadIdx.x >> 6, lane = __lane_id();
  for (int a = 0; a < 3; ++a) {
    r[a] = wave_min(r[a]);
    r[3 + a] = wave_max(r[3 + a]);
  }
  if (lane == 0)
    for (int a = 0; a < 6; ++a) s[w][a] = r[a];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int a = threadIdx.x;
    double v = s[0][a];
    for (int k = 1; k < 4; ++k) v = a < 3 ? fmin(v, s[k][a]) : fmax(v, s[k][a]);
    lohi[6 * blockIdx.x + a] = v;
  }
}

// Morton key per particle (input order) and the 8 digit histograms of all
// keys, for the onesweep passes' global digit offsets.
template <typename Src>
__global__ __launch_bounds__(kBlock) void k_keys_hist(Src src, uint32_t n,
                                                      const uint16_t* __restrict__ sid, const double* __restrict__ lohi,
                                                      uint64_t* __restrict__ keys, uint32_t* __restrict__ hist) {
  __shared__ uint32_t s_h[8 * 256];
  for (int i = threadIdx.x; i < 8 * 256; i += kBlock) s_h[i] = 0;
  __syncthreads();
  const int lane = __lane_id();
  const uint32_t stride = gridDim.x * kBlock;
  // each particle's position and system are loaded one round ahead, so the
  // loads of round r + 1 are in flight while round r computes
  double nx = 0.0, ny = 0.0, nz = 0.0;
  uint32_t ns = 0;
  {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
      bool fin;
      pos_of(src, i, &nx, &ny, &nz, &fin);
      ns = sid[i];
    }
  }
  for (uint32_t i0 = blockIdx.x * kBlock; i0 < n; i0 += stride) {
    // the wave's lanes hold consecutive particles; those below n are a
    // prefix of the wave
    const uint32_t i = i0 + threadIdx.x;
    const bool ok = i < n;
    const double x = nx, y = ny, z = nz;
    const uint32_t sys_i = ns;
    if (i + stride < n) {
      bool fin;
      pos_of(src, i + stride, &nx, &ny, &nz, &fin);
      ns = sid[i + stride];
    }
    uint64_t key = 0;
    if (ok) {
      double c[3], hw;
      domain_from_lohi(lohi + 6 * sys_i, c, &hw);
      const double scale = 2097152.0 / (2.0 * hw);
      const double lx = c[0] - hw, ly = c[1] - hw, lz = c[2] - hw;
      key = spread_bits3(quantize(x, lx, scale)) | (spread_bits3(quantize(y, ly, scale)) << 1) |
            (spread_bits3(quantize(z, lz, scale)) << 2);
      keys[i] = key;
    }
    // Input in (or near) Morton order, as a resident step's is, gives long
    // runs of equal digits across the wave, whose atomics on one LDS word
    // would serialise: the first lane of each run adds the run's length.
    const uint32_t nvalid = static_cast<uint32_t>(__popcll(__ballot(ok)));
    const uint64_t prev = __shfl_up(key, 1);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint32_t d = static_cast<uint32_t>(key >> (8 * b)) & 255u;
      const uint32_t pd = static_cast<uint32_t>(prev >> (8 * b)) & 255u;
      const bool head = ok && (lane == 0 || d != pd);
      const uint64_t heads = __ballot(head);
      const uint64_t after = lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
      const uint32_t end = after ? static_cast<uint32_t>(__ffsll(static_cast<long long>(after)) - 1) : nvalid;
      if (head) atomicAdd(&s_h[b * 256 + d], end - static_cast<uint32_t>(lane));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 8 * 256; i += kBlock)
    if (s_h[i]) atomicAdd(&hist[i], s_h[i]);
}

// Exclusive block scan of one value per thread (256 threads).
__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t* s_wsum, uint32_t* total) {
  const int lane = __lane_id(), w = threadIdx.x >> 6;
  uint32_t x = v;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t u = __shfl_up(x, o);
    if (lane >= o) x += u;
  }
  if (lane == 63) s_wsum[w] = x;
  __syncthreads();
  uint32_t off = 0, tot = 0;
  for (int k = 0; k < kBlock / 64; ++k) {
    if (k < w) off += s_wsum[k];
    tot += s_wsum[k];
  }
  if (total) *total = tot;
  __syncthreads();
  return off + x - v;
}

__global__ __launch_bounds__(kBlock) void k_hist_scan(const uint32_t* __restrict__ hist, uint32_t* __restrict__ dbase) {
  __shared__ uint32_t s_w[4];
  const uint32_t pass = blockIdx.x;
  const uint32_t v = hist[pass * 256 + threadIdx.x];
  dbase[pass * 256 + threadIdx.x] = block_excl_scan(v, s_w, nullptr);
}

// kMode 0: first key pass (payload = input index); 1: key pass; 2: system
// pass (digit of sid[payload], keys carried along).
template <int kMode>
__global__ __launch_bounds__(kBlock) void k_onesweep(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout, uint32_t n,
                                                     uint32_t shift, const uint16_t* __restrict__ sid,
                                                     const uint32_t* __restrict__ dbase, uint32_t* __restrict__ lookback,
                                                     uint32_t* __restrict__ tile_counter) {
  __shared__ uint32_t s_hist[4][256];   // per-wave digit counters, then per-wave exclusive bases
  __shared__ uint32_t s_lexcl[256];     // tile-local exclusive digit offsets
  __shared__ uint32_t s_adj[256];       // global position - local position, per digit
  __shared__ uint64_t s_k[kSortTile];   // the tile in digit order: keys, then (reused) payloads
  __shared__ uint8_t s_d[kSortTile];
  uint32_t* s_v = reinterpret_cast<uint32_t*>(s_k);
  __shared__ uint32_t s_w[4];
  __shared__ uint32_t s_tile;

  const int tid = threadIdx.x, w = tid >> 6, lane = __lane_id();
  if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
  for (int i = tid; i < 4 * 256; i += kBlock) (&s_hist[0][0])[i] = 0;
  __syncthreads();
  const uint32_t tile = s_tile;
  const uint32_t base = tile * kSortTile;

  uint64_t k[kItems];
  uint32_t v[kItems];
  uint32_t r[kItems];
  uint32_t dg[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t idx = base + w * (kSortTile / 4) + j * 64 + lane;
    const bool ok = idx < n;
    k[j] = ok ? kin[idx] : ~0ull;
    v[j] = kMode == 0 ? idx : (ok ? vin[idx] : 0u);
    if (kMode == 2)
      dg[j] = ok ? (static_cast<uint32_t>(sid[v[j]]) >> shift) & 255u : 255u;
    else
      dg[j] = static_cast<uint32_t>(k[j] >> shift) & 255u;
  }
  const uint64_t lt = lanemask_lt();
  // Stable in-tile ranking: items in (wave, item, lane) order follow the
  // input order; lanes with an equal digit are found with 8 ballots.
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t idx = base + w * (kSortTile / 4) + j * 64 + lane;
    const bool ok = idx < n;
    const uint32_t d = dg[j];
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t below = __popcll(peers & lt);
    const uint32_t cnt = s_hist[w][d];
    r[j] = cnt + below;
    __builtin_amdgcn_wave_barrier();
    if (ok && below == 0) s_hist[w][d] = cnt + __popcll(peers);
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // per digit: totals over waves and per-wave exclusive bases
  const uint32_t t = tid;
  uint32_t total = 0;
  {
    uint32_t c0 = s_hist[0][t], c1 = s_hist[1][t], c2 = s_hist[2][t], c3 = s_hist[3][t];
    s_hist[0][t] = 0;
    s_hist[1][t] = c0;
    s_hist[2][t] = c0 + c1;
    s_hist[3][t] = c0 + c1 + c2;
    total = c0 + c1 + c2 + c3;
  }
  uint32_t* my = &lookback[uint64_t(tile) * 256 + t];
  __hip_atomic_store(my, (tile == 0 ? kLbPre : kLbAgg) | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const uint32_t lex = block_excl_scan(total, s_w, nullptr);
  s_lexcl[t] = lex;
  uint32_t excl = 0;
  if (tile > 0) {
    int64_t j = static_cast<int64_t>(tile) - 1;
    for (;;) {
      const uint32_t val = __hip_atomic_load(&lookback[uint64_t(j) * 256 + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const uint32_t f = val & ~kLbVal;
      if (f == 0) continue;
      excl += val & kLbVal;
      if (f == kLbPre) break;
      --j;
    }
    __hip_atomic_store(my, kLbPre | (excl + total), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  s_adj[t] = dbase[t] + excl - lex;
  __syncthreads();
  // keys, then payloads, through the same LDS tile in digit order, so that
  // the stores coalesce (one tile buffer keeps three workgroups per CU)
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t idx = base + w * (kSortTile / 4) + j * 64 + lane;
    if (idx < n) {
      const uint32_t d = dg[j];
      const uint32_t pos = s_lexcl[d] + s_hist[w][d] + r[j];
      s_k[pos] = k[j];
      s_d[pos] = static_cast<uint8_t>(d);
      r[j] = pos;
    }
  }
  __syncthreads();
  const uint32_t tn = n - base < kSortTile ? n - base : kSortTile;
  uint32_t g[kItems];
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t e = tid + j * kBlock;
    g[j] = e < tn ? s_adj[s_d[e]] + e : 0u;
    if (e < tn) kout[g[j]] = s_k[e];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t idx = base + w * (kSortTile / 4) + j * 64 + lane;
    if (idx < n) s_v[r[j]] = v[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kItems; ++j) {
    const uint32_t e = tid + j * kBlock;
    if (e < tn) vout[g[j]] = s_v[e];
  }
}

// Permutes the particles into Morton order and writes the FP64 walk copy
// (x, y, z, m), as sort_particles' final loop does (flat_build_ttg.h:180-188).
// Without out only the walk copy is written (a step's walk writes the
// Morton-ordered particles from the input itself).
__global__ __launch_bounds__(kBlock) void k_gather(const bh_particle* __restrict__ in, bh_particle* __restrict__ out,
                                                   const uint32_t* __restrict__ perm, double4* __restrict__ posm, uint32_t n) {
  // four lanes per particle, one 16-byte quarter each: consecutive lanes
  // read and write consecutive 16-byte pieces
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t k = t >> 2, part = t & 3u;
  if (k >= n) return;  // whole quads leave together
  const uint32_t j = perm[k];
  // the walk copy needs quarters 0-2 only
  uint4 v = make_uint4(0, 0, 0, 0);
  if (out || part < 3) v = reinterpret_cast<const uint4*>(in + j)[part];
  if (out) reinterpret_cast<uint4*>(out + k)[part] = v;
  // quarters: 0 = {id, pad, mass}, 1 = {x, y}, 2 = {z, vx}, 3 = {vy, vz}
  const int lane = __lane_id();
  const uint32_t mlo = static_cast<uint32_t>(__shfl(static_cast<int>(v.z), lane ^ 2));
  const uint32_t mhi = static_cast<uint32_t>(__shfl(static_cast<int>(v.w), lane ^ 2));
  uint4* q = reinterpret_cast<uint4*>(posm + k);
  if (part == 1) q[0] = v;
  if (part == 2) q[1] = make_uint4(v.x, v.y, mlo, mhi);
}

// The walk copy from a carried compact copy (blocks of n x, y, z and m in
// the last resident step's order, which is the order perm indexes). The
// copy holds the same FP64 values as the records, so the result equals
// k_gather's.
// Two lanes per particle: lane half 0 moves x, y, half 1 z, m (one 16-byte
// half of the walk record each).
__global__ __launch_bounds__(kBlock) void k_gather_pos(const double* __restrict__ pos, const uint32_t* __restrict__ perm,
                                                       double4* __restrict__ posm, uint32_t n) {
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t k = t >> 1, half = t & 1u;
  if (k >= n) return;
  const uint64_t j = perm[k];
  const uint64_t b0 = (2ull * half) * n + j;
  reinterpret_cast<double2*>(posm + k)[half] = make_double2(pos[b0], pos[b0 + n]);
}

inline uint32_t grid_for(uint64_t n, uint32_t cap) {
  uint64_t g = (n + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  return static_cast<uint32_t>(g < cap ? g : cap);
}

}  // namespace

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 4096)), dim3(kBlock), 0, s, p, value, n);
  return hipGetLastError();
}

hipError_t launch_zero(const ZeroRanges& r, hipStream_t s) {
  uint64_t most = 0;
  for (int k = 0; k < ZeroRanges::kMax; ++k) most = r.n[k] > most ? r.n[k] : most;
  if (most == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_ranges, dim3(grid_for(most, 4096)), dim3(kBlock), 0, s, r);
  return hipGetLastError();
}

hipError_t launch_fill_sid(const uint4* chunks, uint32_t n_chunks, uint16_t* sid, hipStream_t s) {
  if (n_chunks == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_sid, dim3(n_chunks), dim3(kBlock), 0, s, chunks, sid);
  return hipGetLastError();
}

hipError_t launch_bounds(const bh_particle* p, const uint4* chunks, uint32_t n_chunks, const uint32_t* chunk_first,
                         uint32_t n_sys, double* partial, uint32_t* nonfinite, double* lohi, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds_partial<RecordSrc>, dim3(n_chunks), dim3(kBlock), 0, s, RecordSrc{p}, nullptr, chunks,
                     partial, nonfinite);
  hipLaunchKernelGGL(k_bounds_final, dim3(n_sys), dim3(kBlock), 0, s, partial, chunk_first, lohi);
  return hipGetLastError();
}

hipError_t launch_bounds(const double* pos, uint32_t n, const uint32_t* carried_bad, const uint4* chunks,
                         uint32_t n_chunks, const uint32_t* chunk_first, uint32_t n_sys, double* partial,
                         uint32_t* nonfinite, double* lohi, hipStream_t s) {
  hipLaunchKernelGGL(k_bounds_partial<CompactSrc>, dim3(n_chunks), dim3(kBlock), 0, s, CompactSrc{pos, n}, carried_bad,
                     chunks, partial, nonfinite);
  hipLaunchKernelGGL(k_bounds_final, dim3(n_sys), dim3(kBlock), 0, s, partial, chunk_first, lohi);
  return hipGetLastError();
}

hipError_t launch_keys_hist(const bh_particle* p, uint32_t n, const uint16_t* sid, const double* lohi, uint64_t* keys,
                            uint32_t* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_keys_hist<RecordSrc>, dim3(grid_for(n, 2048)), dim3(kBlock), 0, s, RecordSrc{p}, n, sid, lohi,
                     keys, hist);
  return hipGetLastError();
}

hipError_t launch_keys_hist(const double* pos, uint32_t n, const uint16_t* sid, const double* lohi, uint64_t* keys,
                            uint32_t* hist, hipStream_t s) {
  hipLaunchKernelGGL(k_keys_hist<CompactSrc>, dim3(grid_for(n, 2048)), dim3(kBlock), 0, s, CompactSrc{pos, n}, n, sid,
                     lohi, keys, hist);
  return hipGetLastError();
}

uint64_t radix_lookback_words(uint32_t n) {
  const uint64_t tiles = (uint64_t(n) + kSortTile - 1) / kSortTile;
  return 10 * tiles * 256 + 10;
}

hipError_t launch_radix_sort(uint64_t* keys_a, uint32_t* vals_a, uint64_t* keys_b, uint32_t* vals_b, uint32_t n,
                             int sid_passes, const uint16_t* sid, uint32_t* hist, uint32_t* digit_base,
                             uint32_t* lookback, uint32_t* tile_counters, uint64_t** keys_out, uint32_t** vals_out,
                             hipStream_t s) {
  const uint32_t tiles = (n + kSortTile - 1) / kSortTile;
  const int passes = 8 + sid_passes;
  hipLaunchKernelGGL(k_hist_scan, dim3(passes), dim3(kBlock), 0, s, hist, digit_base);
  uint64_t* kin = keys_a;
  uint32_t* vin = vals_a;
  uint64_t* ko = keys_b;
  uint32_t* vo = vals_b;
  for (int pass = 0; pass < passes; ++pass) {
    uint32_t* lb = lookback + uint64_t(pass) * tiles * 256;
    const uint32_t* db = digit_base + pass * 256;
    if (pass == 0)
      hipLaunchKernelGGL(k_onesweep<0>, dim3(tiles), dim3(kBlock), 0, s, kin, vin, ko, vo, n, 0u, sid, db, lb,
                         tile_counters + pass);
    else if (pass < 8)
      hipLaunchKernelGGL(k_onesweep<1>, dim3(tiles), dim3(kBlock), 0, s, kin, vin, ko, vo, n, 8u * pass, sid, db, lb,
                         tile_counters + pass);
    else
      hipLaunchKernelGGL(k_onesweep<2>, dim3(tiles), dim3(kBlock), 0, s, kin, vin, ko, vo, n, 8u * (pass - 8), sid, db,
                         lb, tile_counters + pass);
    uint64_t* tk = kin;
    kin = ko;
    ko = tk;
    uint32_t* tv = vin;
    vin = vo;
    vo = tv;
  }
  *keys_out = kin;
  *vals_out = vin;
  return hipGetLastError();
}

hipError_t launch_gather(const bh_particle* in, bh_particle* out, const uint32_t* perm, double4* posm, uint32_t n,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_gather, dim3((4ull * n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, in, out, perm, posm, n);
  return hipGetLastError();
}

hipError_t launch_gather(const double* pos, const uint32_t* perm, double4* posm, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_gather_pos, dim3((2ull * n + kBlock - 1) / kBlock), dim3(kBlock), 0, s, pos, perm, posm, n);
  return hipGetLastError();
}

}  // namespace bhk
