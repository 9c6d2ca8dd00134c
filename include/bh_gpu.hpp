// C++ host interface over the C ABI of bh_gpu.h, mirroring the reference's
// grouped-path API (ttg-parsec/src/integrate_ttg.h, traversal_ttg.h) so that
// a caller of bh::step_particles can switch by changing the namespace:
//
//   reference                                   this header
//   bh::Particle        particle.h:7-24         bh_gpu::Particle (same 64-byte layout)
//   bh::ForceParams     traversal_ttg.h:136-141 bh_gpu::ForceParams (same fields, defaults)
//   bh::StepTiming      integrate_ttg.h:24-28   bh_gpu::StepTiming
//   bh::GroupedWorkspace integrate_ttg.h:61-64 bh_gpu::GroupedWorkspace (owns the device context)
//   bh::step_particles  integrate_ttg.h:66-148 bh_gpu::step_particles (grouped branch, one GPU)
//
// Error behaviour follows the reference: an incomplete tree throws
// std::logic_error (integrate_ttg.h:104-106); invalid input throws
// std::invalid_argument / std::runtime_error (main.cc:147); device failures
// throw std::runtime_error with bh_last_error()'s message.
#pragma once

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "bh_gpu.h"

namespace bh_gpu {

using Particle = bh_particle;  // id, pad, mass, pos[3], vel[3]

struct ForceParams {
  double theta = 0.5;
  double eps2 = 0.0;
  double G = 1.0;
  bool use_quadrupole = true;
};

struct StepTiming {
  double tree_build_ms = 0.0;
  double traversal_ms = 0.0;
  double exchange_ms = 0.0;
};

[[noreturn]] inline void throw_status(int rc, const std::string& what) {
  switch (rc) {
    case BH_ERR_INCOMPLETE:
      throw std::logic_error(what);
    case BH_ERR_INVALID:
    case BH_ERR_NONFINITE:
    case BH_ERR_TOO_LARGE:
      throw std::invalid_argument(what);
    default:
      throw std::runtime_error(what);
  }
}

// Storage of the grouped pass kept across steps: here, the device context
// with its grow-only buffers on one GPU.
class GroupedWorkspace {
 public:
  explicit GroupedWorkspace(int device = 0) {
    const int rc = bh_ctx_create(&ctx_, device);
    if (rc != BH_OK) throw_status(rc, "bh_ctx_create failed for device " + std::to_string(device));
  }
  ~GroupedWorkspace() { bh_ctx_destroy(ctx_); }
  GroupedWorkspace(const GroupedWorkspace&) = delete;
  GroupedWorkspace& operator=(const GroupedWorkspace&) = delete;
  bh_ctx* ctx() const { return ctx_; }
  void check(int rc) const {
    if (rc != BH_OK) throw_status(rc, bh_last_error(ctx_));
  }

 private:
  bh_ctx* ctx_ = nullptr;
};

inline bh_force_params to_c(const ForceParams& fp) {
  bh_force_params c{};
  c.theta = fp.theta;
  c.eps2 = fp.eps2;
  c.G = fp.G;
  c.use_quadrupole = fp.use_quadrupole ? 1 : 0;
  return c;
}

// step_particles (integrate_ttg.h:66-69), grouped branch (group_size > 0):
// builds the tree, computes every particle's acceleration, applies
// v += kick_dt * a and, if drift_dt != 0, x += drift_dt * v, and returns the
// particles in Morton order. Without a workspace a temporary context is used.
inline std::vector<Particle> step_particles(std::vector<Particle> particles, const ForceParams& fp, unsigned max_leaf,
                                            std::uint32_t group_size, std::uint32_t target_size, double kick_dt,
                                            double drift_dt, StepTiming* timing = nullptr,
                                            GroupedWorkspace* workspace = nullptr) {
  if (group_size == 0) throw std::invalid_argument("group_size 0 selects the per-node pass, which has no GPU version");
  std::unique_ptr<GroupedWorkspace> local;
  GroupedWorkspace* ws = workspace;
  if (!ws) {
    local = std::make_unique<GroupedWorkspace>(0);
    ws = local.get();
  }
  const bh_force_params c = to_c(fp);
  bh_step_timing t{};
  ws->check(bh_step(ws->ctx(), particles.data(), particles.size(), &c, max_leaf, group_size, target_size, kick_dt,
                    drift_dt, &t));
  if (timing) {
    timing->tree_build_ms = t.tree_build_ms;
    timing->traversal_ms = t.traversal_ms;
    timing->exchange_ms = t.exchange_ms;
  }
  return particles;
}

}  // namespace bh_gpu
