// bh_run: the reference driver (ttg-parsec/src/main.cc) on one MI355X.
//
// Same flags and defaults as main.cc:18-92, same random initial conditions
// (main.cc:97-111: std::mt19937(seed), per particle mass U(0.5,1.5), then
// pos x,y,z U(-spread,spread), then vel x,y,z U(-vel_spread,vel_spread)),
// same leapfrog schedule (main.cc:199-246: a half-kick + drift to prime,
// steps-1 full kick + drift steps, a closing half-kick without drift, i.e.
// steps+1 force evaluations) and the same report lines. Particles stay
// resident on the GPU between steps; they are copied to the host only for
// energy reports and snapshots. Extra flags: --device, --fp64 (FP64 walk),
// --output-csv PATH (final state, in Morton order like the reference's).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <iterator>
#include <iostream>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/bh_gpu.hpp"

namespace {

struct Options {
  std::string input, vtk_prefix, output_csv;
  bool write_vtk = false;
  int steps = 10;
  double dt = 1e-3, theta = 0.5, eps = 0.0, G = 1.0;
  unsigned max_leaf = 8;
  uint32_t group_size = 256, target_size = 32;
  bool use_quadrupole = true;
  int energy_every = 1;
  bool time_steps = false;
  std::size_t n = 1000;
  unsigned seed = 1;
  double spread = 5.0, vel_spread = 0.0;
  int device = 0;
  bool fp64 = false;
};

// The flags of main.cc:18-92 (plus this driver's own), each with the
// setter of its option. A flag without a setter value is a switch.
Options parse_args(int argc, char** argv) {
  Options o;
  using Set = std::function<void(const std::string&)>;
  struct Flag {
    const char* name;
    bool takes_value;
    Set set;
  };
  const Flag flags[] = {
      {"--input", true, [&](const std::string& v) { o.input = v; }},
      {"--steps", true, [&](const std::string& v) { o.steps = std::stoi(v); }},
      {"--dt", true, [&](const std::string& v) { o.dt = std::stod(v); }},
      {"--theta", true, [&](const std::string& v) { o.theta = std::stod(v); }},
      {"--eps", true, [&](const std::string& v) { o.eps = std::stod(v); }},
      {"--G", true, [&](const std::string& v) { o.G = std::stod(v); }},
      {"--max-leaf", true, [&](const std::string& v) { o.max_leaf = static_cast<unsigned>(std::stoul(v)); }},
      {"--target-size", true, [&](const std::string& v) { o.target_size = static_cast<uint32_t>(std::stoul(v)); }},
      {"--group-size", true, [&](const std::string& v) { o.group_size = static_cast<uint32_t>(std::stoul(v)); }},
      {"--no-quadrupole", false, [&](const std::string&) { o.use_quadrupole = false; }},
      {"--vtk-prefix", true, [&](const std::string& v) { o.vtk_prefix = v, o.write_vtk = true; }},
      {"--energy-every", true, [&](const std::string& v) { o.energy_every = std::stoi(v); }},
      {"--time", false, [&](const std::string&) { o.time_steps = true; }},
      {"--n", true, [&](const std::string& v) { o.n = std::stoul(v); }},
      {"--seed", true, [&](const std::string& v) { o.seed = static_cast<unsigned>(std::stoul(v)); }},
      {"--spread", true, [&](const std::string& v) { o.spread = std::stod(v); }},
      {"--vel-spread", true, [&](const std::string& v) { o.vel_spread = std::stod(v); }},
      {"--device", true, [&](const std::string& v) { o.device = std::stoi(v); }},
      {"--fp64", false, [&](const std::string&) { o.fp64 = true; }},
      {"--output-csv", true, [&](const std::string& v) { o.output_csv = v; }},
  };
  int k = 1;
  while (k < argc) {
    const std::string flag = argv[k++];
    const Flag* f = std::find_if(std::begin(flags), std::end(flags), [&](const Flag& x) { return flag == x.name; });
    if (f == std::end(flags)) throw std::runtime_error("unknown argument: " + flag);
    if (!f->takes_value) {
      f->set(std::string());
      continue;
    }
    if (k >= argc) throw std::runtime_error("missing value for " + flag);
    f->set(argv[k++]);
  }
  return o;
}

std::vector<bh_particle> random_particles(const Options& o) {
  std::mt19937 rng(o.seed);
  std::uniform_real_distribution<double> pos(-o.spread, o.spread), vel(-o.vel_spread, o.vel_spread), mass(0.5, 1.5);
  std::vector<bh_particle> p(o.n);
  for (std::size_t i = 0; i < o.n; ++i) {
    p[i] = bh_particle{};
    p[i].id = static_cast<uint32_t>(i);
    p[i].mass = mass(rng);
    for (int a = 0; a < 3; ++a) p[i].pos[a] = pos(rng);
    for (int a = 0; a < 3; ++a) p[i].vel[a] = vel(rng);
  }
  return p;
}

std::vector<bh_particle> read_csv(const std::string& path) {
  bh_particle* buf = nullptr;
  uint64_t n = 0;
  char err[512] = {0};
  const int rc = bh_read_csv(path.c_str(), &buf, &n, err, sizeof err);
  if (rc != BH_OK) throw std::runtime_error(err);
  std::vector<bh_particle> p(buf, buf + n);
  bh_free(buf);
  return p;
}

void write_vtk(const std::string& path, const std::vector<bh_particle>& p) {
  char err[512] = {0};
  if (bh_write_vtk(path.c_str(), p.data(), p.size(), err, sizeof err) != BH_OK) throw std::runtime_error(err);
}

void write_csv(const std::string& path, const std::vector<bh_particle>& p) {
  FILE* f = std::fopen(path.c_str(), "w");
  if (!f) throw std::runtime_error("cannot open output file: " + path);
  for (const auto& q : p)
    std::fprintf(f, "%u,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", q.id, q.mass, q.pos[0], q.pos[1], q.pos[2], q.vel[0],
                 q.vel[1], q.vel[2]);
  std::fclose(f);
}

}  // namespace

int main(int argc, char** argv) {
  try {
    const Options opt = parse_args(argc, argv);
    std::vector<bh_particle> particles = opt.input.empty() ? random_particles(opt) : read_csv(opt.input);
    if (particles.empty()) throw std::runtime_error("input file contains no particles");
    if (opt.input.empty()) {
      std::cout << "generated " << particles.size() << " random particles (seed=" << opt.seed << ", spread=" << opt.spread
                << ", vel_spread=" << opt.vel_spread << ")\n";
    }
    bh_force_params fp{};
    fp.theta = opt.theta;
    fp.eps2 = opt.eps * opt.eps;
    fp.G = opt.G;
    fp.use_quadrupole = opt.use_quadrupole ? 1 : 0;

    bh_gpu::GroupedWorkspace ws(opt.device);
    bh_ctx* ctx = ws.ctx();
    ws.check(bh_set_precision(ctx, opt.fp64 ? BH_FP64 : BH_FP32));
    ws.check(bh_state_upload(ctx, particles.data(), particles.size()));

    auto download = [&]() { ws.check(bh_state_download(ctx, particles.data(), particles.size())); };
    auto report = [&](int step, const char* tag) {
      download();
      double ek = 0.0, ep = 0.0;
      ws.check(bh_energies(ctx, particles.data(), particles.size(), opt.G, &ek, &ep));
      std::cout << "step " << step << tag << " E_K=" << ek << " E_P=" << ep << " E_T=" << (ek + ep) << "\n";
    };
    bh_step_timing total{0, 0, 0};
    auto report_timing = [&](int step, const bh_step_timing& t) {
      std::cout << "step " << step << " timing: tree=" << t.tree_build_ms << "ms traversal=" << t.traversal_ms
                << "ms total=" << (t.tree_build_ms + t.traversal_ms + t.exchange_ms) << "ms\n";
      total.tree_build_ms += t.tree_build_ms;
      total.traversal_ms += t.traversal_ms;
      total.exchange_ms += t.exchange_ms;
    };
    auto step = [&](double kick, double drift) {
      bh_step_timing t{};
      ws.check(bh_state_step(ctx, &fp, opt.max_leaf, opt.group_size, opt.target_size, kick, drift, &t));
      return t;
    };

    if (opt.steps > 0) {
      const auto run_start = std::chrono::steady_clock::now();
      bh_step_timing t = step(opt.dt / 2.0, opt.dt);
      if (opt.energy_every > 0) report(0, "");
      if (opt.time_steps) report_timing(0, t);
      if (opt.write_vtk) {
        download();
        write_vtk(opt.vtk_prefix + "_0.vtk", particles);
      }
      for (int s = 1; s < opt.steps; ++s) {
        t = step(opt.dt, opt.dt);
        if (opt.energy_every > 0 && s % opt.energy_every == 0) report(s, "");
        if (opt.time_steps) report_timing(s, t);
        if (opt.write_vtk) {
          download();
          write_vtk(opt.vtk_prefix + "_" + std::to_string(s) + ".vtk", particles);
        }
      }
      t = step(opt.dt / 2.0, 0.0);
      if (opt.energy_every > 0) report(opt.steps - 1, " (synced)");
      if (opt.time_steps) {
        total.tree_build_ms += t.tree_build_ms;
        total.traversal_ms += t.traversal_ms;
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - run_start).count();
        const int evals = opt.steps + 1;
        std::cout << "timing summary: evals=" << evals << " wall=" << wall << "ms tree=" << total.tree_build_ms << "ms (avg "
                  << total.tree_build_ms / evals << "ms) traversal=" << total.traversal_ms << "ms (avg "
                  << total.traversal_ms / evals << "ms) sum=" << (total.tree_build_ms + total.traversal_ms) << "ms\n";
      }
    }
    if (!opt.output_csv.empty()) {
      download();
      write_csv(opt.output_csv, particles);
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
