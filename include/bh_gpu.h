/*
 * MI355X Barnes-Hut force path: C ABI.
 *
 * Drop-in for the reference's grouped, single-rank force path
 * (devreal/AMT-BH-benchmark, ttg-parsec/src): per-step Morton sort, flat
 * octree build with multipoles, grouped force walk, leapfrog kick/drift.
 * Plain C types only; all device state is owned by a bh_ctx.
 *
 * Reference interfaces replaced (paths relative to the reference repo):
 *   step_particles          ttg-parsec/src/integrate_ttg.h:66-148 (grouped branch, 1 rank: :88-107)
 *   build_flat_tree         ttg-parsec/src/flat_build_ttg.h:276-281
 *   run_grouped_pass        ttg-parsec/src/grouped_ttg.h:297-323
 *   brute_force_accel       ttg-parsec/src/brute_force.h:15-28
 *   kinetic/potential_energy ttg-parsec/src/integrate.h:27-43
 */
#ifndef BH_GPU_H
#define BH_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* bh::Particle (particle.h:7-24): id@0, pad, mass@8, pos@16, vel@40; 64 bytes. */
typedef struct bh_particle {
  uint32_t id;
  uint32_t pad;
  double mass;
  double pos[3];
  double vel[3];
} bh_particle;

/* bh::ForceParams (traversal_ttg.h:136-141). eps2 is the squared softening. */
typedef struct bh_force_params {
  double theta;
  double eps2;
  double G;
  int32_t use_quadrupole;
  int32_t pad;
} bh_force_params;

/* bh::StepTiming (integrate_ttg.h:24-28), milliseconds; exchange_ms is 0 on one GPU. */
typedef struct bh_step_timing {
  double tree_build_ms;
  double traversal_ms;
  double exchange_ms;
} bh_step_timing;

/* bh::WalkNode (flat_tree.h:61-70), 56 bytes; used by bh_tree_export. */
typedef struct bh_walk_node {
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
} bh_walk_node;

/* Status codes. The reference throws (integrate_ttg.h:104-106, main.cc:147);
 * no exception crosses this ABI: every call returns one of these and
 * bh_last_error() holds the message. */
enum bh_status {
  BH_OK = 0,
  BH_ERR_INVALID = 1,     /* bad argument (null pointer, group_size 0, ...) */
  BH_ERR_NONFINITE = 2,   /* a position, velocity or mass is NaN or infinite */
  BH_ERR_INCOMPLETE = 3,  /* the walk met an interior node it could not open (no children, or deeper
                             than the walk's stack); cannot occur on a tree this library built */
  BH_ERR_OOM = 4,         /* device allocation failed */
  BH_ERR_HIP = 5,         /* HIP runtime error */
  BH_ERR_TOO_LARGE = 6,   /* n or max_leaf beyond the supported range */
  BH_ERR_NO_TREE = 7      /* bh_accel / bh_tree_export before bh_build_tree */
};

/* Walk arithmetic. FP32: interactions in FP32, each batch summed in FP32
 * and the batch sums accumulated per particle in FP64 (default). FP64:
 * interactions in FP64. The MAC, tree build, multipoles and kick/drift are
 * FP64 in both. */
enum bh_precision { BH_FP32 = 0, BH_FP64 = 1 };

/* Node numbering of bh_tree_export: the reference's (backbone BFS, then each
 * subtree of <= 4096 particles BFS, flat_build_ttg.h:194-273) or this
 * library's own level-order BFS. */
enum bh_layout { BH_LAYOUT_REFERENCE = 0, BH_LAYOUT_BFS = 1 };

typedef struct bh_ctx bh_ctx;

/* Context on HIP device `device`: one stream, grow-only device buffers (the
 * GroupedWorkspace of integrate_ttg.h:61-64). Not re-entrant: use one
 * context per host thread. */
int bh_ctx_create(bh_ctx** out, int device);
void bh_ctx_destroy(bh_ctx* ctx);
const char* bh_last_error(const bh_ctx* ctx);
int bh_set_precision(bh_ctx* ctx, int precision);
/* Walk schedule (results are the same either way: the interaction lists
 * are identical, sums may round differently): BH_WALK_GROUPED (default) walks
 * once per group of up to 256 particles for the decisions its targets
 * share, then per target; BH_WALK_PER_TARGET walks from the root for
 * every target. */
typedef enum bh_walk_mode { BH_WALK_GROUPED = 0, BH_WALK_PER_TARGET = 1 } bh_walk_mode;
int bh_set_walk_mode(bh_ctx* ctx, int mode);
/* Graph mode (no reference counterpart; the reference's step is a TTG task
 * graph rebuilt every call). A step makes no host decision between its first
 * and last kernel, so with BH_GRAPH_ON every bh_step* call runs its launches
 * as one HIP graph: captured the first time a given launch sequence occurs
 * (same buffers, particle count, system count and table size, precision,
 * walk mode and planned tree shape) and replayed after that; the system
 * table (per-system parameters, kick and drift) is re-read on every replay.
 * Results are bit-identical to BH_GRAPH_OFF (the default). The context keeps
 * the 256 most recently used graphs; switching the mode off releases them.
 * bh_graph_count reports how many are held. */
typedef enum bh_graph_mode { BH_GRAPH_OFF = 0, BH_GRAPH_ON = 1 } bh_graph_mode;
int bh_set_graph_mode(bh_ctx* ctx, int mode);
/* Walk tile (no reference counterpart): the most particles one walk unit
 * (one wavefront's shared walk) holds, BH_MIN_WALK_TILE..BH_MAX_WALK_TILE,
 * default BH_MAX_WALK_TILE. Groups of the reference's group_size are split
 * into units of at most this many particles, and runs of small sibling
 * targets are packed into units up to it. Results do not depend on it (every
 * target gets its own walk's interaction list either way; sums may round
 * differently); smaller tiles give small systems more wavefronts, larger
 * ones share more of the walk. Applies to the single-system entry points
 * and to batch systems whose walk_tile is 0. */
#define BH_MIN_WALK_TILE 16u
#define BH_MAX_WALK_TILE 128u
int bh_set_walk_tile(bh_ctx* ctx, uint32_t tile);
/* Diagnostic (no reference counterpart): caps the walk's stack at `levels`
 * frames (1..22; 22, the default, is every level a tree can have), so that
 * a test can make the walk meet nodes it may not open and check the
 * BH_ERR_INCOMPLETE report (the reference's missing > 0,
 * integrate_ttg.h:104-106); the kernel still finishes normally. */
int bh_debug_walk_stack_limit(bh_ctx* ctx, uint32_t levels);
int bh_graph_count(bh_ctx* ctx, uint32_t* count);
/* Limits: n < 2^30 particles, max_leaf <= 1024. */
uint64_t bh_max_particles(void);

/* step_particles (integrate_ttg.h:66-69), grouped branch on one rank:
 * builds the tree of particles[0..n), walks it and applies
 * v += kick_dt * a; if (drift_dt != 0) x += drift_dt * v
 * to every particle, and returns them IN MORTON ORDER in `particles` (as the
 * reference does, integrate_ttg.h:107). Host memory in and out.
 * group_size must be > 0 (0 selects the reference's per-node path, which
 * this library does not provide). `timing` may be NULL. */
int bh_step(bh_ctx* ctx, bh_particle* particles, uint64_t n, const bh_force_params* fp, uint32_t max_leaf,
            uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt, bh_step_timing* timing);

/* The same on device memory: reads d_in[0..n), writes the Morton-ordered,
 * updated particles to d_out[0..n). d_in and d_out must not overlap; d_in is
 * never written. The work is ordered on the context's stream and the host
 * waits once, when it is complete. */
int bh_step_device(bh_ctx* ctx, const bh_particle* d_in, bh_particle* d_out, uint64_t n, const bh_force_params* fp,
                   uint32_t max_leaf, uint32_t group_size, uint32_t target_size, double kick_dt, double drift_dt,
                   bh_step_timing* timing);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). bh_step_device for n <= BH_NOALLOC_MAX_N particles with every
 * intermediate array (keys, tree, walk scratch) cut from one fixed
 * BH_NOALLOC_SCRATCH_MIB MiB device array that the library holds per GPU
 * from its first context on that GPU until the process ends; the call
 * allocates nothing. Same kernels, same results as bh_step_device, byte for
 * byte. Larger n, or a tree with more nodes than the fixed array holds,
 * returns BH_ERR_TOO_LARGE. Every call on one GPU uses that GPU's single
 * array: calls that overlap on one GPU (other threads or contexts) race and
 * must be serialised by the caller. The call does not leave a tree for
 * bh_tree_export / bh_accel. */
#define BH_NOALLOC_MAX_N 65536u
#define BH_NOALLOC_SCRATCH_MIB 64u
int bh_step_device_noalloc(bh_ctx* ctx, const bh_particle* d_in, bh_particle* d_out, uint64_t n,
                           const bh_force_params* fp, uint32_t max_leaf, uint32_t group_size, uint32_t target_size,
                           double kick_dt, double drift_dt, bh_step_timing* timing);

/* One independent system of a batch: n particles (consecutive in the batch
 * arrays, in table order) and its own step_particles arguments. */
typedef struct bh_system {
  uint64_t n;
  bh_force_params fp;
  uint32_t max_leaf;
  uint32_t group_size;
  uint32_t target_size;
  uint32_t walk_tile; /* 0: the context's walk tile (bh_set_walk_tile) */
  double kick_dt;
  double drift_dt;
} bh_system;

/* step_particles for many independent systems in the same launches (no
 * reference counterpart: the batched path of one GPU). d_in holds the
 * systems' particles back to back in table order; d_out receives each
 * system's updated particles in its own Morton order, in the same ranges.
 * Every system's result is bit-identical to a bh_step_device call on it
 * alone. At most 65,535 systems. */
int bh_step_batch(bh_ctx* ctx, const bh_particle* d_in, bh_particle* d_out, const bh_system* systems,
                  uint32_t n_systems, bh_step_timing* timing);
/* The same on host memory, in place. */
int bh_step_batch_host(bh_ctx* ctx, bh_particle* particles, const bh_system* systems, uint32_t n_systems,
                       bh_step_timing* timing);

/* ---- Continuous-batching server (no reference counterpart) ----
 * Many independent leapfrog runs share the GPU: each server step admits
 * queued requests (in submission order) while the active systems'
 * footprints stay within the memory cap, runs one force evaluation of every
 * active system in one bh_step_batch, and retires the systems whose
 * schedule is complete. A request runs main.cc:199-246's schedule: (dt/2,
 * dt), steps - 1 times (dt, dt), then (dt/2, 0); its result (Morton order)
 * is bit-identical to running that schedule alone with bh_step. One thread
 * per server; the server steps on its context's stream. */
typedef struct bh_server bh_server;
typedef struct bh_request {
  uint64_t n;
  bh_force_params fp;
  uint32_t max_leaf;
  uint32_t group_size;
  uint32_t target_size;
  uint32_t steps; /* leapfrog steps: steps + 1 force evaluations */
  double dt;
} bh_request;
typedef struct bh_completion {
  uint64_t ticket;
  double submit_s, admit_s, done_s; /* server clock, seconds */
  uint32_t evaluations;
  int32_t status;
} bh_completion;
typedef struct bh_server_stats {
  uint32_t active_systems; /* after the step */
  uint32_t pending;
  uint32_t admitted;
  uint32_t completed;
  uint32_t stepped_systems;   /* systems evaluated by this step */
  uint32_t pad;
  uint64_t active_particles;
  uint64_t stepped_particles; /* particles evaluated by this step */
  uint64_t used_bytes;        /* footprint of the active systems */
  double step_ms;
} bh_server_stats;

int bh_server_create(bh_server** out, bh_ctx* ctx, uint64_t mem_cap_bytes, uint32_t max_systems);
void bh_server_destroy(bh_server* server);
const char* bh_server_last_error(const bh_server* server);
/* Device bytes a request is counted for against the cap. */
uint64_t bh_server_footprint(const bh_request* request);
/* Queues a request; `particles` is copied at once, `result` (n particles)
 * must stay valid until the request's completion is polled. */
int bh_server_submit(bh_server* server, const bh_particle* particles, bh_particle* result, const bh_request* request,
                     uint64_t* ticket);
/* One admission + batched evaluation + retirement round. */
int bh_server_step(bh_server* server, bh_server_stats* stats);
/* Completed requests since the last poll, oldest first. */
int bh_server_poll(bh_server* server, bh_completion* out, uint32_t max, uint32_t* count);
int bh_server_pending(const bh_server* server, uint32_t* pending, uint32_t* active);
double bh_server_clock(const bh_server* server);

/* Particles resident in device memory across steps (the caller's
 * GroupedWorkspace reuse, main.cc:203-242, without host copies). A resident
 * step's walk also keeps the new positions in a compact array, from which
 * the next resident step builds its tree when nothing else ran in between. */
int bh_state_upload(bh_ctx* ctx, const bh_particle* host, uint64_t n);
int bh_state_step(bh_ctx* ctx, const bh_force_params* fp, uint32_t max_leaf, uint32_t group_size, uint32_t target_size,
                  double kick_dt, double drift_dt, bh_step_timing* timing);
/* One bh_step_batch over the resident particles (the systems' sizes add up
 * to the resident count). */
int bh_state_step_batch(bh_ctx* ctx, const bh_system* systems, uint32_t n_systems, bh_step_timing* timing);
int bh_state_download(bh_ctx* ctx, bh_particle* host, uint64_t n);
/* Device address of the resident particles (Morton order after a step).
 * The caller may write through it: the next resident step then reads the
 * particle records again. */
bh_particle* bh_state_device_ptr(bh_ctx* ctx);

/* build_flat_tree (flat_build_ttg.h:276-281) of host particles; the tree stays
 * in the context for bh_tree_export / bh_accel. */
int bh_build_tree(bh_ctx* ctx, const bh_particle* host, uint64_t n, uint32_t max_leaf, uint32_t group_size);
/* Sizes of the last built tree; dom4 (nullable) receives centre xyz and half width. */
int bh_tree_sizes(bh_ctx* ctx, uint64_t* n_nodes, uint64_t* n_groups, double* dom4);
/* Copies the last built tree out, nodes numbered per `layout` (bh_layout).
 * keys[n], perm[n] (sorted position -> input index), sorted[n] (particles in
 * Morton order), nodes[n_nodes], center[3*n_nodes], iab[6*n_nodes]
 * (xx,yy,zz,xy,xz,yz about the centre of mass), groups[n_groups] (group
 * roots in the reference's order). Any pointer may be NULL. */
int bh_tree_export(bh_ctx* ctx, int layout, uint64_t* keys, uint32_t* perm, bh_particle* sorted, bh_walk_node* nodes,
                   double* center, double* iab, uint32_t* groups);
/* run_grouped_pass (grouped_ttg.h:297-298) over the last built tree with the
 * update replaced by output: acc[3*n] = G * acceleration of each particle in
 * Morton order; n_pp[n] / n_pc[n] = length of its particle(+monopole) list and
 * its quadrupole-cell count; totals[2] = sum over particles of each. Any
 * output may be NULL. */
int bh_accel(bh_ctx* ctx, const bh_force_params* fp, uint32_t target_size, double* acc, uint32_t* n_pp, uint32_t* n_pc,
             uint64_t* totals);

/* Per-phase device times of the last step or bh_accel call, from HIP events
 * recorded on the context's stream (the reference's StepTiming split,
 * integrate_ttg.h:24-28, refined), and the work the walk did. Has no
 * reference counterpart beyond StepTiming; it exists for measurement. */
typedef struct bh_step_stats {
  double sort_ms;     /* bounds, Morton keys, radix sort, gather to Morton order */
  double tree_ms;     /* topology, geometry, multipoles */
  double targets_ms;  /* walk-target selection */
  double walk_ms;     /* the walk kernel alone, fused kick/drift included */
  uint64_t n_pp;      /* particle (and monopole) interactions evaluated */
  uint64_t n_pc;      /* quadrupole-cell interactions evaluated */
  uint64_t n_nodes;
  uint64_t n_leaves;
  uint64_t n_targets;
  double step_ms;     /* the whole step on the device */
  uint64_t n_units;   /* walk units (groups, packed runs of small groups, or targets) */
  uint32_t graph;     /* 1: the step ran as a graph replay; the four phase times are then 0 */
  uint32_t reruns;    /* times the step ran again because the tree outgrew the planned arrays or levels */
} bh_step_stats;
int bh_last_step_stats(bh_ctx* ctx, bh_step_stats* stats);

/* brute_force_accel (brute_force.h:15-28) for the targets idx[0..m) among
 * host particles[0..n), FP64 on the device; acc[3*m]. idx NULL = all. */
int bh_direct_sum(bh_ctx* ctx, const bh_particle* host, uint64_t n, const uint32_t* idx, uint64_t m, double eps2, double G,
                  double* acc);

/* kinetic_energy and potential_energy (integrate.h:27-43) of host particles
 * [0..n), FP64 on the device: E_K = sum m v^2 / 2; E_P = -G sum_{i<j} m_i m_j / r_ij
 * (unsoftened, O(n^2)). Either output may be NULL (E_P is then not computed). */
int bh_energies(bh_ctx* ctx, const bh_particle* host, uint64_t n, double G, double* kinetic, double* potential);

/* read_csv (io.h:10-14, io.cc:94-178): particles "id,mass,x,y,z,vx,vy,vz", one
 * per line, optional header line. *out is allocated with malloc (free it
 * with bh_free); err (nullable) receives the message of a failure, which
 * names the first malformed line. Host only. */
int bh_read_csv(const char* path, bh_particle** out, uint64_t* n, char* err, size_t err_len);
void bh_free(void* p);
/* write_vtk (io.h:16-18, io.cc:180-195): legacy VTK ASCII PolyData with
 * points, a mass scalar and a velocity vector. Host only. */
int bh_write_vtk(const char* path, const bh_particle* particles, uint64_t n, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* BH_GPU_H */
