/*
 * nbody.h -- C ABI of the MI355X N-body engine (libnbody.so).
 *
 * Drop-in for the reference's headless simulation stages.  The reference has no
 * FFI: its "operator API" is the set of OpenCL kernels that CMyApp::Update
 * (src/MyApp.cpp:550-656) enqueues in a fixed order, wired by
 * CMyApp::SetKernelArgs (src/MyApp.cpp:245-325) and sized by
 * CMyApp::SetKernelConfig (src/MyApp.cpp:71-83).  Each nb_stage_* entry point
 * below replaces one of those enqueues; nb_sim_create/upload replace
 * InitParticles/InitBodyAttributes (src/MyApp.cpp:216-243, 331-504).
 *
 * Conventions: plain pointers and sizes; host arrays are AoS float[3*N]
 * positions/velocities and float[N] masses (the reference's cl_p/cl_v/cl_m
 * host layout); no exceptions cross the ABI; every call returns nb_status.
 * Stage calls are asynchronous on the simulation's HIP stream unless noted.
 * One nb_sim is not internally locked; distinct handles may be used from
 * distinct threads.
 */
#ifndef NBODY_NBODY_H
#define NBODY_NBODY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NB_ABI_VERSION 2

typedef enum nb_status {
    NB_OK = 0,
    NB_E_INVALID_ARG = 1,    /* bad config or argument (e.g. eps <= 0, mass <= 0) */
    NB_E_NO_DEVICE = 2,      /* no HIP device */
    NB_E_HIP = 3,            /* a HIP runtime call failed; see nb_last_error_detail() */
    NB_E_OUT_OF_MEMORY = 4,
    NB_E_TREE_OVERFLOW = 5,  /* cell pool exhausted (reference: build_tree.cl:109-114) */
    NB_E_DEPTH_OVERFLOW = 6, /* tree deeper than NB_MAX_LEVEL (reference: calculate_forces.cl:63-67) */
    NB_E_NONFINITE = 7,      /* NaN/Inf position */
    NB_E_STATE = 8,          /* stage called before its inputs exist */
    NB_E_ADMISSION = 9       /* ensemble: request does not fit the HBM budget */
} nb_status;

/* Enum values are those of src/SimulationConfig.h:10-38. */
enum { NB_LEAPFROG = 1, NB_EULER = 2 };
enum { NB_SPHERE_POS = 0, NB_UNIFORM_POS = 1 };
enum { NB_RANDOM_VEL = 0, NB_STARTING_OUT_VEL = 1, NB_STARTING_IN_VEL = 2, NB_FUNC_ZERO_VEL = 3,
       NB_TANGENT_XZ_VEL = 4 };
enum { NB_BARNES_HUT = 0, NB_BRUTE_FORCE_GLOBAL = 1, NB_BRUTE_FORCE_LOCAL = 2 };

/* Rule sets.  REFERENCE reproduces the tree shape and opening rule of the
 * reference kernels (root hop into slot 0, build_tree.cl:57-58; y half chosen by
 * the x comparison, build_tree.hcl:42; threshold (s/2)^2/theta + eps,
 * calculate_forces.cl:45-62).  STANDARD: cubic root, consistent halving,
 * criterion s/d < theta. */
enum { NB_RULES_REFERENCE = 0, NB_RULES_STANDARD = 1 };
/* Brute-force laws.  PLUMMER: m*d/(|d|^2+eps^2)^1.5.  REFERENCE_UNIT: the
 * reference brute-force law m*n/(|n|^2+e^2), n = normalize(d), e = 0.01 for
 * BRUTE_FORCE_GLOBAL (brute_force_update.cl:54-58) and 0.1 for
 * BRUTE_FORCE_LOCAL (brute_force_update.cl:114-120). */
enum { NB_LAW_PLUMMER = 0, NB_LAW_REFERENCE_UNIT = 1 };
/* Per-target accumulation: plain FP32 or Kahan-compensated FP32. */
enum { NB_ACCUM_PLAIN = 0, NB_ACCUM_KAHAN = 1 };

#define NB_MAX_LEVEL 64   /* deepest cell level (calculate_forces.hcl:1 MAX_DEPTH) */
#define NB_PATH_STRIDE 72 /* bytes per body in nb_export_bodies paths */
#define NB_NUM_STAGES 9

/* Stage indices for nb_stage_times. */
enum { NB_STAGE_BOUNDS = 0, NB_STAGE_RESET = 1, NB_STAGE_BUILD = 2, NB_STAGE_SUMMARIZE = 3,
       NB_STAGE_SORT = 4, NB_STAGE_FORCE = 5, NB_STAGE_UPDATE = 6, NB_STAGE_BRUTE = 7,
       NB_STAGE_COLLISION = 8 };

/* Stepping parameters: SimulationConfig + BarnesHutConfig fields that reach the
 * kernels (src/SimulationConfig.h:85-104, src/BarnesHutConfig.h:58-60), plus
 * the engine's own choices.  nb_config_default() gives the reference defaults. */
typedef struct nb_config {
    int32_t num_bodies;        /* SimulationConfig::num_of_bodies (20000) */
    int32_t algorithm;         /* NB_BARNES_HUT / NB_BRUTE_FORCE_GLOBAL / NB_BRUTE_FORCE_LOCAL */
    int32_t method;            /* NB_EULER (default) or NB_LEAPFROG (kick-drift-kick) */
    int32_t rules;             /* NB_RULES_REFERENCE (default) or NB_RULES_STANDARD */
    int32_t law;               /* brute force: NB_LAW_REFERENCE_UNIT (default) or NB_LAW_PLUMMER */
    int32_t accumulation;      /* NB_ACCUM_PLAIN (default) or NB_ACCUM_KAHAN */
    int32_t max_octree_depth;  /* BarnesHutConfig::max_octree_depth (32); kept, not used by kernels */
    int32_t cell_capacity;     /* 0: reference sizing max(2N, 1024*CUs) rounded up to 16 */
    float gravitational_constant; /* 6.67e-11 */
    float theta;               /* BarnesHutConfig::theta (1.5) */
    float epsilon;             /* BarnesHutConfig::epsilon (0.01) */
    float dt;                  /* time step (the reference derives it from frame time) */
    int32_t collision;         /* SimulationConfig::collision: 1 runs nb_stage_collision after the update */
    int32_t use_graph;         /* nb_step replays a captured HIP graph (default 1) */
    int32_t time_stages;       /* record per-stage HIP events (default 0) */
    int32_t reserved[5];
} nb_config;

typedef struct nb_sim nb_sim;

/* Tree summary after a build. */
typedef struct nb_tree_info {
    int32_t num_cells;   /* cells in use, root included */
    int32_t max_depth;   /* max over bodies of the level of the cell holding them (>= 1) */
    int32_t capacity;    /* cell pool size */
    int32_t status;      /* nb_status of the last step's device work */
} nb_tree_info;

void nb_config_default(nb_config* cfg);
const char* nb_status_string(nb_status s);
const char* nb_last_error_detail(void);
int32_t nb_abi_version(void);
/* Device description: compute units and name (name may be NULL). */
nb_status nb_device_info(int32_t* compute_units, char* name, size_t name_len);
/* Bodies per opening decision of the Barnes-Hut force kernel: the traversal is shared by
 * groups of this many consecutive bodies of the force order (nb_export_force_order), and
 * a cell is accepted for the
 * group only if it is far enough for every body of it (the reference votes over its
 * work-group, calculate_forces.cl:121). */
int32_t nb_force_vote_group(void);
/* Launch-shape tuning (no counterpart in the reference; set by the autotuner).  Process-wide;
 * applies to simulations, ensembles and servers whose batch is (re)described afterwards.  A
 * system's arithmetic depends on bf_min_waves (how finely its sources are split), so results
 * are bit-identical across batches only under the same tuning. */
typedef struct nb_tuning {
    int32_t bf_targets_per_lane; /* brute-force register blocking: 0 automatic, 1 or 2 */
    int32_t bf_min_waves;        /* brute force: split a system's sources until it yields this many
                                    waves (at most 16 waves x 4 lanes per target); default 8192 */
    int32_t reserved[6];
} nb_tuning;
void nb_tuning_default(nb_tuning* t);
nb_status nb_set_tuning(const nb_tuning* t);
nb_status nb_get_tuning(nb_tuning* t);
/* Cell pool size the reference would allocate for n bodies (src/MyApp.cpp:71-83). */
int32_t nb_reference_cell_capacity(int32_t n, int32_t compute_units);

/* Create/destroy.  `stream` is a hipStream_t or NULL (the library creates one). */
nb_status nb_sim_create(const nb_config* cfg, void* stream, nb_sim** out);
nb_status nb_sim_destroy(nb_sim* sim);
/* Host -> device state (InitBodyAttributes' writes, src/MyApp.cpp:357,416,489).
 * vel may be NULL (zero).  Synchronous. */
nb_status nb_sim_upload(nb_sim* sim, const float* pos_xyz, const float* vel_xyz, const float* mass);
/* Device -> host; any pointer may be NULL.  Synchronous. */
nb_status nb_sim_download(nb_sim* sim, float* pos_xyz, float* vel_xyz, float* acc_xyz);
nb_status nb_sim_sync(nb_sim* sim);
/* Status of device work queued so far (synchronises). */
nb_status nb_sim_status(nb_sim* sim);
nb_status nb_sim_clear_status(nb_sim* sim);
/* Elastic recovery from NB_E_TREE_OVERFLOW (no counterpart: the reference resets its
 * error word and the frame's tree is lost, build_tree.cl:109-114).  Re-creates the
 * simulation's arena with a cell pool of `cells` (1 <= cells < 2^24) and the same
 * bodies' positions, velocities, masses and accelerations; clear the status first,
 * then repeat the failed step.  Under NB_EULER a failed step leaves the state at the
 * last good step; under NB_LEAPFROG at the failed step's opening kick and drift.
 * Barnes-Hut only (NB_E_STATE otherwise).  If the new pool cannot be allocated the
 * old one is restored and the allocation error returned. */
nb_status nb_sim_set_cell_capacity(nb_sim* sim, int32_t cells);

/* Stages, in the order CMyApp::Update enqueues them (src/MyApp.cpp:583-598). */
/* hybrid_reduce_root + parallel_reduce_root (kernels/boundary_reduce.cl:6-159);
 * box (nullable) receives {minx,miny,minz,maxx,maxy,maxz} (synchronises if set). */
nb_status nb_stage_bounds(nb_sim* sim, float* box6);
/* init (kernels/init.cl:2-36). */
nb_status nb_stage_tree_reset(nb_sim* sim);
/* build_tree (kernels/build_tree.cl:6-165). */
nb_status nb_stage_tree_build(nb_sim* sim);
/* saturate_tree (kernels/saturate_tree.cl:6-146). */
nb_status nb_stage_summarize(nb_sim* sim);
/* sort (kernels/sort.cl, empty in the reference; intent at calculate_forces.cl:148). */
nb_status nb_stage_sort(nb_sim* sim);
/* calculate_force_local traversal (kernels/calculate_forces.cl:44-148); writes acc. */
nb_status nb_stage_force(nb_sim* sim);
/* Integrator (calculate_forces.cl:150-182): Euler v+=a dt, x+=v dt; under
 * NB_LEAPFROG the closing kick. */
nb_status nb_stage_update(nb_sim* sim);
/* update / update_local all-pairs force (kernels/brute_force_update.cl:3-228); writes acc. */
nb_status nb_stage_brute_force(nb_sim* sim);
/* One GPU's share of a force pass, for callers that shard the targets over G GPUs that
 * each hold the whole tree (rank r takes [r*N/G, (r+1)*N/G)).  The counterpart is the
 * same kernel as above (calculate_force_local, kernels/calculate_forces.cl:44-148;
 * update_local, kernels/brute_force_update.cl:108-228) run over a sub-range of its
 * NDRange; the reference itself never launches one.  Accelerations of the targets
 * [first, first+count) are written, every other body's are left as they were.
 * nb_stage_force_range counts targets in force order (nb_export_force_order); the
 * vote groups stay those of the full pass (64 consecutive bodies from position 0 of that
 * order), so every written acceleration equals the full pass's bit for bit, wherever the
 * range starts or ends.  nb_stage_brute_force_range counts targets in body index order; its
 * results also equal the full pass's.  A range outside [0, N] is NB_E_INVALID_ARG. */
nb_status nb_stage_force_range(nb_sim* sim, int64_t first, int64_t count);
nb_status nb_stage_brute_force_range(nb_sim* sim, int64_t first, int64_t count);
/* update_collision_local (kernels/collision_update.cl:3-83): bodies are spheres of radius
 * 0.015; each body's velocity is reflected about the separation of every other body within
 * 0.03, in index order.  O(N^2); reads positions, writes velocities. */
nb_status nb_stage_collision(nb_sim* sim);
/* nsteps whole steps (BH: bounds..update; BF: brute force + update).  The step counter
 * (nb_sim_step_count) advances only when every step succeeded.  Without stage timing,
 * brute-force simulations of up to 16384 bodies and no collision stage run the integrator
 * inside the brute-force kernel (one launch per Euler step; results bit-identical to the
 * separate stages). */
nb_status nb_step(nb_sim* sim, int32_t nsteps);

/* Tree inspection (synchronises).  Canonical forms: bodies in tree order with
 * their octant paths (NB_PATH_STRIDE bytes each, unused bytes 0xff), cells in
 * depth-first preorder.  Any pointer may be NULL. */
nb_status nb_tree_info_get(nb_sim* sim, nb_tree_info* info);
nb_status nb_export_bodies(nb_sim* sim, int32_t* sorted, uint8_t* paths, int32_t* path_len);
nb_status nb_export_cells(nb_sim* sim, int32_t* level, int32_t* slot, int32_t* count, float* com4);
/* The force order (N body indices).  STANDARD rules: depth-first like `sorted`, but each
 * cell's children taken along a 3-D Hilbert curve (Skilling's index on the octants), so
 * 64 consecutive bodies span less space than in `sorted`.  REFERENCE rules: `sorted`
 * itself (those cells are not boxes in y).  No counterpart in the reference, whose vote
 * groups are the bodies of a work-group in index order (calculate_forces.cl:84-148).  The
 * Barnes-Hut force pass forms its wave-wide opening votes from 64 consecutive bodies of
 * this order; nb_stage_force_range counts in it. */
nb_status nb_export_force_order(nb_sim* sim, int32_t* order);
/* Opening thresholds per level (NB_MAX_LEVEL+1 floats) of the current box. */
nb_status nb_export_dq(nb_sim* sim, float* dq);
/* Interactions the last nb_stage_force performed, summed over bodies (synchronises). */
nb_status nb_force_interactions(nb_sim* sim, int64_t* total);
/* Interactions of each body in the last force pass, N entries in body index order: Barnes-Hut,
 * the bodies and cells its vote group accepted for it; brute force, N - 1.  No counterpart in the
 * reference (it counts nothing); for parity checks on sampled vote groups.  Synchronises. */
nb_status nb_export_interactions(nb_sim* sim, int32_t* per_body);

/* Per-stage milliseconds of the last step (needs cfg.time_stages). */
nb_status nb_stage_times(nb_sim* sim, float* ms);
/* nsteps whole steps queued on the stream without waiting for them (no status
 * check; call nb_sim_sync / nb_sim_status afterwards). */
nb_status nb_step_async(nb_sim* sim, int32_t nsteps);

/* Vertex export, the equivalent of copy_vertices (kernels/copy_vertices.cl:3-17, which fills
 * the GL vertex buffer): (x, y, z, 1) per body, float[4*N].  dst_is_device = 0: dst is host
 * memory (synchronous); 1: dst is device memory, written on the simulation's stream. */
nb_status nb_export_vertices(nb_sim* sim, float* dst_xyzw, int32_t dst_is_device);

/* Checkpoints (no counterpart in the reference, which keeps no state across runs): a
 * self-describing byte blob = nb_checkpoint_header, then float pos[3N], vel[3N], mass[N],
 * acc[3N].  Loading needs a simulation of the same num_bodies; it restores the state, the
 * accelerations (so LEAPFROG continues without a new force pass) and the step counter. */
#define NB_CHECKPOINT_MAGIC 0x4b43424eu /* "NBCK" */
#define NB_CHECKPOINT_VERSION 1
typedef struct nb_checkpoint_header {
    uint32_t magic;
    uint32_t version;
    int32_t num_bodies;
    int32_t have_acc;   /* acc holds a(x) of the stored positions */
    int64_t step;       /* steps taken since the state was first uploaded */
    nb_config config;   /* the simulation's configuration when saved */
} nb_checkpoint_header;
nb_status nb_checkpoint_size(nb_sim* sim, size_t* bytes);
nb_status nb_checkpoint_save(nb_sim* sim, void* buf, size_t len);
nb_status nb_checkpoint_load(nb_sim* sim, const void* buf, size_t len);
/* Steps taken since the last upload or checkpoint load (plus the loaded counter). */
nb_status nb_sim_step_count(nb_sim* sim, int64_t* steps);

/* FP64 conservation diagnostics of the current state, computed on the GPU (no
 * counterpart in the reference).  out[NB_ENERGY_LEN] = kinetic, potential,
 * px, py, pz, Lx, Ly, Lz, total mass.  The pair potential is the one the
 * configured force law derives from: PLUMMER (and Barnes-Hut)
 * -G m_i m_j / sqrt(r^2 + eps^2); REFERENCE_UNIT +G m_i m_j r / (1 + ueps^2).
 * O(N^2) FP64 work; synchronous. */
#define NB_ENERGY_LEN 9
nb_status nb_sim_energy(nb_sim* sim, double* out);

/* nb_sim_energy without allocating: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).  Same
 * kernels, so the same results bit for bit as nb_sim_energy on the same state.  Inputs and
 * output are caller-owned DEVICE arrays: bodies_xyzm = float4 (x, y, z, m) per body and
 * vel_xyzw = float4 (vx, vy, vz, unused) per body, both 16-byte aligned; out =
 * NB_ENERGY_LEN doubles (layout of nb_sim_energy).  law = NB_LAW_PLUMMER (pair potential
 * -G m_i m_j / sqrt(r^2 + eps^2), eps > 0) or NB_LAW_REFERENCE_UNIT (+G m_i m_j r /
 * (1 + eps^2)).  Asynchronous on `stream` (NULL: the null stream).  The per-workgroup
 * partial sums live in a library-owned static device array sized for at most
 * NB_NOALLOC_MAX_BODIES bodies (4.5 MiB per device, held until the process ends); larger n
 * is NB_E_INVALID_ARG.  Every call on one device shares that array: calls that overlap on
 * one device race.  No hipMalloc, hipFree or hipMemset is called. */
#define NB_NOALLOC_MAX_BODIES (1 << 24)
nb_status nb_energy_f64_noalloc(const float* bodies_xyzm, const float* vel_xyzw, int32_t n, int32_t law,
                                double eps, double G, double* out, void* stream);

/* ---------------------------------------------------------------------------
 * Ensembles: many independent systems stepped together.  The reference steps
 * one system per frame (CMyApp::Update, src/MyApp.cpp:550-656); here every
 * stage kernel covers all systems of the batch in one launch (per-system
 * descriptors and block tables), and a whole ensemble step is one HIP graph.
 * Each system's results are bit-identical to stepping it alone.
 * ------------------------------------------------------------------------- */
typedef struct nb_system_spec {
    int32_t num_bodies;
    int32_t algorithm;      /* NB_BARNES_HUT / NB_BRUTE_FORCE_GLOBAL / NB_BRUTE_FORCE_LOCAL */
    int32_t rules;          /* Barnes-Hut rule set */
    int32_t cell_capacity;  /* 0: 2N + 1024 */
    float gravitational_constant;
    float theta;
    float epsilon;
    float dt;
    int32_t collision;      /* 1: the collision stage runs after this system's update */
} nb_system_spec;

typedef struct nb_ensemble nb_ensemble;

/* method, law and accumulation apply to every system of the ensemble. */
nb_status nb_ens_create(const nb_system_spec* specs, int32_t count, int32_t method, int32_t law,
                        int32_t accumulation, int32_t use_graph, void* stream, nb_ensemble** out);
nb_status nb_ens_destroy(nb_ensemble* ens);
nb_status nb_ens_upload(nb_ensemble* ens, int32_t system, const float* pos_xyz, const float* vel_xyz,
                        const float* mass);
nb_status nb_ens_download(nb_ensemble* ens, int32_t system, float* pos_xyz, float* vel_xyz, float* acc_xyz);
/* nsteps batched steps of every system; synchronises and reports the first failing status. */
nb_status nb_ens_step(nb_ensemble* ens, int32_t nsteps);
/* nb_export_interactions for one system of the ensemble. */
nb_status nb_ens_export_interactions(nb_ensemble* ens, int32_t system, int32_t* per_body);
/* per-system nb_status after the last step (count entries). */
nb_status nb_ens_status(nb_ensemble* ens, int32_t* status);
/* HBM bytes the ensemble holds. */
nb_status nb_ens_device_bytes(nb_ensemble* ens, int64_t* bytes);

/* ---------------------------------------------------------------------------
 * Serving: continuous batching of independent simulation requests.  No
 * counterpart in the reference (one system per process, src/main.cpp); built
 * on the ensemble kernels.  A server reserves one HBM arena from a byte budget.
 * A request is a system plus a step count.  Each nb_server_tick admits queued
 * requests in arrival order while a system slot and arena room are free, runs
 * one batched step of every resident system, and retires the ones that are
 * done (or failed), keeping their final state on the host until fetched.
 * A request's result is bit-identical to stepping its system alone.
 * Not thread-safe: one caller thread per server.
 * ------------------------------------------------------------------------- */
typedef struct nb_server nb_server;

typedef struct nb_server_stats {
    int32_t active;       /* systems resident in the arena */
    int32_t queued;       /* submitted, not yet admitted */
    int32_t finished;     /* retired, not yet fetched */
    int32_t reserved;
    int64_t arena_bytes;  /* HBM the server holds */
    int64_t body_steps;   /* sum over ticks of resident bodies stepped */
    int64_t ticks;        /* ticks that ran a step */
} nb_server_stats;

/* method, law and accumulation apply to every request. */
nb_status nb_server_create(int64_t hbm_budget_bytes, int32_t max_systems, int32_t method, int32_t law,
                           int32_t accumulation, void* stream, nb_server** out);
nb_status nb_server_destroy(nb_server* sv);
/* Queue a request (state is copied).  NB_E_ADMISSION if it can never fit the arena. */
nb_status nb_server_submit(nb_server* sv, const nb_system_spec* spec, int32_t steps, const float* pos_xyz,
                           const float* vel_xyz, const float* mass, int64_t* id);
/* Admit, step, retire; *completed = requests retired by this tick. */
nb_status nb_server_tick(nb_server* sv, int32_t* completed);
/* Oldest retired request not yet polled: its id and nb_status.  NB_E_STATE if none. */
nb_status nb_server_poll(nb_server* sv, int64_t* id, int32_t* status);
/* Final state of a retired request (any pointer may be NULL); frees it. */
nb_status nb_server_fetch(nb_server* sv, int64_t id, float* pos_xyz, float* vel_xyz, float* acc_xyz);
nb_status nb_server_stats_get(nb_server* sv, nb_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif
