// nb_api.hip -- the C ABI of include/nbody/nbody.h over nb::Engine (a batch of one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "nb_engine.h"
#include "nbody/nbody.h"

namespace nb {
const char* last_error();
}

using nb::Engine;
using nb::SysSpec;
using nb::SysState;

struct nb_sim {
    nb_config cfg;
    Engine eng;
    int n = 0;
    int cap = 0;
    bool have_state = false;
    bool have_box = false;
    bool tree_clean = true;
    bool have_tree = false;
    bool have_summary = false;
    bool have_sorted = false;
    bool have_acc = false;
    hipGraphExec_t graph = nullptr;  // one step
    hipGraph_t graph_def = nullptr;
    hipGraphExec_t fgraph[2] = {};    // fused brute-force steps: one step, two steps
    hipGraph_t fgraph_def[2] = {};
    float4* pos_b = nullptr;          // second position buffer of fused Euler steps (allocated at create)
    std::vector<hipEvent_t> ev;  // kEvPerStep per timed step
    float stage_ms[NB_NUM_STAGES] = {};
    double* energy_buf = nullptr;  // per-workgroup partial sums + result of nb_sim_energy
    int64_t steps = 0;             // steps since upload (or the loaded checkpoint's count)
};

#define NB_HIPC(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            nb::set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
            return e_ == hipErrorOutOfMemory ? NB_E_OUT_OF_MEMORY : NB_E_HIP;          \
        }                                                                              \
    } while (0)

static nb_status fail(nb_status s, const std::string& msg) {
    nb::set_error(msg);
    return s;
}

extern "C" {

int32_t nb_abi_version(void) { return NB_ABI_VERSION; }

int32_t nb_force_vote_group(void) { return nb::kForceVote; }

const char* nb_last_error_detail(void) { return nb::last_error(); }

const char* nb_status_string(nb_status s) {
    switch (s) {
        case NB_OK: return "ok";
        case NB_E_INVALID_ARG: return "invalid argument";
        case NB_E_NO_DEVICE: return "no HIP device";
        case NB_E_HIP: return "HIP runtime error";
        case NB_E_OUT_OF_MEMORY: return "out of device memory";
        case NB_E_TREE_OVERFLOW: return "cell pool overflow";
        case NB_E_DEPTH_OVERFLOW: return "tree depth overflow";
        case NB_E_NONFINITE: return "non-finite position";
        case NB_E_STATE: return "stage called out of order";
        case NB_E_ADMISSION: return "request does not fit the memory budget";
    }
    return "unknown status";
}

void nb_config_default(nb_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    // src/SimulationConfig.h:85-104, src/BarnesHutConfig.h:58-60
    c->num_bodies = 20000;
    c->algorithm = NB_BRUTE_FORCE_GLOBAL;
    c->method = NB_EULER;
    c->rules = NB_RULES_REFERENCE;
    c->law = NB_LAW_REFERENCE_UNIT;
    c->accumulation = NB_ACCUM_PLAIN;
    c->max_octree_depth = 32;
    c->cell_capacity = 0;
    c->gravitational_constant = 6.67e-11f;
    c->theta = 1.5f;
    c->epsilon = 0.01f;
    c->dt = 1e-3f;
    c->collision = 0;
    c->use_graph = 1;
    c->time_stages = 0;
}

// nodes = max(2N, 1024*CUs) rounded up to a multiple of 16 (src/MyApp.cpp:71-83).
int32_t nb_reference_cell_capacity(int32_t n, int32_t compute_units) {
    long long nodes = 2LL * n;
    long long floor_nodes = 1024LL * compute_units;
    if (nodes < floor_nodes) nodes = floor_nodes;
    while (nodes & 15) ++nodes;
    return (int32_t)std::min<long long>(nodes, 0x7fffffff);
}

nb_status nb_device_info(int32_t* compute_units, char* name, size_t name_len) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NB_E_NO_DEVICE, "no HIP device");
    int dev = 0;
    NB_HIPC(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    NB_HIPC(hipGetDeviceProperties(&prop, dev));
    if (compute_units) *compute_units = prop.multiProcessorCount;
    if (name && name_len) {
        std::snprintf(name, name_len, "%s", prop.gcnArchName);
    }
    return NB_OK;
}

static nb_status validate(const nb_config* c) {
    if (!c) return fail(NB_E_INVALID_ARG, "config is NULL");
    if (c->num_bodies < 0) return fail(NB_E_INVALID_ARG, "num_bodies < 0");
    if (c->algorithm < 0 || c->algorithm > 2) return fail(NB_E_INVALID_ARG, "bad algorithm");
    if (c->method != NB_EULER && c->method != NB_LEAPFROG) return fail(NB_E_INVALID_ARG, "bad method");
    if (c->rules != NB_RULES_REFERENCE && c->rules != NB_RULES_STANDARD) return fail(NB_E_INVALID_ARG, "bad rules");
    if (c->law != NB_LAW_PLUMMER && c->law != NB_LAW_REFERENCE_UNIT) return fail(NB_E_INVALID_ARG, "bad law");
    if (c->accumulation != NB_ACCUM_PLAIN && c->accumulation != NB_ACCUM_KAHAN)
        return fail(NB_E_INVALID_ARG, "bad accumulation");
    if (c->collision != 0 && c->collision != 1) return fail(NB_E_INVALID_ARG, "collision must be 0 or 1");
    if (!std::isfinite(c->dt) || !std::isfinite(c->gravitational_constant) || !std::isfinite(c->theta))
        return fail(NB_E_INVALID_ARG, "non-finite dt, G or theta");
    const bool needs_eps = c->algorithm == NB_BARNES_HUT || c->law == NB_LAW_PLUMMER;
    if (needs_eps && !(c->epsilon > 0.0f && std::isfinite(c->epsilon)))
        return fail(NB_E_INVALID_ARG, "epsilon must be > 0 (self-interaction is finite only then)");
    if (c->cell_capacity < 0) return fail(NB_E_INVALID_ARG, "cell_capacity < 0");
    return NB_OK;
}

static SysSpec spec_of(const nb_config* c, int cap) {
    SysSpec s;
    s.n = c->num_bodies;
    s.cap = cap;
    s.rules = c->rules;
    s.algo = c->algorithm;
    s.law = c->law;
    s.accum = c->accumulation;
    s.method = c->method;
    s.G = c->gravitational_constant;
    s.theta = c->theta;
    s.eps = c->epsilon;
    s.dt = c->dt;
    s.collision = c->collision;
    return s;
}

static bool is_bh(const nb_sim* s) { return s->cfg.algorithm == NB_BARNES_HUT; }

// Untimed brute-force steps of a small system without collision fuse the
// integrator into the brute-force kernel (one launch per Euler step instead of
// two).  Euler moves positions, which the kernel still reads as sources, so
// fused Euler steps alternate between node4 and pos_b: two steps return to
// node4, a single step copies back with a kernel.  Leapfrog's opening kick and
// drift stay a launch of their own; the closing kick is fused.  Results are
// bit-identical to the separate launches.
static bool fusable(const nb_sim* sim) {
    return !is_bh(sim) && !sim->cfg.collision && sim->n > 0 && sim->n <= nb::kBfFuseMax;
}

static void drop_graph(nb_sim* sim) {
    if (sim->graph) (void)hipGraphExecDestroy(sim->graph);
    if (sim->graph_def) (void)hipGraphDestroy(sim->graph_def);
    sim->graph = nullptr;
    sim->graph_def = nullptr;
    for (int k = 0; k < 2; ++k) {
        if (sim->fgraph[k]) (void)hipGraphExecDestroy(sim->fgraph[k]);
        if (sim->fgraph_def[k]) (void)hipGraphDestroy(sim->fgraph_def[k]);
        sim->fgraph[k] = nullptr;
        sim->fgraph_def[k] = nullptr;
    }
}

nb_status nb_sim_create(const nb_config* cfg, void* stream, nb_sim** out) {
    if (!out) return fail(NB_E_INVALID_ARG, "out is NULL");
    *out = nullptr;
    nb_status st = validate(cfg);
    if (st) return st;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NB_E_NO_DEVICE, "no HIP device");
    nb_sim* sim = new nb_sim();
    sim->cfg = *cfg;
    sim->n = cfg->num_bodies;
    int cus = 0, dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    sim->cap = cfg->algorithm == NB_BARNES_HUT
                   ? (cfg->cell_capacity > 0 ? cfg->cell_capacity : nb_reference_cell_capacity(sim->n, cus))
                   : 0;
    st = sim->eng.init({spec_of(cfg, sim->cap)}, stream);
    // the fused Euler steps' second position buffer exists for the simulation's whole
    // life, so that stepping calls never allocate
    if (!st && fusable(sim) && cfg->method == NB_EULER) {
        const hipError_t e = hipMalloc(&sim->pos_b, sizeof(float4) * (size_t)sim->n);
        if (e != hipSuccess) {
            sim->pos_b = nullptr;
            st = fail(e == hipErrorOutOfMemory ? NB_E_OUT_OF_MEMORY : NB_E_HIP,
                      std::string("hipMalloc(second position buffer): ") + hipGetErrorString(e));
        }
    }
    if (st) {
        delete sim;
        return st;
    }
    *out = sim;
    return NB_OK;
}

nb_status nb_sim_destroy(nb_sim* sim) {
    if (!sim) return NB_OK;
    if (sim->eng.stream) (void)hipStreamSynchronize(sim->eng.stream);
    drop_graph(sim);
    for (auto& e : sim->ev)
        if (e) (void)hipEventDestroy(e);
    if (sim->energy_buf) (void)hipFree(sim->energy_buf);
    if (sim->pos_b) (void)hipFree(sim->pos_b);
    delete sim;
    return NB_OK;
}

nb_status nb_sim_upload(nb_sim* sim, const float* pos, const float* vel, const float* mass) {
    if (!sim || (sim->n > 0 && (!pos || !mass))) return fail(NB_E_INVALID_ARG, "NULL argument");
    const int n = sim->n;
    std::vector<float4> h4(std::max(n, 1)), hv(std::max(n, 1));
    for (int i = 0; i < n; ++i) {
        const float m = mass[i];
        const bool ok = std::isfinite(m) && (sim->cfg.algorithm == NB_BARNES_HUT ? m > 0.0f : m >= 0.0f);
        if (!ok) return fail(NB_E_INVALID_ARG, "mass " + std::to_string(i) + " must be finite and > 0");
        h4[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], m);
        hv[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f) : make_float4(0, 0, 0, 0);
    }
    Engine& e = sim->eng;
    if (n > 0) {
        NB_HIPC(hipMemcpyAsync(e.node4 + e.hdesc[0].node_off, h4.data(), sizeof(float4) * n, hipMemcpyHostToDevice,
                               e.stream));
        NB_HIPC(hipMemcpyAsync(e.vel4, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, e.stream));
        NB_HIPC(hipMemsetAsync(e.acc4, 0, sizeof(float4) * n, e.stream));
    }
    NB_HIPC(hipStreamSynchronize(e.stream));
    sim->have_state = true;
    sim->have_box = sim->have_tree = sim->have_summary = sim->have_sorted = sim->have_acc = false;
    sim->steps = 0;
    return NB_OK;
}

nb_status nb_sim_download(nb_sim* sim, float* pos, float* vel, float* acc) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    const int n = sim->n;
    if (n == 0) return NB_OK;
    Engine& e = sim->eng;
    std::vector<float4> h(n);
    auto get = [&](const float4* src, float* dst) -> nb_status {
        if (!dst) return NB_OK;
        NB_HIPC(hipMemcpyAsync(h.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost, e.stream));
        NB_HIPC(hipStreamSynchronize(e.stream));
        for (int i = 0; i < n; ++i) {
            dst[3 * i] = h[i].x;
            dst[3 * i + 1] = h[i].y;
            dst[3 * i + 2] = h[i].z;
        }
        return NB_OK;
    };
    nb_status st = get(e.node4 + e.hdesc[0].node_off, pos);
    if (!st) st = get(e.vel4, vel);
    if (!st) st = get(e.acc4, acc);
    return st;
}

nb_status nb_sim_sync(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    NB_HIPC(hipStreamSynchronize(sim->eng.stream));
    return NB_OK;
}

static nb_status status_of_bits(int bits) {
    if (bits & nb::kStStall) return fail(NB_E_HIP, "a bounded device wait ran out (inconsistent tree)");
    if (bits & nb::kStNonfinite) return fail(NB_E_NONFINITE, "non-finite body position");
    if (bits & nb::kStDepthOverflow) return fail(NB_E_DEPTH_OVERFLOW, "tree deeper than NB_MAX_LEVEL");
    if (bits & nb::kStTreeOverflow) return fail(NB_E_TREE_OVERFLOW, "cell pool exhausted");
    return NB_OK;
}

nb_status nb_sim_status(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    std::vector<SysState> s;
    nb_status st = sim->eng.download_state(s);
    if (st) return st;
    return status_of_bits(s[0].status);
}

nb_status nb_sim_clear_status(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    NB_HIPC(hipStreamSynchronize(sim->eng.stream));
    NB_HIPC(hipMemset(&sim->eng.d_state[0].status, 0, sizeof(int)));
    return NB_OK;
}

nb_status nb_sim_set_cell_capacity(nb_sim* sim, int32_t cells) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    if (!is_bh(sim)) return fail(NB_E_STATE, "the cell pool belongs to Barnes-Hut simulations");
    if (cells < 1 || cells >= (1 << 24)) return fail(NB_E_INVALID_ARG, "cell capacity must be in [1, 2^24)");
    Engine& e = sim->eng;
    NB_HIPC(hipStreamSynchronize(e.stream));
    // the bodies' state goes through the host; the new arena starts with an empty pool
    const int n = sim->n;
    std::vector<float4> x(std::max(n, 1)), v(std::max(n, 1)), a(std::max(n, 1));
    if (n > 0) {
        NB_HIPC(hipMemcpy(x.data(), e.node4 + e.hdesc[0].node_off, sizeof(float4) * n, hipMemcpyDeviceToHost));
        NB_HIPC(hipMemcpy(v.data(), e.vel4, sizeof(float4) * n, hipMemcpyDeviceToHost));
        NB_HIPC(hipMemcpy(a.data(), e.acc4, sizeof(float4) * n, hipMemcpyDeviceToHost));
    }
    void* user_stream = e.own_stream ? nullptr : (void*)e.stream;
    drop_graph(sim);
    e.release();
    nb_config c = sim->cfg;
    c.cell_capacity = cells;
    nb_status st = e.init({spec_of(&c, cells)}, user_stream);
    if (st) {  // keep a usable simulation: the old pool size, the staged state
        const std::string why = nb::last_error();
        e.release();
        if (e.init({spec_of(&sim->cfg, sim->cap)}, user_stream) != NB_OK) {
            sim->have_state = false;
            return fail(st, "cell pool resize failed (" + why + ") and the old pool could not be restored");
        }
    } else {
        sim->cfg = c;
        sim->cap = cells;
    }
    if (n > 0) {
        NB_HIPC(hipMemcpy(e.node4 + e.hdesc[0].node_off, x.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        NB_HIPC(hipMemcpy(e.vel4, v.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
        NB_HIPC(hipMemcpy(e.acc4, a.data(), sizeof(float4) * n, hipMemcpyHostToDevice));
    }
    sim->have_box = sim->have_tree = sim->have_summary = sim->have_sorted = false;
    sim->tree_clean = true;
    return st;
}

static nb_status need(bool cond, const char* what) {
    return cond ? NB_OK : fail(NB_E_STATE, std::string("stage order: ") + what);
}

nb_status nb_stage_bounds(nb_sim* sim, float* box6) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_state && is_bh(sim), "bounds needs uploaded state and Barnes-Hut");
    if (st) return st;
    sim->eng.bounds();
    sim->have_box = true;
    if (box6) {
        std::vector<SysState> s;
        st = sim->eng.download_state(s);
        if (st) return st;
        std::memcpy(box6, s[0].box, sizeof(float) * 6);
        return status_of_bits(s[0].status);
    }
    return NB_OK;
}

nb_status nb_stage_tree_reset(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(is_bh(sim), "reset needs Barnes-Hut");
    if (st) return st;
    sim->eng.tree_reset();
    sim->tree_clean = true;
    sim->have_tree = sim->have_summary = sim->have_sorted = false;
    return NB_OK;
}

nb_status nb_stage_tree_build(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_box && sim->tree_clean, "build needs bounds and a reset tree");
    if (st) return st;
    sim->eng.build();
    sim->tree_clean = false;
    sim->have_tree = true;
    return NB_OK;
}

nb_status nb_stage_summarize(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_tree, "summarize needs a built tree");
    if (st) return st;
    sim->eng.summarize();
    sim->have_summary = true;
    return NB_OK;
}

nb_status nb_stage_sort(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_summary, "sort needs a summarized tree");
    if (st) return st;
    sim->eng.sort();
    sim->have_sorted = true;
    return NB_OK;
}

nb_status nb_stage_force(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_sorted, "force needs a sorted tree");
    if (st) return st;
    sim->eng.force();
    sim->have_acc = true;
    return NB_OK;
}

nb_status nb_stage_brute_force(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_state && !is_bh(sim), "brute force needs uploaded state and a brute-force algorithm");
    if (st) return st;
    sim->eng.brute();
    sim->have_acc = true;
    return NB_OK;
}

static nb_status check_range(const nb_sim* sim, int64_t first, int64_t count) {
    if (first < 0 || count < 0 || first > sim->n || count > sim->n - first)
        return fail(NB_E_INVALID_ARG, "target range [" + std::to_string(first) + ", " +
                                          std::to_string(first + count) + ") outside [0, " + std::to_string(sim->n) + ")");
    return NB_OK;
}

nb_status nb_stage_force_range(nb_sim* sim, int64_t first, int64_t count) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_sorted, "force needs a sorted tree");
    if (!st) st = check_range(sim, first, count);
    if (st) return st;
    if (count > 0) sim->eng.force_range(0, (int)first, (int)(first + count));
    sim->have_acc = true;
    return NB_OK;
}

nb_status nb_stage_brute_force_range(nb_sim* sim, int64_t first, int64_t count) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_state && !is_bh(sim), "brute force needs uploaded state and a brute-force algorithm");
    if (!st) st = check_range(sim, first, count);
    if (st) return st;
    if (count > 0) sim->eng.brute_range(0, (int)first, (int)(first + count));
    sim->have_acc = true;
    return NB_OK;
}

nb_status nb_stage_collision(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_state, "collision needs uploaded state");
    if (st) return st;
    sim->eng.collide(true);  // on request, whatever the configuration says
    return NB_OK;
}

nb_status nb_stage_update(nb_sim* sim) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    nb_status st = need(sim->have_acc, "update needs accelerations");
    if (st) return st;
    sim->eng.update(sim->cfg.method == NB_EULER ? 0 : 1);
    sim->have_box = false;
    ++sim->steps;
    return NB_OK;
}

// One whole step on the stream.  With timing, events e[0..7] bracket the stages:
// BH: e0 bounds e1 reset e2 build e3 summarize e4 sort e5 force e6 update e7;
// BF: e5 brute e6 update e7.
static void enqueue_force(nb_sim* sim, hipEvent_t* e) {
    Engine& g = sim->eng;
    auto mark = [&](int i) {
        if (e) (void)hipEventRecord(e[i], g.stream);
    };
    if (is_bh(sim)) {
        if (e) {  // per-stage events: the stages launched one by one
            mark(0);
            g.bounds();
            mark(1);
            g.tree_reset();
            mark(2);
        } else {
            g.bounds_reset();
        }
        g.build();
        mark(3);
        g.summarize();
        mark(4);
        g.sort();
        mark(5);
        g.force();
    } else {
        mark(5);
        g.brute();
    }
    mark(6);
}

static void enqueue_step(nb_sim* sim, hipEvent_t* e) {
    Engine& g = sim->eng;
    const bool lf = sim->cfg.method == NB_LEAPFROG;
    if (lf) g.update(2);
    enqueue_force(sim, e);
    g.update(lf ? 1 : 0);
    if (e) (void)hipEventRecord(e[7], g.stream);
    if (sim->cfg.collision) g.collide();  // after the integrator, as CMyApp::Update orders it
    if (e) (void)hipEventRecord(e[8], g.stream);
}

static constexpr size_t kEvPerStep = 9;

static nb_status ensure_events(nb_sim* sim, size_t nsteps) {
    while (sim->ev.size() < kEvPerStep * nsteps) {
        hipEvent_t e = nullptr;
        NB_HIPC(hipEventCreate(&e));
        sim->ev.push_back(e);
    }
    return NB_OK;
}

// Sum of each stage's event-measured time over the last nsteps timed steps.
static nb_status read_stage_times(nb_sim* sim, int nsteps) {
    std::fill(sim->stage_ms, sim->stage_ms + NB_NUM_STAGES, 0.0f);
    for (int s = 0; s < nsteps; ++s) {
        hipEvent_t* e = sim->ev.data() + kEvPerStep * (size_t)s;
        NB_HIPC(hipEventSynchronize(e[8]));
        float ms = 0.0f;
        if (is_bh(sim)) {
            for (int i = 0; i < 7; ++i) {
                NB_HIPC(hipEventElapsedTime(&ms, e[i], e[i + 1]));
                sim->stage_ms[i] += ms;
            }
        } else {
            NB_HIPC(hipEventElapsedTime(&ms, e[5], e[6]));
            sim->stage_ms[NB_STAGE_BRUTE] += ms;
            NB_HIPC(hipEventElapsedTime(&ms, e[6], e[7]));
            sim->stage_ms[NB_STAGE_UPDATE] += ms;
        }
        if (sim->cfg.collision) {
            NB_HIPC(hipEventElapsedTime(&ms, e[7], e[8]));
            sim->stage_ms[NB_STAGE_COLLISION] += ms;
        }
    }
    return NB_OK;
}

static void enqueue_fused(nb_sim* sim, int k) {
    Engine& g = sim->eng;
    if (sim->cfg.method == NB_LEAPFROG) {
        for (int i = 0; i < k; ++i) {
            g.update(2);
            g.brute_step(1, g.node4, nullptr);
        }
        return;
    }
    g.brute_step(0, g.node4, sim->pos_b);
    if (k == 2)
        g.brute_step(0, sim->pos_b, g.node4);
    else
        nb::launch_copy4(g.stream, sim->n, sim->pos_b, g.node4);
}

// k = 0: one plain step; k = 1, 2: that many fused steps
static nb_status build_graph(nb_sim* sim, int k, hipGraph_t* def, hipGraphExec_t* exec) {
    Engine& e = sim->eng;
    NB_HIPC(hipStreamBeginCapture(e.stream, hipStreamCaptureModeThreadLocal));
    if (k == 0)
        enqueue_step(sim, nullptr);
    else
        enqueue_fused(sim, k);
    const hipError_t le = hipGetLastError();  // a launch that failed to be recorded
    hipGraph_t g = nullptr;
    hipError_t ce = hipStreamEndCapture(e.stream, &g);
    if (ce == hipSuccess && le != hipSuccess) ce = le;
    if (ce != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail(NB_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
    }
    hipGraphExec_t ex = nullptr;
    ce = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (ce != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(NB_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ce));
    }
    *def = g;
    *exec = ex;
    return NB_OK;
}

// nsteps untimed steps on the stream (no synchronisation)
static nb_status enqueue_steps(nb_sim* sim, int nsteps) {
    Engine& e = sim->eng;
    if (fusable(sim)) {
        if (sim->cfg.method == NB_EULER && !sim->pos_b) return fail(NB_E_STATE, "no second position buffer");
        if (sim->cfg.use_graph) {
            for (int k = 1; k <= 2; ++k)
                if (!sim->fgraph[k - 1]) {
                    nb_status st = build_graph(sim, k, &sim->fgraph_def[k - 1], &sim->fgraph[k - 1]);
                    if (st) return st;
                }
            for (int i = 0; i + 1 < nsteps; i += 2) NB_HIPC(hipGraphLaunch(sim->fgraph[1], e.stream));
            if (nsteps & 1) NB_HIPC(hipGraphLaunch(sim->fgraph[0], e.stream));
        } else {
            for (int i = 0; i + 1 < nsteps; i += 2) enqueue_fused(sim, 2);
            if (nsteps & 1) enqueue_fused(sim, 1);
        }
    } else if (sim->cfg.use_graph) {
        if (!sim->graph) {
            nb_status st = build_graph(sim, 0, &sim->graph_def, &sim->graph);
            if (st) return st;
        }
        for (int i = 0; i < nsteps; ++i) NB_HIPC(hipGraphLaunch(sim->graph, e.stream));
    } else {
        for (int i = 0; i < nsteps; ++i) enqueue_step(sim, nullptr);
    }
    NB_HIPC(hipGetLastError());
    return NB_OK;
}

static void after_steps(nb_sim* sim, int nsteps) {
    sim->have_box = false;
    sim->tree_clean = false;
    sim->have_tree = sim->have_summary = sim->have_sorted = is_bh(sim);
    sim->have_acc = true;
    sim->steps += nsteps;
}

nb_status nb_step(nb_sim* sim, int32_t nsteps) {
    nb::TraceRange tr("nb_step");
    if (!sim || nsteps < 0) return fail(NB_E_INVALID_ARG, "bad arguments");
    if (!sim->have_state) return fail(NB_E_STATE, "nb_step needs uploaded state");
    if (nsteps == 0 || sim->n == 0) return NB_OK;
    Engine& e = sim->eng;
    if (sim->cfg.method == NB_LEAPFROG && !sim->have_acc) {
        enqueue_force(sim, nullptr);  // kick-drift-kick needs a(x_0)
        sim->have_acc = true;
    }
    const bool timed = sim->cfg.time_stages != 0;
    if (timed) {
        nb_status st = ensure_events(sim, (size_t)nsteps);
        if (st) return st;
        for (int i = 0; i < nsteps; ++i) enqueue_step(sim, sim->ev.data() + kEvPerStep * (size_t)i);
        NB_HIPC(hipGetLastError());
    } else {
        nb_status st = enqueue_steps(sim, nsteps);
        if (st) return st;
    }
    NB_HIPC(hipStreamSynchronize(e.stream));
    if (timed) {
        nb_status st = read_stage_times(sim, nsteps);
        if (st) return st;
    }
    const nb_status st = nb_sim_status(sim);
    // a failed step leaves the state at the last good step and the step counter unchanged
    after_steps(sim, st == NB_OK ? nsteps : 0);
    return st;
}

nb_status nb_step_async(nb_sim* sim, int32_t nsteps) {
    nb::TraceRange tr("nb_step_async");
    if (!sim || nsteps < 0) return fail(NB_E_INVALID_ARG, "bad arguments");
    if (!sim->have_state) return fail(NB_E_STATE, "nb_step_async needs uploaded state");
    if (nsteps == 0 || sim->n == 0) return NB_OK;
    if (sim->cfg.method == NB_LEAPFROG && !sim->have_acc) {
        enqueue_force(sim, nullptr);
        sim->have_acc = true;
    }
    nb_status st = enqueue_steps(sim, nsteps);
    if (st) return st;
    after_steps(sim, nsteps);
    return NB_OK;
}

nb_status nb_sim_energy(nb_sim* sim, double* out) {
    if (!sim || !out) return fail(NB_E_INVALID_ARG, "NULL argument");
    if (!sim->have_state) return fail(NB_E_STATE, "no state uploaded");
    Engine& g = sim->eng;
    const int n = sim->n;
    if (n == 0) {
        std::fill(out, out + NB_ENERGY_LEN, 0.0);
        return NB_OK;
    }
    const int nblk = (n + 255) / 256;
    if (!sim->energy_buf)
        NB_HIPC(hipMalloc(&sim->energy_buf, sizeof(double) * nb::kEnergyLen * (size_t)(nblk + 1)));
    const bool plummer = sim->cfg.algorithm == NB_BARNES_HUT || sim->cfg.law == NB_LAW_PLUMMER;
    const double eps = plummer ? sim->cfg.epsilon : (sim->cfg.algorithm == NB_BRUTE_FORCE_LOCAL ? 0.1 : 0.01);
    double* res = sim->energy_buf + (size_t)nb::kEnergyLen * nblk;
    nb::launch_energy(g.stream, n, g.node4, g.vel4, plummer ? 0 : 1, eps, sim->cfg.gravitational_constant,
                      sim->energy_buf, res);
    NB_HIPC(hipGetLastError());
    NB_HIPC(hipMemcpyAsync(out, res, sizeof(double) * nb::kEnergyLen, hipMemcpyDeviceToHost, g.stream));
    NB_HIPC(hipStreamSynchronize(g.stream));
    return NB_OK;
}

nb_status nb_export_vertices(nb_sim* sim, float* dst, int32_t dst_is_device) {
    if (!sim || (!dst && sim->n > 0)) return fail(NB_E_INVALID_ARG, "NULL argument");
    if (!sim->have_state) return fail(NB_E_STATE, "no state uploaded");
    Engine& g = sim->eng;
    const int n = sim->n;
    if (n == 0) return NB_OK;
    if (dst_is_device) {
        nb::launch_vertices(g.stream, n, g.node4, reinterpret_cast<float4*>(dst));
        NB_HIPC(hipGetLastError());
        return NB_OK;
    }
    NB_HIPC(hipMemcpyAsync(dst, g.node4, sizeof(float4) * n, hipMemcpyDeviceToHost, g.stream));
    NB_HIPC(hipStreamSynchronize(g.stream));
    for (int i = 0; i < n; ++i) dst[4 * i + 3] = 1.0f;
    return NB_OK;
}

nb_status nb_sim_step_count(nb_sim* sim, int64_t* steps) {
    if (!sim || !steps) return fail(NB_E_INVALID_ARG, "NULL argument");
    *steps = sim->steps;
    return NB_OK;
}

static size_t checkpoint_bytes(int n) { return sizeof(nb_checkpoint_header) + sizeof(float) * 10 * (size_t)n; }

nb_status nb_checkpoint_size(nb_sim* sim, size_t* bytes) {
    if (!sim || !bytes) return fail(NB_E_INVALID_ARG, "NULL argument");
    *bytes = checkpoint_bytes(sim->n);
    return NB_OK;
}

nb_status nb_checkpoint_save(nb_sim* sim, void* buf, size_t len) {
    if (!sim || !buf) return fail(NB_E_INVALID_ARG, "NULL argument");
    if (!sim->have_state) return fail(NB_E_STATE, "no state uploaded");
    const int n = sim->n;
    if (len < checkpoint_bytes(n)) return fail(NB_E_INVALID_ARG, "buffer smaller than nb_checkpoint_size");
    NB_HIPC(hipStreamSynchronize(sim->eng.stream));
    nb_checkpoint_header h;
    std::memset(&h, 0, sizeof(h));
    h.magic = NB_CHECKPOINT_MAGIC;
    h.version = NB_CHECKPOINT_VERSION;
    h.num_bodies = n;
    h.have_acc = sim->have_acc ? 1 : 0;
    h.step = sim->steps;
    h.config = sim->cfg;
    char* out = static_cast<char*>(buf);
    std::memcpy(out, &h, sizeof(h));
    float* pos = reinterpret_cast<float*>(out + sizeof(h));
    float* vel = pos + 3 * (size_t)n;
    float* mass = vel + 3 * (size_t)n;
    float* acc = mass + (size_t)n;
    nb_status st = nb_sim_download(sim, pos, vel, acc);
    if (st) return st;
    if (n > 0) {
        std::vector<float4> h4(n);
        Engine& g = sim->eng;
        NB_HIPC(hipMemcpyAsync(h4.data(), g.node4, sizeof(float4) * n, hipMemcpyDeviceToHost, g.stream));
        NB_HIPC(hipStreamSynchronize(g.stream));
        for (int i = 0; i < n; ++i) mass[i] = h4[i].w;
    }
    return NB_OK;
}

nb_status nb_checkpoint_load(nb_sim* sim, const void* buf, size_t len) {
    if (!sim || !buf) return fail(NB_E_INVALID_ARG, "NULL argument");
    if (len < sizeof(nb_checkpoint_header)) return fail(NB_E_INVALID_ARG, "truncated checkpoint");
    nb_checkpoint_header h;
    std::memcpy(&h, buf, sizeof(h));
    if (h.magic != NB_CHECKPOINT_MAGIC || h.version != NB_CHECKPOINT_VERSION)
        return fail(NB_E_INVALID_ARG, "not a checkpoint of this version");
    if (h.num_bodies != sim->n) return fail(NB_E_INVALID_ARG, "checkpoint holds a different number of bodies");
    const int n = sim->n;
    if (len < checkpoint_bytes(n)) return fail(NB_E_INVALID_ARG, "truncated checkpoint");
    const float* pos = reinterpret_cast<const float*>(static_cast<const char*>(buf) + sizeof(h));
    const float* vel = pos + 3 * (size_t)n;
    const float* mass = vel + 3 * (size_t)n;
    const float* acc = mass + (size_t)n;
    nb_status st = nb_sim_upload(sim, pos, vel, mass);
    if (st) return st;
    if (n > 0 && h.have_acc) {
        std::vector<float4> a4(n);
        for (int i = 0; i < n; ++i) a4[i] = make_float4(acc[3 * i], acc[3 * i + 1], acc[3 * i + 2], 0.0f);
        Engine& g = sim->eng;
        NB_HIPC(hipMemcpyAsync(g.acc4, a4.data(), sizeof(float4) * n, hipMemcpyHostToDevice, g.stream));
        NB_HIPC(hipStreamSynchronize(g.stream));
    }
    sim->have_acc = h.have_acc != 0;
    sim->steps = h.step;
    return NB_OK;
}

nb_status nb_stage_times(nb_sim* sim, float* ms) {
    if (!sim || !ms) return fail(NB_E_INVALID_ARG, "NULL argument");
    std::memcpy(ms, sim->stage_ms, sizeof(float) * NB_NUM_STAGES);
    return NB_OK;
}

nb_status nb_tree_info_get(nb_sim* sim, nb_tree_info* info) {
    if (!sim || !info) return fail(NB_E_INVALID_ARG, "NULL argument");
    std::vector<SysState> s;
    nb_status st = sim->eng.download_state(s);
    if (st) return st;
    info->num_cells = std::min(s[0].ncells, sim->cap);
    info->max_depth = s[0].max_depth;
    info->capacity = sim->cap;
    info->status = status_of_bits(s[0].status);
    return NB_OK;
}

nb_status nb_export_dq(nb_sim* sim, float* dq) {
    if (!sim || !dq) return fail(NB_E_INVALID_ARG, "NULL argument");
    std::vector<SysState> s;
    nb_status st = sim->eng.download_state(s);
    if (st) return st;
    std::memcpy(dq, s[0].dq, sizeof(float) * nb::kDqLen);
    return NB_OK;
}

}  // extern "C"

// Host copy of the tree for the canonical exports.
struct HostTree {
    int n = 0, ncells = 0;
    std::vector<int> child, count;
    std::vector<float4> cells;
};

static nb_status fetch_tree(nb_sim* sim, HostTree& t) {
    if (!sim->have_tree) return fail(NB_E_STATE, "no tree to export");
    Engine& e = sim->eng;
    std::vector<SysState> s;
    nb_status st = e.download_state(s);
    if (st) return st;
    if (s[0].status) return status_of_bits(s[0].status);
    t.n = sim->n;
    t.ncells = std::min(s[0].ncells, sim->cap);
    t.child.resize(8 * (size_t)t.ncells);
    t.count.resize(t.ncells);
    t.cells.resize(t.ncells);
    const nb::SysDesc& d = e.hdesc[0];
    nb::launch_cell_export(e.stream, d.n, t.ncells, e.child + 8 * (size_t)d.cell_off, e.cdata + 8 * (size_t)d.cell_off,
                           e.slotoff + 8 * (size_t)d.cell_off, e.node4 + d.node_off + d.n, e.ccount + d.cell_off);
    NB_HIPC(hipGetLastError());
    NB_HIPC(hipMemcpyAsync(t.child.data(), e.child + 8 * (size_t)d.cell_off, sizeof(int) * 8 * t.ncells,
                           hipMemcpyDeviceToHost, e.stream));
    NB_HIPC(hipMemcpyAsync(t.count.data(), e.ccount + d.cell_off, sizeof(int) * t.ncells, hipMemcpyDeviceToHost,
                           e.stream));
    NB_HIPC(hipMemcpyAsync(t.cells.data(), e.node4 + d.node_off + d.n, sizeof(float4) * t.ncells,
                           hipMemcpyDeviceToHost, e.stream));
    NB_HIPC(hipStreamSynchronize(e.stream));
    return NB_OK;
}

// Depth-first preorder, slots in order (the canonical form shared with the oracle).
template <class BodyFn, class CellFn>
static void dfs(const HostTree& t, BodyFn on_body, CellFn on_cell) {
    if (t.n == 0 || t.ncells == 0) return;
    int stk_node[NB_MAX_LEVEL + 2], stk_next[NB_MAX_LEVEL + 2];
    uint8_t digits[NB_MAX_LEVEL + 2];
    int sp = 0;
    stk_node[0] = 0;
    stk_next[0] = 0;
    on_cell(0, 0, -1);
    while (sp >= 0) {
        const int node = stk_node[sp];
        const int k = stk_next[sp];
        if (k >= 8) {
            --sp;
            continue;
        }
        stk_next[sp] = k + 1;
        const int v = t.child[8 * (size_t)node + k];
        if (v < 0) continue;
        digits[sp] = (uint8_t)k;
        if (v < t.n) {
            on_body(v, digits, sp + 1);
        } else {
            if (sp + 1 > NB_MAX_LEVEL) continue;  // cannot happen: the build refuses deeper cells
            on_cell(v - t.n, sp + 1, k);
            ++sp;
            stk_node[sp] = v - t.n;
            stk_next[sp] = 0;
        }
    }
}

extern "C" {

nb_status nb_export_bodies(nb_sim* sim, int32_t* sorted, uint8_t* paths, int32_t* path_len) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    if (sorted) {
        if (!sim->have_sorted) return fail(NB_E_STATE, "no sorted order to export");
        Engine& e = sim->eng;
        if (sim->n > 0) {
            NB_HIPC(hipMemcpyAsync(sorted, e.sorted, sizeof(int) * sim->n, hipMemcpyDeviceToHost, e.stream));
            NB_HIPC(hipStreamSynchronize(e.stream));
        }
    }
    if (paths || path_len) {
        HostTree t;
        nb_status st = fetch_tree(sim, t);
        if (st) return st;
        dfs(
            t,
            [&](int b, const uint8_t* dg, int len) {
                if (paths) {
                    std::memset(paths + (size_t)b * NB_PATH_STRIDE, 0xff, NB_PATH_STRIDE);
                    std::memcpy(paths + (size_t)b * NB_PATH_STRIDE, dg, (size_t)len);
                }
                if (path_len) path_len[b] = len;
            },
            [&](int, int, int) {});
    }
    return NB_OK;
}

nb_status nb_export_force_order(nb_sim* sim, int32_t* order) {
    if (!sim || !order) return fail(NB_E_INVALID_ARG, "NULL argument");
    if (!sim->have_sorted) return fail(NB_E_STATE, "no force order to export");
    Engine& e = sim->eng;
    if (sim->n > 0) {
        NB_HIPC(hipMemcpyAsync(order, e.forder, sizeof(int) * sim->n, hipMemcpyDeviceToHost, e.stream));
        NB_HIPC(hipStreamSynchronize(e.stream));
    }
    return NB_OK;
}

nb_status nb_export_cells(nb_sim* sim, int32_t* level, int32_t* slot, int32_t* count, float* com4) {
    if (!sim) return fail(NB_E_INVALID_ARG, "NULL sim");
    HostTree t;
    nb_status st = fetch_tree(sim, t);
    if (st) return st;
    if (!sim->have_summary && (count || com4)) return fail(NB_E_STATE, "counts need a summarized tree");
    int idx = 0;
    dfs(
        t, [&](int, const uint8_t*, int) {},
        [&](int c, int lvl, int k) {
            if (level) level[idx] = lvl;
            if (slot) slot[idx] = k;
            if (count) count[idx] = t.count[c];
            if (com4) {
                com4[4 * idx] = t.cells[c].x;
                com4[4 * idx + 1] = t.cells[c].y;
                com4[4 * idx + 2] = t.cells[c].z;
                com4[4 * idx + 3] = t.cells[c].w;
            }
            ++idx;
        });
    return NB_OK;
}

}  // extern "C"

// Per-body interaction counts of the last force pass: the force kernels store them in acc.w.
static nb_status interactions_of(Engine& e, int body_off, int n, int32_t* out) {
    if (n <= 0) return NB_OK;
    std::vector<float4> h(n);
    NB_HIPC(hipMemcpyAsync(h.data(), e.acc4 + body_off, sizeof(float4) * n, hipMemcpyDeviceToHost, e.stream));
    NB_HIPC(hipStreamSynchronize(e.stream));
    for (int i = 0; i < n; ++i) out[i] = (int32_t)h[i].w;
    return NB_OK;
}

extern "C" {

nb_status nb_export_interactions(nb_sim* sim, int32_t* per_body) {
    if (!sim || (!per_body && sim->n > 0)) return fail(NB_E_INVALID_ARG, "NULL argument");
    return interactions_of(sim->eng, 0, sim->n, per_body);
}

nb_status nb_force_interactions(nb_sim* sim, int64_t* total) {
    if (!sim || !total) return fail(NB_E_INVALID_ARG, "NULL argument");
    Engine& e = sim->eng;
    std::vector<float4> h(std::max(sim->n, 1));
    if (sim->n > 0) {
        NB_HIPC(hipMemcpyAsync(h.data(), e.acc4, sizeof(float4) * sim->n, hipMemcpyDeviceToHost, e.stream));
        NB_HIPC(hipStreamSynchronize(e.stream));
    }
    int64_t s = 0;
    for (int i = 0; i < sim->n; ++i) s += (int64_t)h[i].w;
    *total = s;
    return NB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// ensembles

struct nb_ensemble {
    Engine eng;
    std::vector<nb_system_spec> specs;
    int method = NB_EULER;
    bool use_graph = true;
    bool have_acc = false;
    std::vector<bool> uploaded;
    hipGraphExec_t graph = nullptr;
    hipGraph_t graph_def = nullptr;
};

static void ens_enqueue_force(nb_ensemble* e) { e->eng.force_all(); }

static void ens_enqueue_step(nb_ensemble* e) {
    const bool lf = e->method == NB_LEAPFROG;
    if (lf) e->eng.update(2);
    ens_enqueue_force(e);
    e->eng.update(lf ? 1 : 0);
    if (e->eng.any_collision) e->eng.collide();  // the systems that ask for it, after the integrator
}

extern "C" {

nb_status nb_ens_create(const nb_system_spec* specs, int32_t count, int32_t method, int32_t law,
                        int32_t accumulation, int32_t use_graph, void* stream, nb_ensemble** out) {
    if (!out || !specs || count < 1) return fail(NB_E_INVALID_ARG, "bad arguments");
    *out = nullptr;
    if (method != NB_EULER && method != NB_LEAPFROG) return fail(NB_E_INVALID_ARG, "bad method");
    if (law != NB_LAW_PLUMMER && law != NB_LAW_REFERENCE_UNIT) return fail(NB_E_INVALID_ARG, "bad law");
    if (accumulation != NB_ACCUM_PLAIN && accumulation != NB_ACCUM_KAHAN)
        return fail(NB_E_INVALID_ARG, "bad accumulation");
    std::vector<SysSpec> sys(count);
    for (int i = 0; i < count; ++i) {
        nb_status st = nb::spec_from_system(specs[i], method, law, accumulation, &sys[i]);
        if (st) return fail(st, "system " + std::to_string(i) + ": " + nb::last_error());
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NB_E_NO_DEVICE, "no HIP device");
    nb_ensemble* e = new nb_ensemble();
    e->specs.assign(specs, specs + count);
    e->method = method;
    e->use_graph = use_graph != 0;
    e->uploaded.assign(count, false);
    nb_status st = e->eng.init(sys, stream);
    if (st) {
        delete e;
        return st;
    }
    *out = e;
    return NB_OK;
}

nb_status nb_ens_destroy(nb_ensemble* e) {
    if (!e) return NB_OK;
    if (e->eng.stream) (void)hipStreamSynchronize(e->eng.stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    if (e->graph_def) (void)hipGraphDestroy(e->graph_def);
    delete e;
    return NB_OK;
}

nb_status nb_ens_upload(nb_ensemble* e, int32_t sys, const float* pos, const float* vel, const float* mass) {
    if (!e || sys < 0 || sys >= (int)e->specs.size()) return fail(NB_E_INVALID_ARG, "bad system index");
    const nb::SysDesc& d = e->eng.hdesc[sys];
    const int n = d.n;
    if (n > 0 && (!pos || !mass)) return fail(NB_E_INVALID_ARG, "NULL argument");
    std::vector<float4> h4(std::max(n, 1)), hv(std::max(n, 1));
    for (int i = 0; i < n; ++i) {
        const float m = mass[i];
        const bool ok = std::isfinite(m) && (d.algo == NB_BARNES_HUT ? m > 0.0f : m >= 0.0f);
        if (!ok) return fail(NB_E_INVALID_ARG, "mass " + std::to_string(i) + " must be finite and > 0");
        h4[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], m);
        hv[i] = vel ? make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0f) : make_float4(0, 0, 0, 0);
    }
    Engine& g = e->eng;
    if (n > 0) {
        NB_HIPC(hipMemcpyAsync(g.node4 + d.node_off, h4.data(), sizeof(float4) * n, hipMemcpyHostToDevice, g.stream));
        NB_HIPC(hipMemcpyAsync(g.vel4 + d.body_off, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, g.stream));
        NB_HIPC(hipMemsetAsync(g.acc4 + d.body_off, 0, sizeof(float4) * n, g.stream));
    }
    NB_HIPC(hipStreamSynchronize(g.stream));
    e->uploaded[sys] = true;
    e->have_acc = false;
    return NB_OK;
}

nb_status nb_ens_download(nb_ensemble* e, int32_t sys, float* pos, float* vel, float* acc) {
    if (!e || sys < 0 || sys >= (int)e->specs.size()) return fail(NB_E_INVALID_ARG, "bad system index");
    Engine& g = e->eng;
    const nb::SysDesc& d = g.hdesc[sys];
    const int n = d.n;
    if (n == 0) return NB_OK;
    std::vector<float4> h(n);
    auto get = [&](const float4* src, float* dst) -> nb_status {
        if (!dst) return NB_OK;
        NB_HIPC(hipMemcpyAsync(h.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost, g.stream));
        NB_HIPC(hipStreamSynchronize(g.stream));
        for (int i = 0; i < n; ++i) {
            dst[3 * i] = h[i].x;
            dst[3 * i + 1] = h[i].y;
            dst[3 * i + 2] = h[i].z;
        }
        return NB_OK;
    };
    nb_status st = get(g.node4 + d.node_off, pos);
    if (!st) st = get(g.vel4 + d.body_off, vel);
    if (!st) st = get(g.acc4 + d.body_off, acc);
    return st;
}

nb_status nb_ens_export_interactions(nb_ensemble* e, int32_t sys, int32_t* per_body) {
    if (!e || sys < 0 || sys >= (int)e->specs.size()) return fail(NB_E_INVALID_ARG, "bad system index");
    const nb::SysDesc& d = e->eng.hdesc[sys];
    if (!per_body && d.n > 0) return fail(NB_E_INVALID_ARG, "NULL argument");
    return interactions_of(e->eng, d.body_off, d.n, per_body);
}

nb_status nb_ens_step(nb_ensemble* e, int32_t nsteps) {
    nb::TraceRange tr("nb_ens_step");
    if (!e || nsteps < 0) return fail(NB_E_INVALID_ARG, "bad arguments");
    for (bool u : e->uploaded)
        if (!u) return fail(NB_E_STATE, "every system needs uploaded state");
    Engine& g = e->eng;
    if (nsteps == 0) return NB_OK;
    if (e->method == NB_LEAPFROG && !e->have_acc) {
        ens_enqueue_force(e);
        e->have_acc = true;
    }
    if (e->use_graph) {
        if (!e->graph) {
            NB_HIPC(hipStreamBeginCapture(g.stream, hipStreamCaptureModeThreadLocal));
            ens_enqueue_step(e);
            const hipError_t le = hipGetLastError();  // a launch that failed to be recorded
            hipGraph_t gr = nullptr;
            hipError_t ce = hipStreamEndCapture(g.stream, &gr);
            if (ce == hipSuccess && le != hipSuccess) ce = le;
            if (ce != hipSuccess) {
                if (gr) (void)hipGraphDestroy(gr);
                return fail(NB_E_HIP, std::string("graph capture: ") + hipGetErrorString(ce));
            }
            ce = hipGraphInstantiate(&e->graph, gr, nullptr, nullptr, 0);
            if (ce != hipSuccess) {
                (void)hipGraphDestroy(gr);
                return fail(NB_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ce));
            }
            e->graph_def = gr;
        }
        for (int i = 0; i < nsteps; ++i) NB_HIPC(hipGraphLaunch(e->graph, g.stream));
    } else {
        for (int i = 0; i < nsteps; ++i) ens_enqueue_step(e);
    }
    NB_HIPC(hipGetLastError());
    NB_HIPC(hipStreamSynchronize(g.stream));
    e->have_acc = true;
    std::vector<SysState> s;
    nb_status st = g.download_state(s);
    if (st) return st;
    for (const SysState& x : s)
        if (x.status) return status_of_bits(x.status);
    return NB_OK;
}

nb_status nb_ens_status(nb_ensemble* e, int32_t* status) {
    if (!e || !status) return fail(NB_E_INVALID_ARG, "NULL argument");
    std::vector<SysState> s;
    nb_status st = e->eng.download_state(s);
    if (st) return st;
    for (size_t i = 0; i < s.size(); ++i) status[i] = (int32_t)status_of_bits(s[i].status);
    return NB_OK;
}

nb_status nb_ens_device_bytes(nb_ensemble* e, int64_t* bytes) {
    if (!e || !bytes) return fail(NB_E_INVALID_ARG, "NULL argument");
    *bytes = (int64_t)e->eng.arena_bytes;
    return NB_OK;
}

}  // extern "C"

namespace nb {

nb_status status_from_bits(int bits) { return status_of_bits(bits); }

nb_status spec_from_system(const nb_system_spec& s, int method, int law, int accum, SysSpec* out) {
    nb_config c;
    nb_config_default(&c);
    c.num_bodies = s.num_bodies;
    c.algorithm = s.algorithm;
    c.method = method;
    c.rules = s.rules;
    c.law = law;
    c.accumulation = accum;
    c.gravitational_constant = s.gravitational_constant;
    c.theta = s.theta;
    c.epsilon = s.epsilon;
    c.dt = s.dt;
    c.cell_capacity = s.cell_capacity;
    c.collision = s.collision;
    nb_status st = validate(&c);
    if (st) return st;
    const int cap = c.algorithm == NB_BARNES_HUT ? (c.cell_capacity > 0 ? c.cell_capacity : 2 * c.num_bodies + 1024) : 0;
    *out = spec_of(&c, cap);
    return NB_OK;
}

}  // namespace nb
