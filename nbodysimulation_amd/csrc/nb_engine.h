// nb_engine.h -- host runtime of the N-body engine: device arena, block tables,
// stage launches for a batch of independent systems.
//
// The arena is reserved once (capacities in nodes, bodies, cells and system
// slots); `activate` then describes which systems occupy which ranges and
// rebuilds the block tables, so systems can join and leave (serving) without
// reallocating device memory.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "nb_internal.h"
#include "nbody/nbody.h"

#include <rocprofiler-sdk-roctx/roctx.h>

namespace nb {

// kernel launchers (nb_kernels.hip)
void launch_bounds(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                   const float4* node4, float* partial);
// the bounds pass (nbounds workgroups of table wb) in one launch with the force pass's dispatch
// order (as launch_force_order) computed by one more workgroup when olen is in (0, kOrderSmall]
// (otherwise not computed); the root's re-seeding is left to k_build_top
void launch_bounds_order(hipStream_t s, int nbounds, const BlkWork* wb, const SysDesc* sd, SysState* st,
                         const float4* node4, float* partial, const BlkWork* og, int olen, int* keys, BlkWork* oout);
// one system's per-cell (COM, mass) (cells: its node4 past the bodies) and body counts, for the
// exports (the step writes them for the root only)
void launch_cell_export(hipStream_t s, int n, int ncells, const int* child, const float4* cdata,
                        const int* slotoff, float4* cells, int* ccount);
// slots NULL: 0..nsys-1; sd and slotoff NULL: the roots' slots and counts are left alone
void launch_tree_init(hipStream_t s, int nsys, const int* slots, const SysDesc* sd, SysState* st, int* slotoff,
                      int* child, int* cpend);
// fblist: the build's fallback list, which takes the phase's bodies when the LDS tree overflows
void launch_build_top(hipStream_t s, int nsys, const int* slots, const SysDesc* sd, SysState* st,
                      const float4* node4, int* child, int* cparent, int* cpend, float4* cdata, int* slotoff,
                      int stride, int* fblist);
// list: the queued phase takes its bodies from this per-system list (SysState::nfb entries)
void launch_build(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                  const float4* node4, int* child, int* cparent, int* cpend, float4* cdata, int* slotoff, int stride,
                  int exclude, const int* list);
// the bucketed insertion phases (nb_build.hip)
void launch_classify(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                     const float4* node4, const int* child, int* slotoff, int* ccount, int* bkey, int* brank,
                     float4* ppos, const int* sorder, int stride, int exclude);
void launch_frames(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                   const float4* node4, const int* child, int* slotoff, const int* ccount, int* kbase, float4* lpos,
                   int4* lmeta, int2* bframe);
void launch_scatter(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                    const float4* ppos, const int* kbase, const int* bkey, const int* brank, const int* ccount,
                    const int* sorder, float4* lpos, int4* lmeta, int* fblist, int stride, int exclude);
// last: no later bucketed phase (the new cells' per-slot counts are not cleared)
// last: the last phase, whose frames also summarize their subtrees (when no earlier phase
// left bodies for the queued insertion, which could land in them)
void launch_frame_build(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                        float4* node4, int* child, int* cparent, int* cpend, int* ccount, int* slotoff,
                        float4* cdata, const float4* lpos, const int4* lmeta, const int2* bframe, int* fblist,
                        bool last);
void launch_summarize(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                      float4* node4, const int* child, const int* cparent, int* cpend, int* ccount, int* slotoff,
                      float4* cdata);
void launch_sort(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work, const float4* node4,
                 const int* child, const int* slotoff, int* sorted, int* forder, const int* sorder);
void launch_force_order(hipStream_t s, const BlkWork* in, int len, int* keys, BlkWork* out);
// pair: the instance for systems under REFERENCE rules (walks two cells per scalar-load
// round trip); otherwise the one for STANDARD rules.  Each skips the other's systems.
void launch_force_bh(hipStream_t s, int nblk, bool kahan, bool pair, const SysDesc* sd, SysState* st,
                     const BlkWork* work, const float4* node4, const int* child, const float4* cdata,
                     const int* order, float4* acc4, int* sorder, int* keys);
// mode >= 0: also integrate (k_update's mode 0 or 1); mode 0 writes the moved
// positions to out4 (same layout as node4), never in place
void launch_brute(hipStream_t s, int nblk, int law, bool kahan, int tb, const SysDesc* sd, const SysState* st,
                  const BlkWork* work, const float4* node4, float4* acc4, float4* vel4, float4* out4, int mode);
void launch_range_table(hipStream_t s, BlkWork* out, int nblk, int sys, int start, int stop, int lo, int hi,
                        int per);
// all = false: only the systems whose SysDesc asks for the collision stage
void launch_collision(hipStream_t s, int nblk, bool all, const SysDesc* sd, const SysState* st, const BlkWork* work,
                      const float4* node4, float4* vel4);
void launch_copy4(hipStream_t s, int n, const float4* src, float4* dst);
void launch_vertices(hipStream_t s, int n, const float4* node4, float4* out);
void launch_energy(hipStream_t s, int n, const float4* node4, const float4* vel4, int law, double eps, double G,
                   double* partial, double* out);
void launch_update(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                   float4* node4, float4* vel4, const float4* acc4, int mode);

// Host-side roctx range around a stage launch or an API call, visible to
// rocprofv3 --marker-trace next to the kernels it launched (a no-op without a
// profiler).  Under stream capture it marks the capture, not the replays.
struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

void set_error(const std::string& msg);
const char* last_error();

// Process-wide launch-shape tuning (nb_set_tuning).
const nb_tuning& tuning();

// One system, as requested by the caller.
struct SysSpec {
    int n = 0;
    int cap = 0;  // cells (Barnes-Hut only)
    int rules = 0, algo = 0, law = 1, accum = 0, method = 2;
    float G = 1.f, theta = 0.5f, eps = 0.01f, dt = 1e-3f;
    int collision = 0;  // the collision stage runs after this system's update
};

// A system placed in the arena.
struct Slot {
    int slot = 0;  // index into the SysDesc / SysState arrays
    SysSpec spec;
    long long node_off = 0, body_off = 0, cell_off = 0;
};

// Arena capacities.
struct Capacity {
    long long nodes = 0, bodies = 0, cells = 0;
    int systems = 0;
};

// Bytes of device memory per node, body and cell in the arena layout.
constexpr size_t kBytesPerNode = 16;                     // node4
constexpr size_t kBytesPerBody = 16 + 16 + 4 + 4 + 4 + 32 + 16 + 8 + 4;  // vel4, acc4, sorted, forder, sorder, lpos+lmeta, ppos, bframe, fblist
constexpr size_t kBytesPerCell = 32 + 4 * 3 + 32 + 128 + 32;      // child, parent/pend/count, slotoff, cdata, kbase

struct Engine {
#ifndef NB_BUILD_PHASES
#define NB_BUILD_PHASES 3
#endif
    static constexpr int kBuildPhases = NB_BUILD_PHASES;  // insertion phases of the build (at most 4)

    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    // a batch that holds both kinds of systems runs its brute-force launch on `side`, beside
    // the Barnes-Hut stages on `stream` (forked and joined by the two events); created by
    // activate for the first such batch
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;

    Capacity capc;
    std::vector<SysDesc> hdesc;  // indexed by slot
    std::vector<Slot> active;
    bool any_bh = false, any_bf = false;
    bool any_bh_ref = false, any_bh_std = false;  // Barnes-Hut systems under REFERENCE / STANDARD rules
    bool any_top_global = false;  // a Barnes-Hut system whose first insertion phase exceeds kTopBodies
    int bf_tb = 1;  // brute-force targets per lane of the current batch (1 or 2)
    int law = 1;
    bool kahan = false;

    void* arena = nullptr;
    size_t arena_bytes = 0;
    float4 *node4 = nullptr, *vel4 = nullptr, *acc4 = nullptr, *cdata = nullptr;
    int *sorted = nullptr, *forder = nullptr, *sorder = nullptr, *child = nullptr, *cparent = nullptr, *cpend = nullptr, *ccount = nullptr,
        *slotoff = nullptr;
    // build scratch of the bucketed phases: per slot a bucket's place in the frame list; the
    // frame list (position; body, key, level), frames, fallback list
    int *kbase = nullptr, *fblist = nullptr;
    float4 *lpos = nullptr, *ppos = nullptr;
    int4* lmeta = nullptr;
    int2* bframe = nullptr;
    SysDesc* d_desc = nullptr;
    SysState* d_state = nullptr;
    float* d_partial = nullptr;
    int* d_slots = nullptr;  // slots of the active systems (tree_init touches only these)
    int* gkey = nullptr;     // per force workgroup: dispatch-order key (k_force_bh -> the next pass's ordering)
    bool order_done = false; // host: this step's bounds_reset() computed the force pass's dispatch order
    bool lpt = false;        // force workgroups dispatched longest first (few waves per SIMD)
    int nactive = 0;

    // block tables (device) with their capacities and current lengths
    // T_RANGE (target-range force launches) and T_FORDER (T_GROUPS in dispatch
    // order) are filled on the device
    // T_KEYS, T_FRAMES: (system, block index in the system, the system's block count)
    enum Table { T_BOUNDS, T_BODY, T_CELLS, T_GROUPS, T_BRUTE, T_UPDATE, T_PHASE0, T_PHASE1, T_PHASE2, T_PHASE3,
                 T_RANGE, T_FORDER, T_KEYS, T_FRAMES, T_COUNT };
    BlkWork* tab[T_COUNT] = {};
    int tab_cap[T_COUNT] = {};
    int tab_len[T_COUNT] = {};

    ~Engine();
    // Reserve the arena; `law`/`kahan` apply to every system.
    nb_status reserve(const Capacity& c, int law_, bool kahan_, void* user_stream);
    // Describe the systems now resident (their ranges must lie inside the arena)
    // and rebuild the block tables.  Synchronous on the engine's stream.
    nb_status activate(const std::vector<Slot>& slots);
    // Prepare a newly placed system: clear its cell slots and reset its state.
    nb_status reset_slot(const Slot& s);
    // Static batch: reserve exactly what `systems` need and place them in order.
    nb_status init(const std::vector<SysSpec>& systems, void* user_stream);
    void release();

    // whole-batch stage launches (asynchronous on `stream`)
    void bounds();
    void tree_reset();  // an empty root per system (no pass clears the other cells: their creators write them)
    // bounds() of a step with the force pass's dispatch order in one launch, the root's
    // re-seeding (tree_reset()) done by build()'s first kernel: only where build() follows at once
    void bounds_reset();
    void build();
    void summarize();
    void sort();
    void force();
    void brute(hipStream_t on = nullptr);  // on: the stream to launch on (default: stream)
    // the force pass of every resident system: the Barnes-Hut stages, and brute force beside
    // them on `side` when the batch holds both kinds (joined back into `stream`)
    void force_all();
    // Brute force with the integrator of `mode` (0 Euler, 1 closing kick) fused:
    // positions are read from `in` (node4 layout) and, for mode 0, the moved ones
    // written to `out`; velocities are updated in place.
    void brute_step(int mode, const float4* in, float4* out);
    // One system's force pass over the targets [lo, hi), 0 <= lo < hi <= n: tree
    // order for Barnes-Hut (vote groups stay those of the full pass), body index
    // for brute force.  Results equal the full pass's for those targets.
    void force_range(int sys, int lo, int hi);
    void brute_range(int sys, int lo, int hi);
    void update(int mode);
    // collision stage of the systems that ask for it (all: of every system)
    void collide(bool all = false);
    bool any_collision = false;

    nb_status download_state(std::vector<SysState>& out);
};

size_t align_up(size_t x, size_t a);

// nb_api.hip: validation of one system of a batch, and the nb_status of a
// system's device status bits (also sets the last-error text).
nb_status spec_from_system(const nb_system_spec& s, int method, int law, int accum, SysSpec* out);
nb_status status_from_bits(int bits);

}  // namespace nb
