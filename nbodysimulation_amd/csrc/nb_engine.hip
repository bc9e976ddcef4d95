// nb_engine.hip -- host runtime: one HBM arena, per-slot descriptors, block tables, stage launches.
#include "nb_engine.h"

#include <algorithm>
#include <cstring>

namespace nb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Insertion phases of the build: every 4096th body, then every 64th, then the rest.
#ifndef NB_PHASE_STRIDES
#define NB_PHASE_STRIDES 4096, 64, 1
#endif
static constexpr int kPhaseStride[Engine::kBuildPhases] = {NB_PHASE_STRIDES};
static constexpr int kBoundsPerBlock = 2048;
// Bucketed phases: blocks per system of the frame packing (over the slots of the tree so
// far) and of the frame builds (one workgroup per frame in flight; frames hold up to 512 bodies)
static constexpr int kKeysBodiesPerBlock = 4096;
static constexpr int kFramesBodiesPerBlock = 256;

#define NB_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            set_error(std::string(#call) + ": " + hipGetErrorString(e_));              \
            return e_ == hipErrorOutOfMemory ? NB_E_OUT_OF_MEMORY : NB_E_HIP;          \
        }                                                                              \
    } while (0)

Engine::~Engine() { release(); }

void Engine::release() {
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
    if (side) (void)hipStreamDestroy(side);
    ev_fork = ev_join = nullptr;
    side = nullptr;
    if (arena) (void)hipFree(arena);
    arena = nullptr;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    own_stream = false;
}

// Block table over `count` items of one system, `per` items per workgroup.
static void add_ranges(std::vector<BlkWork>& t, int sys, int count, int per) {
    for (int b = 0; b < count; b += per) t.push_back(BlkWork{sys, b, std::min(count, b + per), 0});
}

static nb_tuning g_tuning = {0, 8192, {0, 0, 0, 0, 0, 0}};
const nb_tuning& tuning() { return g_tuning; }

}  // namespace nb

extern "C" {

void nb_tuning_default(nb_tuning* t) {
    if (!t) return;
    *t = nb_tuning{0, 8192, {0, 0, 0, 0, 0, 0}};
}

nb_status nb_set_tuning(const nb_tuning* t) {
    if (!t || t->bf_targets_per_lane < 0 || t->bf_targets_per_lane > 2 || t->bf_min_waves < 1 ||
        t->bf_min_waves > (1 << 24)) {
        nb::set_error("bad tuning");
        return NB_E_INVALID_ARG;
    }
    nb::g_tuning = *t;
    return NB_OK;
}

nb_status nb_get_tuning(nb_tuning* t) {
    if (!t) {
        nb::set_error("NULL tuning");
        return NB_E_INVALID_ARG;
    }
    *t = nb::g_tuning;
    return NB_OK;
}

}  // extern "C"

namespace nb {

// Brute force shape of one system, from N and the tuning alone (so a system sums
// in the same order in any batch): S = 2^lg_s waves split a target group's sources
// (at most the 16 waves of a workgroup), then W = 2^lg_w lanes split each target's
// sources (at most 4), until one target per lane gives bf_min_waves waves (default
// 8192: two 16-wave workgroups per CU on 256 CUs) or the limits are reached.
static void bf_shape(int n, int& lg_s, int& lg_w) {
    lg_s = 0;
    lg_w = 0;
    auto waves = [&]() {
        const long long per_wave = 64 >> lg_w;
        return ((n + per_wave - 1) / per_wave) << lg_s;
    };
    const long long target = g_tuning.bf_min_waves;
    while (lg_s < 4 && waves() < target) ++lg_s;
    while (lg_w < 2 && waves() < target) ++lg_w;
}

static int blocks_bound(long long items, int per, int systems) {
    return (int)std::min<long long>((items + per - 1) / per + 2LL * systems, 0x7fffffff);
}

nb_status Engine::reserve(const Capacity& c, int law_, bool kahan_, void* user_stream) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device");
        return NB_E_NO_DEVICE;
    }
    if (c.nodes > 0x7fffffffLL || c.bodies > 0x7fffffffLL || c.cells * 8 > 0x7fffffffLL * 4LL || c.systems < 1) {
        set_error("arena too large for 32-bit numbering, or no system slots");
        return NB_E_INVALID_ARG;
    }
    NB_HIP(hipGetDevice(&device));
    NB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    if (user_stream) {
        stream = (hipStream_t)user_stream;
        own_stream = false;
    } else {
        NB_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        own_stream = true;
    }
    capc = c;
    law = law_;
    kahan = kahan_;
    hdesc.assign(c.systems, SysDesc{});
    const int S = c.systems;
    tab_cap[T_BOUNDS] = blocks_bound(c.bodies, kBoundsPerBlock, S);
    tab_cap[T_BODY] = blocks_bound(c.bodies, 256, S);
    tab_cap[T_CELLS] = blocks_bound(c.cells, 256, S);
    tab_cap[T_GROUPS] = blocks_bound(c.bodies, kForceVote * kForceWaves, S);
    tab_cap[T_BRUTE] = blocks_bound(c.bodies, 16, S);
    tab_cap[T_UPDATE] = blocks_bound(c.bodies, 256, S);
    tab_cap[T_RANGE] = blocks_bound(c.bodies, 16, S);  // the smallest share per block of either force kernel
    tab_cap[T_FORDER] = tab_cap[T_GROUPS];
    tab_cap[T_KEYS] = blocks_bound(c.bodies, kKeysBodiesPerBlock, S);
    tab_cap[T_FRAMES] = blocks_bound(c.bodies, kFramesBodiesPerBlock, S);
    for (int ph = 0; ph < kBuildPhases; ++ph)
        tab_cap[T_PHASE0 + ph] =
            blocks_bound((c.bodies + kPhaseStride[ph] - 1) / kPhaseStride[ph], build_block(kPhaseStride[ph]), S);

    struct Region { void** dst; size_t bytes; };
    const size_t nn = (size_t)std::max<long long>(c.nodes, 1), nb = (size_t)std::max<long long>(c.bodies, 1),
                 nc = (size_t)std::max<long long>(c.cells, 1);
    std::vector<Region> regs = {
        {(void**)&node4, sizeof(float4) * nn},
        {(void**)&vel4, sizeof(float4) * nb},
        {(void**)&acc4, sizeof(float4) * nb},
        {(void**)&sorted, sizeof(int) * nb},
        {(void**)&forder, sizeof(int) * nb},
        {(void**)&sorder, sizeof(int) * nb},
        {(void**)&child, sizeof(int) * 8 * nc},
        {(void**)&cparent, sizeof(int) * nc},
        {(void**)&cpend, sizeof(int) * nc},
        {(void**)&ccount, sizeof(int) * nc},
        {(void**)&slotoff, sizeof(int) * 8 * nc},
        {(void**)&cdata, sizeof(float4) * 8 * nc},
        {(void**)&kbase, sizeof(int) * 8 * nc},
        {(void**)&lpos, sizeof(float4) * nb},
        {(void**)&ppos, sizeof(float4) * nb},
        {(void**)&lmeta, sizeof(int4) * nb},
        {(void**)&bframe, sizeof(int2) * nb},
        {(void**)&fblist, sizeof(int) * nb},
        {(void**)&d_desc, sizeof(SysDesc) * S},
        {(void**)&d_state, sizeof(SysState) * S},
        {(void**)&d_partial, sizeof(float) * 7 * (size_t)tab_cap[T_BOUNDS]},
        {(void**)&d_slots, sizeof(int) * S},
        {(void**)&gkey, sizeof(int) * (size_t)std::max(tab_cap[T_GROUPS], 1)},
    };
    for (int t = 0; t < T_COUNT; ++t) regs.push_back({(void**)&tab[t], sizeof(BlkWork) * (size_t)tab_cap[t]});
    size_t total = 0;
    for (auto& r : regs) total = align_up(total, 256) + r.bytes;
    arena_bytes = align_up(total, 256);
    NB_HIP(hipMalloc(&arena, arena_bytes));
    size_t off = 0;
    for (auto& r : regs) {
        off = align_up(off, 256);
        *r.dst = (char*)arena + off;
        off += r.bytes;
    }
    NB_HIP(hipMemsetAsync(node4, 0, sizeof(float4) * nn, stream));
    NB_HIP(hipMemsetAsync(vel4, 0, sizeof(float4) * nb, stream));
    NB_HIP(hipMemsetAsync(acc4, 0, sizeof(float4) * nb, stream));
    NB_HIP(hipMemsetAsync(sorted, 0, sizeof(int) * nb, stream));
    NB_HIP(hipMemsetAsync(forder, 0, sizeof(int) * nb, stream));
    NB_HIP(hipMemsetAsync(sorder, 0, sizeof(int) * nb, stream));
    NB_HIP(hipMemsetAsync(child, 0xff, sizeof(int) * 8 * nc, stream));
    NB_HIP(hipMemsetAsync(cpend, 0, sizeof(int) * nc, stream));
    NB_HIP(hipMemsetAsync(ccount, 0, sizeof(int) * nc, stream));
    NB_HIP(hipMemsetAsync(slotoff, 0, sizeof(int) * 8 * nc, stream));
    NB_HIP(hipMemsetAsync(d_state, 0, sizeof(SysState) * S, stream));
    NB_HIP(hipMemsetAsync(gkey, 0, sizeof(int) * (size_t)std::max(tab_cap[T_GROUPS], 1), stream));
    launch_tree_init(stream, S, nullptr, nullptr, d_state, nullptr, nullptr, nullptr);  // slotoff is cleared above
    NB_HIP(hipGetLastError());
    NB_HIP(hipStreamSynchronize(stream));
    return NB_OK;
}

nb_status Engine::activate(const std::vector<Slot>& slots) {
    std::vector<BlkWork> t[T_COUNT];
    std::vector<int> ids;
    std::vector<char> seen(capc.systems, 0);
    long long bf_waves2 = 0;  // brute-force waves of the batch at two targets per lane
    any_bh = any_bf = any_collision = false;
    any_bh_ref = any_bh_std = any_top_global = false;
    long long bh_bodies = 0;  // the batch's Barnes-Hut bodies: the queued build phase shares its budget by them
    for (const Slot& sl : slots)
        if (sl.spec.algo == NB_BARNES_HUT && sl.spec.n > 0) bh_bodies += sl.spec.n;
    for (const Slot& sl : slots) {
        if (sl.slot < 0 || sl.slot >= capc.systems || seen[sl.slot]) {
            set_error("slot index out of range or used twice");
            return NB_E_INVALID_ARG;
        }
        seen[sl.slot] = 1;
        const SysSpec& p = sl.spec;
        const bool bh = p.algo == NB_BARNES_HUT;
        if (sl.node_off + p.n + (bh ? p.cap : 0) > capc.nodes || sl.body_off + p.n > capc.bodies ||
            sl.cell_off + (bh ? p.cap : 0) > capc.cells) {
            set_error("system range outside the arena");
            return NB_E_INVALID_ARG;
        }
        if (bh && p.cap >= (1 << 24)) {
            set_error("cell capacity must be below 2^24 per system");
            return NB_E_INVALID_ARG;
        }
        SysDesc& d = hdesc[sl.slot];
        d = SysDesc{};
        d.n = p.n;
        d.cap = bh ? p.cap : 0;
        d.node_off = (int)sl.node_off;
        d.body_off = (int)sl.body_off;
        d.cell_off = (int)sl.cell_off;
        d.rules = p.rules;
        d.algo = p.algo;
        d.law = p.law;
        d.accum = p.accum;
        d.method = p.method;
        d.G = p.G;
        d.theta = p.theta;
        d.eps = p.eps;
        d.dt = p.dt;
        d.ueps = p.algo == NB_BRUTE_FORCE_LOCAL ? 0.1f : 0.01f;
        d.collision = p.collision ? 1 : 0;
        any_collision = any_collision || (p.collision && d.n > 0);
        const int s = sl.slot;
        ids.push_back(s);
        if (bh && d.n > 0) {
            any_bh = true;
            (d.rules == 0 ? any_bh_ref : any_bh_std) = true;
            if ((d.n + kPhaseStride[0] - 1) / kPhaseStride[0] > kTopBodies) any_top_global = true;
            d.bbox_first = (int)t[T_BOUNDS].size();
            add_ranges(t[T_BOUNDS], s, d.n, kBoundsPerBlock);
            d.bbox_blocks = (int)t[T_BOUNDS].size() - d.bbox_first;
            add_ranges(t[T_BODY], s, d.n, 256);
            add_ranges(t[T_CELLS], s, d.cap, 256);
            d.grp_first = (int)t[T_GROUPS].size();
            add_ranges(t[T_GROUPS], s, d.n, kForceVote * kForceWaves);
            for (int ph = 0; ph < kBuildPhases; ++ph) {
                const int blk = build_block(kPhaseStride[ph]);
                long long items = (d.n + kPhaseStride[ph] - 1) / kPhaseStride[ph];
                // the queued last phase takes its bodies from the system's queue, so its
                // workgroup count is free: 5/8 of the CUs' worth of workgroups contend less
                // for the same slots than one per 1024 bodies (1M: build 0.42 -> 0.35 ms).  A
                // batch shares twice the CU count among its systems by their bodies, at least
                // one workgroup each and at most the single system's 5/8 (512 systems: 91.6 ->
                // 109 ensemble steps/s against one workgroup per 1024 bodies of every system;
                // 107 with 1x, 110 with 4x the CU count)
                if (kPhaseStride[ph] == 1) {
                    const long long cap = std::max(1, cus * 5 / 8), budget = 2LL * cus;
                    const long long share = std::max<long long>(1, (budget * d.n + bh_bodies - 1) / bh_bodies);
                    items = std::min<long long>(items, std::min(share, cap) * blk);
                }
                add_ranges(t[T_PHASE0 + ph], s, (int)items, blk);
            }
            // grid-stride tables of the bucketed phases: block i of the system's k
            const int nk = std::max(1, (d.n + kKeysBodiesPerBlock - 1) / kKeysBodiesPerBlock);
            for (int i = 0; i < nk; ++i) t[T_KEYS].push_back(BlkWork{s, i, nk, 0});
            const int nfr = std::max(1, (d.n + kFramesBodiesPerBlock - 1) / kFramesBodiesPerBlock);
            for (int i = 0; i < nfr; ++i) t[T_FRAMES].push_back(BlkWork{s, i, nfr, 0});
        }
        if (!bh && d.n > 0) {
            any_bf = true;
            bf_shape(d.n, d.bf_split_log2, d.bf_lanes_log2);
            const long long per_wave2 = 128 >> d.bf_lanes_log2;
            bf_waves2 += ((d.n + per_wave2 - 1) / per_wave2) << d.bf_split_log2;
        }
        add_ranges(t[T_UPDATE], s, d.n, 256);
    }
    // two targets per lane only when the batch still fills every CU twice over
    bf_tb = g_tuning.bf_targets_per_lane ? g_tuning.bf_targets_per_lane
                                         : (bf_waves2 >= g_tuning.bf_min_waves ? 2 : 1);
    for (const Slot& sl : slots) {
        const SysDesc& d = hdesc[sl.slot];
        if (d.algo != NB_BARNES_HUT && d.n > 0)
            add_ranges(t[T_BRUTE], sl.slot, d.n, (kBfWaves >> d.bf_split_log2) * (64 >> d.bf_lanes_log2) * bf_tb);
    }
    for (int k = 0; k < T_COUNT; ++k) {
        if ((int)t[k].size() > tab_cap[k]) {
            set_error("block table capacity exceeded");
            return NB_E_INVALID_ARG;
        }
        tab_len[k] = (int)t[k].size();
        if (!t[k].empty())
            NB_HIP(hipMemcpyAsync(tab[k], t[k].data(), sizeof(BlkWork) * t[k].size(), hipMemcpyHostToDevice, stream));
    }
    // longest-first dispatch only where a launch holds at most four waves per SIMD slot
    // (8 waves x 4 SIMDs per CU): with more, the last workgroups matter less than locality
    lpt = tab_len[T_GROUPS] <= kOrderMax && (long long)tab_len[T_GROUPS] * kForceWaves <= 4LL * 32 * cus;
    if (any_bh && any_bf && !side) {
        NB_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
        NB_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        NB_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    NB_HIP(hipMemcpyAsync(d_desc, hdesc.data(), sizeof(SysDesc) * capc.systems, hipMemcpyHostToDevice, stream));
    if (!ids.empty()) NB_HIP(hipMemcpyAsync(d_slots, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, stream));
    NB_HIP(hipStreamSynchronize(stream));
    order_done = false;
    nactive = (int)ids.size();
    active = slots;
    return NB_OK;
}

nb_status Engine::reset_slot(const Slot& s) {
    const bool bh = s.spec.algo == NB_BARNES_HUT;
    if (bh && s.spec.cap > 0) {
        NB_HIP(hipMemsetAsync(child + 8 * s.cell_off, 0xff, sizeof(int) * 8 * (size_t)s.spec.cap, stream));
        NB_HIP(hipMemsetAsync(cpend + s.cell_off, 0, sizeof(int) * (size_t)s.spec.cap, stream));
    }
    SysState z;
    std::memset(&z, 0, sizeof(z));
    z.ncells = 1;
    z.max_depth = 1;
    NB_HIP(hipMemcpyAsync(d_state + s.slot, &z, sizeof(SysState), hipMemcpyHostToDevice, stream));
    NB_HIP(hipStreamSynchronize(stream));
    return NB_OK;
}

nb_status Engine::init(const std::vector<SysSpec>& systems, void* user_stream) {
    Capacity c;
    c.systems = std::max<int>((int)systems.size(), 1);
    std::vector<Slot> slots;
    for (size_t i = 0; i < systems.size(); ++i) {
        const SysSpec& p = systems[i];
        const int cap = p.algo == NB_BARNES_HUT ? p.cap : 0;
        Slot s;
        s.slot = (int)i;
        s.spec = p;
        s.spec.cap = cap;
        s.node_off = c.nodes;
        s.body_off = c.bodies;
        s.cell_off = c.cells;
        c.nodes += (long long)p.n + cap;
        c.bodies += p.n;
        c.cells += cap;
        slots.push_back(s);
    }
    const int law_ = systems.empty() ? 1 : systems.back().law;
    const bool kahan_ = !systems.empty() && systems.back().accum == NB_ACCUM_KAHAN;
    nb_status st = reserve(c, law_, kahan_, user_stream);
    if (st) return st;
    return activate(slots);
}

void Engine::bounds() {
    TraceRange tr("nb:bounds");
    launch_bounds(stream, tab_len[T_BOUNDS], d_desc, d_state, tab[T_BOUNDS], node4, d_partial);
}
void Engine::bounds_reset() {
    TraceRange tr("nb:bounds_reset");
    const bool ord = lpt && tab_len[T_GROUPS] <= kOrderSmall;
    launch_bounds_order(stream, tab_len[T_BOUNDS], tab[T_BOUNDS], d_desc, d_state, node4, d_partial, tab[T_GROUPS],
                        ord ? tab_len[T_GROUPS] : 0, gkey, tab[T_FORDER]);
    order_done = ord;
}
void Engine::tree_reset() {
    TraceRange tr("nb:tree_reset");
    launch_tree_init(stream, nactive, d_slots, d_desc, d_state, slotoff, child, cpend);
}
// The first phase (every 4096th body) in LDS (k_build_top), or by insertion in global memory
// for a system with more first-phase bodies than the LDS build takes; each later phase
// bucketed (nb_build.hip) against the tree of the phases before it; then the queued insertion
// of the bodies that any phase left (see below).
void Engine::build() {
    TraceRange tr("nb:build");
    launch_build_top(stream, nactive, d_slots, d_desc, d_state, node4, child, cparent, cpend, cdata, slotoff,
                     kPhaseStride[0], fblist);
    if (any_top_global)  // only systems over kTopBodies first-phase bodies (S->top 0) take part
        launch_build(stream, tab_len[T_PHASE0], d_desc, d_state, tab[T_PHASE0], node4, child, cparent, cpend,
                     cdata, slotoff, kPhaseStride[0], 0, nullptr);
    for (int ph = 1; ph < kBuildPhases; ++ph) {
        // the phase's bodies: a strided table, or every body (skipping the earlier phases')
        const bool strided = kPhaseStride[ph] > 1;
        const Table bt = strided ? (Table)(T_PHASE0 + ph) : T_BODY;
        const int stride = kPhaseStride[ph], exclude = kPhaseStride[ph - 1];
        launch_classify(stream, tab_len[bt], d_desc, d_state, tab[bt], node4, child, slotoff, ccount, sorted, forder,
                        ppos, sorder, stride, exclude);
        launch_frames(stream, tab_len[T_KEYS], d_desc, d_state, tab[T_KEYS], node4, child, slotoff, ccount, kbase,
                      lpos, lmeta, bframe);
        launch_scatter(stream, tab_len[bt], d_desc, d_state, tab[bt], ppos, kbase, sorted, forder, ccount, sorder,
                       lpos, lmeta, fblist, stride, exclude);
        // the last phase's frames over one workgroup per 256 bodies; a strided phase's few
        // frames over the keys table (one workgroup per 4096 bodies)
        const bool last = ph == kBuildPhases - 1;
        const Table ft = last ? T_FRAMES : T_KEYS;
        launch_frame_build(stream, tab_len[ft], d_desc, d_state, tab[ft], node4, child, cparent, cpend, ccount,
                           slotoff, cdata, lpos, lmeta, bframe, fblist, last);
    }
    // the bodies every phase left (the fallback list: a frame over the LDS pool or the depth
    // limit, a bucket over the frame budget, a first phase over the LDS cell pool), inserted
    // last; usually none, and the launch returns at once
    launch_build(stream, tab_len[T_PHASE0 + kBuildPhases - 1], d_desc, d_state, tab[T_PHASE0 + kBuildPhases - 1],
                 node4, child, cparent, cpend, cdata, slotoff, 1, 0, fblist);
}
void Engine::summarize() {
    TraceRange tr("nb:summarize");
    launch_summarize(stream, tab_len[T_CELLS], d_desc, d_state, tab[T_CELLS], node4, child, cparent, cpend, ccount,
                     slotoff, cdata);
}
void Engine::sort() {
    TraceRange tr("nb:sort");
    launch_sort(stream, tab_len[T_BODY], d_desc, d_state, tab[T_BODY], node4, child, slotoff, sorted, forder,
                sorder);
}
void Engine::force() {
    TraceRange tr("nb:force");
    // the dispatch order from the last full pass's keys: already computed by this step's
    // bounds_reset(), or here
    if (lpt && !order_done) launch_force_order(stream, tab[T_GROUPS], tab_len[T_GROUPS], gkey, tab[T_FORDER]);
    order_done = false;
    for (int pair = 0; pair < 2; ++pair)
        if (pair ? any_bh_ref : any_bh_std)
            launch_force_bh(stream, tab_len[T_GROUPS], kahan, pair != 0, d_desc, d_state,
                            tab[lpt ? T_FORDER : T_GROUPS], node4, child, cdata, forder, acc4, sorder,
                            lpt ? gkey : nullptr);
}
void Engine::force_range(int sys, int lo, int hi) {
    TraceRange tr("nb:force_range");
    const int per = kForceVote * kForceWaves;
    const int start = lo / kForceVote * kForceVote;
    const int stop = (int)std::min<long long>(hdesc[sys].n, ((long long)hi + kForceVote - 1) / kForceVote * kForceVote);
    const int nblk = (stop - start + per - 1) / per;
    launch_range_table(stream, tab[T_RANGE], nblk, sys, start, stop, lo, hi, per);
    launch_force_bh(stream, nblk, kahan, hdesc[sys].rules == 0, d_desc, d_state, tab[T_RANGE], node4, child, cdata,
                    forder, acc4, nullptr, nullptr);
}
void Engine::brute_range(int sys, int lo, int hi) {
    TraceRange tr("nb:brute_force_range");
    const SysDesc& d = hdesc[sys];
    const int per = (kBfWaves >> d.bf_split_log2) * (64 >> d.bf_lanes_log2) * bf_tb;
    const int nblk = (hi - lo + per - 1) / per;
    launch_range_table(stream, tab[T_RANGE], nblk, sys, lo, hi, lo, hi, per);
    launch_brute(stream, nblk, law, kahan, bf_tb, d_desc, d_state, tab[T_RANGE], node4, acc4, vel4, nullptr, -1);
}
void Engine::brute(hipStream_t on) {
    TraceRange tr("nb:brute_force");
    launch_brute(on ? on : stream, tab_len[T_BRUTE], law, kahan, bf_tb, d_desc, d_state, tab[T_BRUTE], node4, acc4,
                 vel4, nullptr, -1);
}
// The brute-force kernel reads and writes only brute-force systems' positions and
// accelerations, which no Barnes-Hut stage touches, so in a mixed batch it runs beside the
// latency-bound summarize and sort instead of after the force pass (forked after the build,
// whose last phase fills every CU in a batch: 512-system ensemble 109.4-109.9 -> 111.0
// steps/s; forked before the bounds, 110.6-110.8).  Errors of the fork and join surface
// through hipGetLastError (or the end of a graph capture), like those of the launches.
void Engine::force_all() {
    const bool beside = any_bh && any_bf && side;
    if (any_bh) {
        bounds_reset();
        build();
    }
    if (beside) {
        (void)hipEventRecord(ev_fork, stream);
        (void)hipStreamWaitEvent(side, ev_fork, 0);
        brute(side);
        (void)hipEventRecord(ev_join, side);
    }
    if (any_bh) {
        summarize();
        sort();
        force();
    }
    if (beside)
        (void)hipStreamWaitEvent(stream, ev_join, 0);
    else if (any_bf)
        brute();
}
void Engine::brute_step(int mode, const float4* in, float4* out) {
    TraceRange tr("nb:brute_force_step");
    launch_brute(stream, tab_len[T_BRUTE], law, kahan, bf_tb, d_desc, d_state, tab[T_BRUTE], in, acc4, vel4, out,
                 mode);
}
void Engine::update(int mode) {
    TraceRange tr("nb:update");
    launch_update(stream, tab_len[T_UPDATE], d_desc, d_state, tab[T_UPDATE], node4, vel4, acc4, mode);
}

void Engine::collide(bool all) {
    TraceRange tr("nb:collision");
    launch_collision(stream, tab_len[T_UPDATE], all, d_desc, d_state, tab[T_UPDATE], node4, vel4);
}

nb_status Engine::download_state(std::vector<SysState>& out) {
    out.resize(capc.systems);
    NB_HIP(hipMemcpyAsync(out.data(), d_state, sizeof(SysState) * capc.systems, hipMemcpyDeviceToHost, stream));
    NB_HIP(hipStreamSynchronize(stream));
    return NB_OK;
}

}  // namespace nb
