// nb_engine.hip -- host runtime: one HBM arena per batch, block tables, stage launches.
#include "nb_engine.h"

#include <algorithm>
#include <cstring>

namespace nb {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

static constexpr int kPhaseStride[Engine::kBuildPhases] = {1024, 32, 1};

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
    if (arena) (void)hipFree(arena);
    arena = nullptr;
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    own_stream = false;
}

// Block table over `count` items per system, `per` items per workgroup.
static void add_ranges(std::vector<BlkWork>& t, int sys, int count, int per) {
    for (int b = 0; b < count; b += per) t.push_back(BlkWork{sys, b, std::min(count, b + per), 0});
}

nb_status Engine::init(const std::vector<SysSpec>& systems, void* user_stream) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        set_error("no HIP device");
        return NB_E_NO_DEVICE;
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
    spec = systems;
    nsys = (int)systems.size();
    hdesc.assign(nsys, SysDesc{});
    tot_nodes = tot_bodies = tot_cells = 0;
    any_bh = any_bf = false;
    std::vector<BlkWork> tb, tbu, tc, tg, tf, tu;
    std::vector<BlkWork> tph[kBuildPhases];
    for (int s = 0; s < nsys; ++s) {
        const SysSpec& p = systems[s];
        SysDesc& d = hdesc[s];
        const bool bh = p.algo == NB_BARNES_HUT;
        d.n = p.n;
        d.cap = bh ? p.cap : 0;
        d.node_off = (int)tot_nodes;
        d.body_off = (int)tot_bodies;
        d.cell_off = (int)tot_cells;
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
        tot_nodes += (long long)d.n + d.cap;
        tot_bodies += d.n;
        tot_cells += d.cap;
        if (bh && d.n > 0) {
            any_bh = true;
            d.bbox_first = (int)tb.size();
            // ~2048 bodies per bounds workgroup (8 per lane)
            add_ranges(tb, s, d.n, 2048);
            d.bbox_blocks = (int)tb.size() - d.bbox_first;
            add_ranges(tbu, s, d.n, 256);
            for (int ph = 0; ph < kBuildPhases; ++ph) {
                const int stride = kPhaseStride[ph];
                add_ranges(tph[ph], s, (d.n + stride - 1) / stride, 1024);  // k_build workgroup size
            }
            add_ranges(tc, s, d.cap, 256);
            add_ranges(tg, s, d.n, 256);
        }
        if (!bh && d.n > 0) {
            any_bf = true;
            add_ranges(tf, s, d.n, 256);
        }
        add_ranges(tu, s, d.n, 256);
        law = p.law;
        kahan = p.accum == NB_ACCUM_KAHAN;
    }
    if (tot_nodes > 0x7fffffffLL || tot_cells * 8 > 0x7fffffffLL * 4LL) {
        set_error("batch too large for 32-bit node numbering");
        return NB_E_INVALID_ARG;
    }
    for (const SysDesc& d : hdesc) {
        if (d.cap >= (1 << 24)) {
            set_error("cell capacity must be below 2^24 per system");
            return NB_E_INVALID_ARG;
        }
    }
    nb_bounds = (int)tb.size();
    nb_build = (int)tbu.size();
    nb_cells = (int)tc.size();
    nb_groups = (int)tg.size();
    nb_brute = (int)tf.size();
    nb_update = (int)tu.size();

    // one arena, 256-byte aligned regions
    struct Region { void** dst; size_t bytes; };
    const size_t nn = (size_t)std::max<long long>(tot_nodes, 1), nbd = (size_t)std::max<long long>(tot_bodies, 1),
                 nc = (size_t)std::max<long long>(tot_cells, 1);
    std::vector<Region> regs = {
        {(void**)&node4, sizeof(float4) * nn},
        {(void**)&vel4, sizeof(float4) * nbd},
        {(void**)&acc4, sizeof(float4) * nbd},
        {(void**)&sorted, sizeof(int) * nbd},
        {(void**)&child, sizeof(int) * 8 * nc},
        {(void**)&cparent, sizeof(int) * nc},
        {(void**)&cpend, sizeof(int) * nc},
        {(void**)&ccount, sizeof(int) * nc},
        {(void**)&slotoff, sizeof(int) * 8 * nc},
        {(void**)&cdata, sizeof(float4) * 8 * nc},
        {(void**)&t_phase[0], sizeof(BlkWork) * std::max<size_t>(tph[0].size(), 1)},
        {(void**)&t_phase[1], sizeof(BlkWork) * std::max<size_t>(tph[1].size(), 1)},
        {(void**)&t_phase[2], sizeof(BlkWork) * std::max<size_t>(tph[2].size(), 1)},
        {(void**)&d_desc, sizeof(SysDesc) * std::max(nsys, 1)},
        {(void**)&d_state, sizeof(SysState) * std::max(nsys, 1)},
        {(void**)&d_partial, sizeof(float) * 7 * std::max(nb_bounds, 1)},
        {(void**)&t_bounds, sizeof(BlkWork) * std::max(nb_bounds, 1)},
        {(void**)&t_build, sizeof(BlkWork) * std::max(nb_build, 1)},
        {(void**)&t_cells, sizeof(BlkWork) * std::max(nb_cells, 1)},
        {(void**)&t_groups, sizeof(BlkWork) * std::max(nb_groups, 1)},
        {(void**)&t_brute, sizeof(BlkWork) * std::max(nb_brute, 1)},
        {(void**)&t_update, sizeof(BlkWork) * std::max(nb_update, 1)},
    };
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
    NB_HIP(hipMemsetAsync(vel4, 0, sizeof(float4) * nbd, stream));
    NB_HIP(hipMemsetAsync(acc4, 0, sizeof(float4) * nbd, stream));
    NB_HIP(hipMemsetAsync(child, 0xff, sizeof(int) * 8 * nc, stream));
    NB_HIP(hipMemsetAsync(cpend, 0, sizeof(int) * nc, stream));
    NB_HIP(hipMemsetAsync(sorted, 0, sizeof(int) * nbd, stream));
    NB_HIP(hipMemsetAsync(ccount, 0, sizeof(int) * nc, stream));
    NB_HIP(hipMemsetAsync(slotoff, 0, sizeof(int) * 8 * nc, stream));
    NB_HIP(hipMemsetAsync(d_state, 0, sizeof(SysState) * std::max(nsys, 1), stream));
    NB_HIP(hipMemcpyAsync(d_desc, hdesc.data(), sizeof(SysDesc) * nsys, hipMemcpyHostToDevice, stream));
    auto up = [&](BlkWork* dst, const std::vector<BlkWork>& v) -> hipError_t {
        if (v.empty()) return hipSuccess;
        return hipMemcpyAsync(dst, v.data(), sizeof(BlkWork) * v.size(), hipMemcpyHostToDevice, stream);
    };
    NB_HIP(up(t_bounds, tb));
    NB_HIP(up(t_build, tbu));
    for (int ph = 0; ph < kBuildPhases; ++ph) {
        nb_phase[ph] = (int)tph[ph].size();
        NB_HIP(up(t_phase[ph], tph[ph]));
    }
    NB_HIP(up(t_cells, tc));
    NB_HIP(up(t_groups, tg));
    NB_HIP(up(t_brute, tf));
    NB_HIP(up(t_update, tu));
    launch_tree_init(stream, nsys, d_state);
    NB_HIP(hipGetLastError());
    NB_HIP(hipStreamSynchronize(stream));
    return NB_OK;
}

void Engine::bounds() {
    launch_bounds(stream, nb_bounds, d_desc, d_state, t_bounds, node4, d_partial);
}
void Engine::tree_reset() {
    launch_tree_clear(stream, nb_cells, d_desc, d_state, t_cells, child, cpend);
    launch_tree_init(stream, nsys, d_state);
}
// Insertion phases: every 1024th body, then every 32nd, then the rest.
void Engine::build() {
    for (int ph = 0; ph < kBuildPhases; ++ph)
        launch_build(stream, nb_phase[ph], d_desc, d_state, t_phase[ph], node4, child, cparent, cpend,
                     kPhaseStride[ph], ph == 0 ? 0 : kPhaseStride[ph - 1]);
}
void Engine::summarize() {
    launch_summarize(stream, nb_cells, d_desc, d_state, t_cells, node4, child, cparent, cpend, ccount, slotoff,
                     cdata);
}
void Engine::sort() {
    launch_sort(stream, nb_build, d_desc, d_state, t_build, node4, child, slotoff, sorted);
}
void Engine::force() {
    launch_force_bh(stream, nb_groups, kahan, d_desc, d_state, t_groups, node4, child, cdata, sorted, acc4);
}
void Engine::brute() {
    launch_brute(stream, nb_brute, law, kahan, d_desc, d_state, t_brute, node4, acc4);
}
void Engine::update(int mode) {
    launch_update(stream, nb_update, d_desc, d_state, t_update, node4, vel4, acc4, mode);
}

nb_status Engine::download_state(std::vector<SysState>& out) {
    out.resize(nsys);
    NB_HIP(hipMemcpyAsync(out.data(), d_state, sizeof(SysState) * nsys, hipMemcpyDeviceToHost, stream));
    NB_HIP(hipStreamSynchronize(stream));
    return NB_OK;
}

}  // namespace nb
