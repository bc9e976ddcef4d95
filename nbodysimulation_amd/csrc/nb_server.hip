// nb_server.hip -- continuous-batching N-body server over one HBM arena.
//
// Requests (a system plus a number of steps) are queued on submit.  A tick is a
// pipeline: (1) the requests staged by the previous tick join the resident
// systems; (2) one batched step of every resident system is launched on the
// compute stream; (3) while it runs, queued requests are admitted in arrival
// order while system slots and arena ranges (nodes, bodies, cells: the HBM
// budget) are free, packed into a library-owned pinned staging buffer and
// uploaded on a separate copy stream into ranges the running step does not
// touch (they join at the next tick); (4) the tick waits for the step and
// retires the systems that have finished their steps (or failed), downloading
// their final state through a pinned buffer and returning their ranges.
// Systems therefore join while others are mid-run and share every launch; each
// system's arithmetic is independent of its neighbours, so its result is
// bit-identical to running it alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "nb_engine.h"
#include "nbody/nbody.h"

namespace {

// First-fit allocator of [off, off+len) ranges over [0, total), coalescing on release.
class RangeAlloc {
public:
    void init(long long total) {
        free_.clear();
        if (total > 0) free_[0] = total;
    }
    bool alloc(long long len, long long* off) {
        if (len <= 0) {
            *off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                *off = it->first;
                const long long rest = it->second - len, start = it->first + len;
                free_.erase(it);
                if (rest > 0) free_[start] = rest;
                return true;
            }
        }
        return false;
    }
    void release(long long off, long long len) {
        if (len <= 0) return;
        auto it = free_.emplace(off, len).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<long long, long long> free_;
};

struct Request {
    int64_t id = 0;
    nb::SysSpec spec;
    int steps_left = 0;
    std::vector<float> pos, vel, mass, acc;
    nb::Slot slot;
    int status = NB_OK;
};

}  // namespace

struct nb_server {
    nb::Engine eng;
    RangeAlloc nodes, bodies, cells;
    std::vector<int> free_slots;
    std::deque<std::unique_ptr<Request>> queue;
    std::vector<std::unique_ptr<Request>> active;
    std::vector<std::unique_ptr<Request>> staged;  // uploaded during the last tick; join at the next
    // admission uploads: a copy stream and two pinned staging halves used alternately
    // (one is packed while the other's copy may still run), each with its copy's event
    hipStream_t up = nullptr;
    float4* stage[2] = {};
    size_t stage_cap = 0;  // float4 per half
    hipEvent_t stage_ev[2] = {};
    int stage_half = 0;
    float4* dl = nullptr;  // pinned buffer of retirement downloads
    size_t dl_cap = 0;     // float4
    std::deque<int64_t> finished_order;
    std::map<int64_t, std::unique_ptr<Request>> finished;
    int64_t next_id = 1;
    int64_t body_steps = 0, steps = 0;
    int method = NB_EULER;
    int law = NB_LAW_PLUMMER;
    int accum = NB_ACCUM_PLAIN;
    bool dirty = true;
};

static nb_status sfail(nb_status s, const std::string& m) {
    nb::set_error(m);
    return s;
}

#define NB_HIPS(call)                                                                  \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            nb::set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
            return e_ == hipErrorOutOfMemory ? NB_E_OUT_OF_MEMORY : NB_E_HIP;          \
        }                                                                              \
    } while (0)

static int cells_for(const nb::SysSpec& s) { return s.algo == NB_BARNES_HUT ? s.cap : 0; }

// Place r in the arena if a slot and all three ranges are free.
static bool try_place(nb_server* sv, Request& r) {
    if (sv->free_slots.empty()) return false;
    const long long nn = (long long)r.spec.n + cells_for(r.spec), nb = r.spec.n, nc = cells_for(r.spec);
    long long on = 0, ob = 0, oc = 0;
    if (!sv->nodes.alloc(nn, &on)) return false;
    if (!sv->bodies.alloc(nb, &ob)) {
        sv->nodes.release(on, nn);
        return false;
    }
    if (!sv->cells.alloc(nc, &oc)) {
        sv->nodes.release(on, nn);
        sv->bodies.release(ob, nb);
        return false;
    }
    r.slot.slot = sv->free_slots.back();
    sv->free_slots.pop_back();
    r.slot.spec = r.spec;
    r.slot.node_off = on;
    r.slot.body_off = ob;
    r.slot.cell_off = oc;
    return true;
}

static void unplace(nb_server* sv, Request& r) {
    sv->nodes.release(r.slot.node_off, (long long)r.spec.n + cells_for(r.spec));
    sv->bodies.release(r.slot.body_off, r.spec.n);
    sv->cells.release(r.slot.cell_off, cells_for(r.spec));
    sv->free_slots.push_back(r.slot.slot);
}

constexpr size_t kStateF4 = sizeof(nb::SysState) / sizeof(float4);
static_assert(sizeof(nb::SysState) % sizeof(float4) == 0, "SysState fills whole float4s");

// Pinned host memory of at least `want` float4 in *buf (grown, never shrunk).
static nb_status ensure_pinned(float4** buf, size_t* cap, size_t want) {
    if (*cap >= want) return NB_OK;
    if (*buf) (void)hipHostFree(*buf);
    *buf = nullptr;
    *cap = 0;
    const size_t n = std::max(want, (size_t)1 << 16);
    NB_HIPS(hipHostMalloc(reinterpret_cast<void**>(buf), sizeof(float4) * n, hipHostMallocDefault));
    *cap = n;
    return NB_OK;
}

// Queue the upload of a placed request on the copy stream: a fresh slot state, its
// cell slots cleared, positions+masses and velocities, zero accelerations.  The
// source is one half of the pinned staging buffer; the half's previous copy is
// waited for first.  Nothing the running step reads is written.
static nb_status stage_request(nb_server* sv, Request& r) {
    nb::Engine& g = sv->eng;
    const int n = r.spec.n;
    const size_t want = 2 * (size_t)n + kStateF4;
    if (want > sv->stage_cap) {  // grow both halves: their copies must be done first
        NB_HIPS(hipStreamSynchronize(sv->up));
        for (int h = 0; h < 2; ++h) {
            size_t cap = sv->stage_cap;
            nb_status st = ensure_pinned(&sv->stage[h], &cap, want);
            if (st) return st;
            if (h == 1) sv->stage_cap = cap;
        }
    }
    const int half = sv->stage_half;
    sv->stage_half ^= 1;
    NB_HIPS(hipEventSynchronize(sv->stage_ev[half]));
    float4* h4 = sv->stage[half];
    float4* hv = h4 + n;
    nb::SysState* z = reinterpret_cast<nb::SysState*>(hv + n);
    std::memset(z, 0, sizeof(*z));
    z->ncells = 1;
    z->max_depth = 1;
    for (int i = 0; i < n; ++i) {
        h4[i] = make_float4(r.pos[3 * i], r.pos[3 * i + 1], r.pos[3 * i + 2], r.mass[i]);
        hv[i] = make_float4(r.vel[3 * i], r.vel[3 * i + 1], r.vel[3 * i + 2], 0.0f);
    }
    const nb::Slot& sl = r.slot;
    if (r.spec.algo == NB_BARNES_HUT && r.spec.cap > 0) {
        NB_HIPS(hipMemsetAsync(g.child + 8 * sl.cell_off, 0xff, sizeof(int) * 8 * (size_t)r.spec.cap, sv->up));
        NB_HIPS(hipMemsetAsync(g.cpend + sl.cell_off, 0, sizeof(int) * (size_t)r.spec.cap, sv->up));
    }
    NB_HIPS(hipMemcpyAsync(g.d_state + sl.slot, z, sizeof(*z), hipMemcpyHostToDevice, sv->up));
    if (n > 0) {
        NB_HIPS(hipMemcpyAsync(g.node4 + sl.node_off, h4, sizeof(float4) * n, hipMemcpyHostToDevice, sv->up));
        NB_HIPS(hipMemcpyAsync(g.vel4 + sl.body_off, hv, sizeof(float4) * n, hipMemcpyHostToDevice, sv->up));
        NB_HIPS(hipMemsetAsync(g.acc4 + sl.body_off, 0, sizeof(float4) * n, sv->up));
    }
    NB_HIPS(hipEventRecord(sv->stage_ev[half], sv->up));
    return NB_OK;
}

// The final state of a retired request, through the pinned download buffer.
static nb_status download_request(nb_server* sv, Request& r) {
    nb::Engine& g = sv->eng;
    const int n = r.spec.n;
    r.acc.assign(3 * (size_t)n, 0.0f);
    if (n == 0) return NB_OK;
    nb_status st = ensure_pinned(&sv->dl, &sv->dl_cap, 3 * (size_t)n);
    if (st) return st;
    float4* h = sv->dl;
    NB_HIPS(hipMemcpyAsync(h, g.node4 + r.slot.node_off, sizeof(float4) * n, hipMemcpyDeviceToHost, g.stream));
    NB_HIPS(hipMemcpyAsync(h + n, g.vel4 + r.slot.body_off, sizeof(float4) * n, hipMemcpyDeviceToHost, g.stream));
    NB_HIPS(hipMemcpyAsync(h + 2 * (size_t)n, g.acc4 + r.slot.body_off, sizeof(float4) * n, hipMemcpyDeviceToHost,
                           g.stream));
    NB_HIPS(hipStreamSynchronize(g.stream));
    std::vector<float>* dst[3] = {&r.pos, &r.vel, &r.acc};
    for (int a = 0; a < 3; ++a) {
        const float4* src = h + (size_t)a * n;
        float* o = dst[a]->data();
        for (int i = 0; i < n; ++i) {
            o[3 * i] = src[i].x;
            o[3 * i + 1] = src[i].y;
            o[3 * i + 2] = src[i].z;
        }
    }
    return NB_OK;
}

static void enqueue_force(nb::Engine& g) { g.force_all(); }

extern "C" {

nb_status nb_server_create(int64_t hbm_budget_bytes, int32_t max_systems, int32_t method, int32_t law,
                           int32_t accumulation, void* stream, nb_server** out) {
    if (!out || max_systems < 1 || hbm_budget_bytes <= 0) return sfail(NB_E_INVALID_ARG, "bad arguments");
    if (method != NB_EULER && method != NB_LEAPFROG) return sfail(NB_E_INVALID_ARG, "bad method");
    if (law != NB_LAW_PLUMMER && law != NB_LAW_REFERENCE_UNIT) return sfail(NB_E_INVALID_ARG, "bad law");
    if (accumulation != NB_ACCUM_PLAIN && accumulation != NB_ACCUM_KAHAN)
        return sfail(NB_E_INVALID_ARG, "bad accumulation");
    *out = nullptr;
    // A Barnes-Hut system of N bodies takes N + cap nodes, N bodies, cap = 2N + 1024 cells.
    const double per_body =
        3.0 * nb::kBytesPerNode + nb::kBytesPerBody + 2.0 * nb::kBytesPerCell;
    nb::Capacity c;
    c.systems = max_systems;
    c.bodies = (long long)((double)hbm_budget_bytes / per_body);
    c.nodes = 3 * c.bodies + 1024LL * max_systems;
    c.cells = 2 * c.bodies + 1024LL * max_systems;
    if (c.bodies < 1) return sfail(NB_E_INVALID_ARG, "budget too small");
    nb_server* sv = new nb_server();
    sv->method = method;
    sv->law = law;
    sv->accum = accumulation;
    nb_status st = sv->eng.reserve(c, law, accumulation == NB_ACCUM_KAHAN, stream);
    if (!st && hipStreamCreateWithFlags(&sv->up, hipStreamNonBlocking) != hipSuccess) {
        sv->up = nullptr;
        st = sfail(NB_E_HIP, "hipStreamCreateWithFlags(copy stream) failed");
    }
    for (int h = 0; h < 2 && !st; ++h)
        if (hipEventCreateWithFlags(&sv->stage_ev[h], hipEventDisableTiming) != hipSuccess) {
            sv->stage_ev[h] = nullptr;
            st = sfail(NB_E_HIP, "hipEventCreateWithFlags failed");
        }
    if (st) {
        nb_server_destroy(sv);
        return st;
    }
    sv->nodes.init(c.nodes);
    sv->bodies.init(c.bodies);
    sv->cells.init(c.cells);
    for (int s = max_systems - 1; s >= 0; --s) sv->free_slots.push_back(s);
    *out = sv;
    return NB_OK;
}

nb_status nb_server_destroy(nb_server* sv) {
    if (!sv) return NB_OK;
    if (sv->eng.stream) (void)hipStreamSynchronize(sv->eng.stream);
    if (sv->up) (void)hipStreamSynchronize(sv->up);
    for (int h = 0; h < 2; ++h) {
        if (sv->stage_ev[h]) (void)hipEventDestroy(sv->stage_ev[h]);
        if (sv->stage[h]) (void)hipHostFree(sv->stage[h]);
    }
    if (sv->dl) (void)hipHostFree(sv->dl);
    if (sv->up) (void)hipStreamDestroy(sv->up);
    delete sv;
    return NB_OK;
}

nb_status nb_server_submit(nb_server* sv, const nb_system_spec* spec, int32_t steps, const float* pos,
                           const float* vel, const float* mass, int64_t* id) {
    if (!sv || !spec || steps < 0 || !id) return sfail(NB_E_INVALID_ARG, "bad arguments");
    auto r = std::make_unique<Request>();
    nb_status vs = nb::spec_from_system(*spec, sv->method, sv->law, sv->accum, &r->spec);
    if (vs) return vs;
    const int n = r->spec.n;
    const bool bh = r->spec.algo == NB_BARNES_HUT;
    if (n > 0 && (!pos || !mass)) return sfail(NB_E_INVALID_ARG, "NULL state");
    const nb::Capacity& c = sv->eng.capc;
    if ((long long)n + r->spec.cap > c.nodes || n > c.bodies || r->spec.cap > c.cells)
        return sfail(NB_E_ADMISSION, "request larger than the whole HBM budget");
    for (int i = 0; i < n; ++i) {
        const bool ok = std::isfinite(mass[i]) && (bh ? mass[i] > 0.0f : mass[i] >= 0.0f);
        if (!ok) return sfail(NB_E_INVALID_ARG, "mass must be finite and > 0");
    }
    r->pos.assign(pos, pos + 3 * (size_t)n);
    r->vel.assign(3 * (size_t)n, 0.0f);
    if (vel) r->vel.assign(vel, vel + 3 * (size_t)n);
    r->mass.assign(mass, mass + n);
    r->steps_left = steps;
    r->id = sv->next_id++;
    *id = r->id;
    sv->queue.push_back(std::move(r));
    return NB_OK;
}

nb_status nb_server_tick(nb_server* sv, int32_t* completed) {
    nb::TraceRange tr("nb_server_tick");
    if (!sv) return sfail(NB_E_INVALID_ARG, "NULL server");
    if (completed) *completed = 0;
    int done = 0;
    nb::Engine& g = sv->eng;
    // (1) the requests staged by the last tick join (their uploads were queued on the
    // copy stream while that tick's step ran)
    std::vector<nb::Slot> fresh;
    if (!sv->staged.empty()) {
        NB_HIPS(hipStreamSynchronize(sv->up));
        for (auto& r : sv->staged) {
            fresh.push_back(r->slot);
            sv->active.push_back(std::move(r));
        }
        sv->staged.clear();
        sv->dirty = true;
    }
    if (sv->method == NB_LEAPFROG && !fresh.empty()) {
        // leapfrog starts from the force at the initial positions
        nb_status st = g.activate(fresh);
        if (st) return st;
        enqueue_force(g);
        sv->dirty = true;
    }
    if (sv->dirty) {
        std::vector<nb::Slot> slots;
        for (auto& a : sv->active) slots.push_back(a->slot);
        nb_status st = g.activate(slots);
        if (st) return st;
        sv->dirty = false;
    }
    // (2) one step of every resident system, asynchronously
    const bool stepping = !sv->active.empty();
    if (stepping) {
        const bool lf = sv->method == NB_LEAPFROG;
        if (lf) g.update(2);
        enqueue_force(g);
        g.update(lf ? 1 : 0);
        if (g.any_collision) g.collide();  // the requests that ask for it, after the integrator
        NB_HIPS(hipGetLastError());
    }
    // (3) admit in arrival order while a slot and arena room are free; the uploads run
    // on the copy stream beside the step
    while (!sv->queue.empty()) {
        Request& r = *sv->queue.front();
        if (r.steps_left == 0) {  // nothing to run: retired on arrival with its input state
            r.acc.assign(3 * (size_t)r.spec.n, 0.0f);
            sv->finished_order.push_back(r.id);
            sv->finished[r.id] = std::move(sv->queue.front());
            sv->queue.pop_front();
            ++done;
            continue;
        }
        if (!try_place(sv, r)) break;
        nb_status st = stage_request(sv, r);
        if (st) {  // the request stays queued; its ranges go back to the pools
            unplace(sv, r);
            if (stepping) (void)hipStreamSynchronize(g.stream);
            return st;
        }
        sv->staged.push_back(std::move(sv->queue.front()));
        sv->queue.pop_front();
    }
    // (4) wait for the step, then retire
    if (stepping) {
        std::vector<nb::SysState> s;
        nb_status st = g.download_state(s);
        if (st) return st;
        ++sv->steps;
        std::vector<std::unique_ptr<Request>> keep;
        for (size_t i = 0; i < sv->active.size(); ++i) {
            auto& a = sv->active[i];
            const int bits = s[a->slot.slot].status;
            if (!bits) {
                --a->steps_left;
                sv->body_steps += a->spec.n;
            }
            if (bits || a->steps_left <= 0) {
                a->status = bits ? nb::status_from_bits(bits) : NB_OK;
                st = download_request(sv, *a);
                if (st) {  // keep this request and the unvisited ones in the active list
                    for (size_t j = i; j < sv->active.size(); ++j) keep.push_back(std::move(sv->active[j]));
                    sv->active = std::move(keep);
                    return st;
                }
                unplace(sv, *a);
                sv->finished_order.push_back(a->id);
                sv->finished[a->id] = std::move(a);
                sv->dirty = true;
                ++done;
            } else {
                keep.push_back(std::move(a));
            }
        }
        sv->active = std::move(keep);
    }
    if (completed) *completed = done;
    return NB_OK;
}

nb_status nb_server_poll(nb_server* sv, int64_t* id, int32_t* status) {
    if (!sv || !id) return sfail(NB_E_INVALID_ARG, "bad arguments");
    if (sv->finished_order.empty()) return sfail(NB_E_STATE, "no finished request");
    *id = sv->finished_order.front();
    sv->finished_order.pop_front();
    if (status) *status = sv->finished[*id]->status;
    return NB_OK;
}

nb_status nb_server_fetch(nb_server* sv, int64_t id, float* pos, float* vel, float* acc) {
    if (!sv) return sfail(NB_E_INVALID_ARG, "NULL server");
    auto it = sv->finished.find(id);
    if (it == sv->finished.end()) return sfail(NB_E_STATE, "request not finished or already fetched");
    Request& r = *it->second;
    const size_t k = 3 * (size_t)r.spec.n;
    if (pos) std::memcpy(pos, r.pos.data(), sizeof(float) * k);
    if (vel) std::memcpy(vel, r.vel.data(), sizeof(float) * k);
    if (acc) std::memcpy(acc, r.acc.data(), sizeof(float) * k);
    sv->finished.erase(it);
    return NB_OK;
}

nb_status nb_server_stats_get(nb_server* sv, nb_server_stats* out) {
    if (!sv || !out) return sfail(NB_E_INVALID_ARG, "bad arguments");
    out->active = (int32_t)(sv->active.size() + sv->staged.size());  // staged requests join at the next tick
    out->queued = (int32_t)sv->queue.size();
    out->finished = (int32_t)sv->finished.size();
    out->reserved = 0;
    out->arena_bytes = (int64_t)sv->eng.arena_bytes;
    out->body_steps = sv->body_steps;
    out->ticks = sv->steps;
    return NB_OK;
}

}  // extern "C"
