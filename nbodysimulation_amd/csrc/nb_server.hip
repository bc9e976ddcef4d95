// nb_server.hip -- continuous-batching N-body server over one HBM arena.
//
// Requests (a system plus a number of steps) are queued on submit.  Every tick
// admits queued requests in arrival order while system slots and arena ranges
// (nodes, bodies, cells: the HBM budget) are free, runs one batched step of
// every resident system, and retires the systems that have finished their
// steps (or failed), staging their final state on the host and returning their
// ranges.  Systems therefore join while others are mid-run and share every
// launch; each system's arithmetic is independent of its neighbours, so its
// result is bit-identical to running it alone.
#include <hip/hip_runtime.h>

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

static nb_status upload_request(nb_server* sv, Request& r) {
    nb::Engine& g = sv->eng;
    nb_status st = g.reset_slot(r.slot);
    if (st) return st;
    const int n = r.spec.n;
    if (n == 0) return NB_OK;
    std::vector<float4> h4(n), hv(n);
    for (int i = 0; i < n; ++i) {
        h4[i] = make_float4(r.pos[3 * i], r.pos[3 * i + 1], r.pos[3 * i + 2], r.mass[i]);
        hv[i] = make_float4(r.vel[3 * i], r.vel[3 * i + 1], r.vel[3 * i + 2], 0.0f);
    }
    NB_HIPS(hipMemcpyAsync(g.node4 + r.slot.node_off, h4.data(), sizeof(float4) * n, hipMemcpyHostToDevice, g.stream));
    NB_HIPS(hipMemcpyAsync(g.vel4 + r.slot.body_off, hv.data(), sizeof(float4) * n, hipMemcpyHostToDevice, g.stream));
    NB_HIPS(hipMemsetAsync(g.acc4 + r.slot.body_off, 0, sizeof(float4) * n, g.stream));
    NB_HIPS(hipStreamSynchronize(g.stream));
    return NB_OK;
}

static nb_status download_request(nb_server* sv, Request& r) {
    nb::Engine& g = sv->eng;
    const int n = r.spec.n;
    r.acc.assign(3 * (size_t)n, 0.0f);
    if (n == 0) return NB_OK;
    std::vector<float4> h(n);
    auto get = [&](const float4* src, std::vector<float>& dst) -> nb_status {
        NB_HIPS(hipMemcpyAsync(h.data(), src, sizeof(float4) * n, hipMemcpyDeviceToHost, g.stream));
        NB_HIPS(hipStreamSynchronize(g.stream));
        for (int i = 0; i < n; ++i) {
            dst[3 * i] = h[i].x;
            dst[3 * i + 1] = h[i].y;
            dst[3 * i + 2] = h[i].z;
        }
        return NB_OK;
    };
    nb_status st = get(g.node4 + r.slot.node_off, r.pos);
    if (!st) st = get(g.vel4 + r.slot.body_off, r.vel);
    if (!st) st = get(g.acc4 + r.slot.body_off, r.acc);
    return st;
}

static void enqueue_force(nb::Engine& g) {
    if (g.any_bh) {
        g.bounds();
        g.tree_reset();
        g.build();
        g.summarize();
        g.sort();
        g.force();
    }
    if (g.any_bf) g.brute();
}

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
    if (st) {
        delete sv;
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
    std::vector<nb::Slot> fresh;
    // admit in arrival order while a slot and arena room are free
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
        nb_status st = upload_request(sv, r);
        if (st) {  // the request stays queued; its ranges go back to the pools
            unplace(sv, r);
            return st;
        }
        fresh.push_back(r.slot);
        sv->active.push_back(std::move(sv->queue.front()));
        sv->queue.pop_front();
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
    if (!sv->active.empty()) {
        const bool lf = sv->method == NB_LEAPFROG;
        if (lf) g.update(2);
        enqueue_force(g);
        g.update(lf ? 1 : 0);
        if (g.any_collision) g.collide();  // the requests that ask for it, after the integrator
        NB_HIPS(hipGetLastError());
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
    out->active = (int32_t)sv->active.size();
    out->queued = (int32_t)sv->queue.size();
    out->finished = (int32_t)sv->finished.size();
    out->reserved = 0;
    out->arena_bytes = (int64_t)sv->eng.arena_bytes;
    out->body_steps = sv->body_steps;
    out->ticks = sv->steps;
    return NB_OK;
}

}  // extern "C"
