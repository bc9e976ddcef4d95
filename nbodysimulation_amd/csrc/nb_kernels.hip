// nb_kernels.hip -- gfx950 kernels of the N-body stage pipeline.
//
// Stages (reference kernel each one replaces):
//   k_bounds      hybrid_reduce_root + parallel_reduce_root  kernels/boundary_reduce.cl:6-159
//   k_tree_clear  init (child reset)                         kernels/init.cl:2-36
//   k_tree_init   init (root, bottom)                        kernels/init.cl:22-34
//   k_bounds_clear  k_bounds + k_tree_clear (+ the force pass's dispatch order) in one launch
//                 of a step (the root's re-seeding then done by k_build_top)
//   k_build       build_tree                                 kernels/build_tree.cl:6-165
//   k_summarize   saturate_tree                              kernels/saturate_tree.cl:6-146
//   k_sort        sort (empty in the reference)              kernels/sort.cl
//   k_force_bh    calculate_force_local traversal            kernels/calculate_forces.cl:44-148
//   k_brute       update / update_local all-pairs            kernels/brute_force_update.cl:3-228
//   k_update      integrator branch                          kernels/calculate_forces.cl:150-182
//   k_collision   update_collision_local                     kernels/collision_update.cl:3-83
//   k_vertices    copy_vertices                              kernels/copy_vertices.cl:3-17
//   k_range_table block table of one target range (no counterpart: one GPU's share of a force pass)
//   k_force_order longest-first dispatch order of the force workgroups (no counterpart)
//   k_energy_f64, k_energy_final   FP64 conservation diagnostics (no counterpart)
//
// Memory-model notes (gfx950 has one L2 per XCD):
//   * Data that one workgroup writes and another reads inside the same kernel is
//     accessed with agent-scope atomics (sc1 loads/stores).  The insertion kernel
//     reads child slots that way and relies on compare-and-swap for correctness: a
//     read that races a writer returns an older value of that slot, which makes the
//     CAS fail and the body retry.  No pass clears the cell pool between steps: a
//     cell's creator writes all eight slots before the cell becomes reachable.
//   * A lock holder publishes a new subtree with one release store.
//   * No lane ever waits inside an iteration for another lane of its own wave:
//     every retry loop makes one attempt per iteration.
#include <hip/hip_runtime.h>
#include <math.h>

#include "nb_device.h"
#include "nb_internal.h"

namespace nb {

// ---------------------------------------------------------------------------
// bounds: exact min/max over all bodies of each system, canonical +0, plus the
// per-level opening thresholds of calculate_forces.cl:45-62.  One pass: each
// workgroup reduces its range (DPP inside a wave, LDS across waves), the last
// workgroup of a system (ticket) reduces the partials.

// One workgroup of the bounds pass (256 lanes): w its range, blk its place in the bounds table.
__device__ __forceinline__ void bounds_block(const SysDesc* __restrict__ sd, SysState* st, const BlkWork w, int blk,
                                             const float4* __restrict__ node4, float* partial) {
    __shared__ float red[4][7];
    __shared__ int is_last;
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    const float4* nodes = node4 + d.node_off;
    float lo0 = INFINITY, lo1 = INFINITY, lo2 = INFINITY;
    float hi0 = -INFINITY, hi1 = -INFINITY, hi2 = -INFINITY;
    float bad = 0.0f;
    // eight positions per lane in flight at once (issued together, then reduced), not one
    // load and its wait per body
    constexpr int kPer = 8;
    for (int i0 = w.begin + (int)threadIdx.x; i0 < w.end; i0 += kPer * (int)blockDim.x) {
        float4 q[kPer];
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            const int i = i0 + j * (int)blockDim.x;
            q[j] = nodes[min(i, w.end - 1)];
        }
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
            if (i0 + j * (int)blockDim.x >= w.end) continue;
            const float4 p = q[j];
            if (finite3(p)) {
                lo0 = fminf(lo0, p.x); lo1 = fminf(lo1, p.y); lo2 = fminf(lo2, p.z);
                hi0 = fmaxf(hi0, p.x); hi1 = fmaxf(hi1, p.y); hi2 = fmaxf(hi2, p.z);
            } else {
                bad = 1.0f;
            }
        }
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    bad = wave_max(bad);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[wave][0] = lo0; red[wave][1] = lo1; red[wave][2] = lo2;
        red[wave][3] = hi0; red[wave][4] = hi1; red[wave][5] = hi2; red[wave][6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float r[7];
        for (int k = 0; k < 7; ++k) r[k] = red[0][k];
        for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
            for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], red[q][k]);
            for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], red[q][k]);
        }
        for (int k = 0; k < 7; ++k) st_cf(partial + 7 * (size_t)blk + k, r[k]);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        const int t = __hip_atomic_fetch_add(&S->ticket, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == d.bbox_blocks - 1);
    }
    __syncthreads();
    if (!is_last) return;
    // last workgroup of the system: reduce the partials (coherent loads)
    lo0 = lo1 = lo2 = INFINITY;
    hi0 = hi1 = hi2 = -INFINITY;
    bad = 0.0f;
    for (int q = threadIdx.x; q < d.bbox_blocks; q += blockDim.x) {
        const float* pp = partial + 7 * (size_t)(d.bbox_first + q);
        lo0 = fminf(lo0, ld_cf(pp)); lo1 = fminf(lo1, ld_cf(pp + 1)); lo2 = fminf(lo2, ld_cf(pp + 2));
        hi0 = fmaxf(hi0, ld_cf(pp + 3)); hi1 = fmaxf(hi1, ld_cf(pp + 4)); hi2 = fmaxf(hi2, ld_cf(pp + 5));
        bad = fmaxf(bad, ld_cf(pp + 6));
    }
    lo0 = wave_min(lo0); lo1 = wave_min(lo1); lo2 = wave_min(lo2);
    hi0 = wave_max(hi0); hi1 = wave_max(hi1); hi2 = wave_max(hi2);
    bad = wave_max(bad);
    __syncthreads();
    if (lane == 0) {
        red[wave][0] = lo0; red[wave][1] = lo1; red[wave][2] = lo2;
        red[wave][3] = hi0; red[wave][4] = hi1; red[wave][5] = hi2; red[wave][6] = bad;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    float r[7];
    for (int k = 0; k < 7; ++k) r[k] = red[0][k];
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) {
        for (int k = 0; k < 3; ++k) r[k] = fminf(r[k], red[q][k]);
        for (int k = 3; k < 7; ++k) r[k] = fmaxf(r[k], red[q][k]);
    }
    for (int k = 0; k < 6; ++k) S->box[k] = r[k] + 0.0f;
    if (r[6] != 0.0f) atomicOr(&S->status, kStNonfinite);
    // opening thresholds per level (calculate_forces.cl:45-62 for REFERENCE rules)
    float rad = 0.5f * fmaxf(fmaxf(S->box[3] - S->box[0], S->box[4] - S->box[1]), S->box[5] - S->box[2]);
    if (d.rules == 0) {
        float q = d.theta > 0.0f ? rad * rad / d.theta : rad * rad;
        for (int l = 0; l < kDqLen; ++l) {
            if (l) q = 0.25f * q;
            S->dq[l] = q + d.eps;
        }
    } else if (d.theta > 0.0f) {
        float s = rad / d.theta;
        float e2 = d.eps * d.eps;
        float q = s * s;
        for (int l = 0; l < kDqLen; ++l) {
            if (l) q = 0.25f * q;
            S->dq[l] = q + e2;
        }
    } else {
        for (int l = 0; l < kDqLen; ++l) S->dq[l] = INFINITY;
    }
    S->ticket = 0;
}

__global__ __launch_bounds__(256) void k_bounds(const SysDesc* __restrict__ sd, SysState* st,
                                                const BlkWork* __restrict__ work,
                                                const float4* __restrict__ node4, float* partial) {
    bounds_block(sd, st, work[blockIdx.x], blockIdx.x, node4, partial);
}

// ---------------------------------------------------------------------------
// reset: clear the slots of the cells the previous build used (so every unused
// cell holds -1 slots), then re-seed the root.

__global__ void k_tree_init(int nsys, const int* __restrict__ slots, const SysDesc* __restrict__ sd, SysState* st,
                            int* slotoff, int* child, int* cpend) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nsys) return;
    const int s = slots ? slots[k] : k;
    if (sd && slotoff && sd[s].algo == 0 && sd[s].cap > 0) {  // an empty root: slots, counts
        int4* so = reinterpret_cast<int4*>(slotoff + 8 * (size_t)sd[s].cell_off);
        so[0] = make_int4(0, 0, 0, 0);
        so[1] = make_int4(0, 0, 0, 0);
        int4* c4 = reinterpret_cast<int4*>(child + 8 * (size_t)sd[s].cell_off);
        c4[0] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
        c4[1] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
        cpend[sd[s].cell_off] = 0;
    }
    st[s].ncells = 1;
    st[s].max_depth = 1;  // build_tree.cl:28
    st[s].bq = 0;
    st[s].top = 0;
    st[s].nfb = 0;  // the build's fallback list: its phases append, its queued insertion empties it
}

// ---------------------------------------------------------------------------
// build: lock-free insertion (build_tree.cl:42-161), one body per lane.
// Each round a lane descends through cells and then makes one attempt:
//   empty slot  -> CAS its body in;
//   body q      -> CAS the slot to -2 (lock), count the cells k that separate
//                  the two bodies;
//   locked slot -> back off, retry next round.
// After the attempt the workgroup re-converges and allocates all its new cells
// with a single atomicAdd (prefix sums of k over lanes and waves; one counter
// touched once per workgroup and round instead of once per split), each lock holder writes its chain
// with coherent stores, waits for them to complete (s_waitcnt) and publishes
// the chain by storing the first cell into the locked slot.
// The build runs in phases over strided subsets of the bodies (every 4096th,
// then every 64th, then the rest): the top levels of the tree form with little
// contention, later phases descend through them.  The topology does not depend
// on insertion order.
// The descent reads slots coherently: no pass clears the cell pool between steps (every
// cell's creator writes all its slots), so a cached copy of a line that holds a cell
// created during this launch could still carry the previous step's pointers.  This
// insertion serves only the bodies the LDS and bucketed builds leave (usually none) and
// the first phase of systems too large for the LDS build.

__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

template <int BLK>  // workgroup size: build_block(stride)
__global__ __launch_bounds__(BLK) void k_build(const SysDesc* __restrict__ sd, SysState* st,
                                               const BlkWork* __restrict__ work,
                                               const float4* __restrict__ node4, int* child, int* cparent,
                                               int* cpend, float4* cdata, int* slotoff, int stride, int exclude,
                                               const int* __restrict__ list) {
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    if (S->status) return;
    if (exclude == 0 && !list && S->top) return;  // the first phase was built in LDS (k_build_top)
    const int n = d.n;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    // The last phase (stride 1) takes its bodies from a per-system queue: a lane whose
    // insertion is complete claims the next body at the round's allocation, so that no
    // lane idles while others of its workgroup retry under contention.  Queue entry j
    // is the j-th body that the earlier phases left out (not a multiple of `exclude`),
    // or, with a list (the bucketed phases' fallback), the list's j-th body.
    const bool queued = stride == 1;
    const int nq = list ? min(max(ld_c(&S->nfb), 0), n)
                        : (exclude > 1 ? n - (n + exclude - 1) / exclude : (exclude == 1 ? 0 : n));
    const int* blist = list ? list + d.body_off : nullptr;
    if (queued && nq == 0) return;  // nothing queued (the usual case of the bucketed phases' fallback)
    int* C = child + 8 * (size_t)d.cell_off;
    int* par = cparent + d.cell_off;
    int* pend = cpend + d.cell_off;
    float4* CD = cdata + 8 * (size_t)d.cell_off;
    float rlx, rly, rlz, rhx, rhy, rhz;
    root_box(d.rules, S->box, rlx, rly, rlz, rhx, rhy, rhz);
    float px = 0.f, py = 0.f, pz = 0.f, pm = 0.f;
    float lx = rlx, ly = rly, lz = rlz, hx = rhx, hy = rhy, hz = rhz;
    int b = 0, node = 0, lvl = 0, dg = 0;
    int depth = 1;
    int tries = 0;
    bool active = false;
    auto start = [&](int body) {
        b = body;
        const float4 p = node4[d.node_off + b];
        px = p.x; py = p.y; pz = p.z; pm = p.w;
        lx = rlx; ly = rly; lz = rlz; hx = rhx; hy = rhy; hz = rhz;
        node = 0;
        lvl = 0;
        dg = level_digit(d.rules, 0, px, py, pz, lx, ly, lz, hx, hy, hz);
        tries = 0;
        active = true;
    };
    constexpr int kWaves = BLK / 64;
    __shared__ int s_wsum[kWaves];
    __shared__ int s_widle[kWaves];
    __shared__ int s_base, s_qbase;
    // lanes that want a body get consecutive queue entries (wave ranks, then wave offsets)
    auto queue_entry = [&](bool want) {
        const unsigned long long m = __ballot(want);
        int off = s_qbase + __popcll(m & ((1ull << lane) - 1));
        for (int q = 0; q < wave; ++q) off += s_widle[q];
        return off;
    };
    auto body_of = [&](int j) {
        return blist ? min(max(blist[j], 0), n - 1) : (exclude > 1 ? j + j / (exclude - 1) + 1 : j);
    };
    bool drained = false;  // thread 0: the queue has no entries left
    if (queued) {
        if (lane == 0) s_widle[wave] = 64;
        __syncthreads();
        if (threadIdx.x == 0) {
            s_qbase = atomicAdd(&S->bq, BLK);
            drained = s_qbase >= nq;
        }
        __syncthreads();
        const int j = queue_entry(true);
        if (j < nq) start(body_of(j));
    } else {
        const int bi = w.begin + threadIdx.x;  // index in this phase's strided subset
        const int bs = bi * stride;
        if (bi < w.end && bs < n && (exclude == 0 || bs % exclude != 0)) start(bs);
    }
    // block-synchronous rounds: one attempt per lane, then one allocation for the block
    while (__syncthreads_or(active)) {
        int k = 0, slot = 0, v = 0;
        float4 q4 = make_float4(0.f, 0.f, 0.f, 0.f);
        if (active) {
            if (++tries > (1 << 22)) {  // cannot happen on a consistent pool; never spin forever
                atomicOr(&S->status, kStStall);
                active = false;
            } else {
                slot = 8 * node + dg;
                v = ld_c(C + slot);
                while (v >= n) {
                    node = v - n;
                    ++lvl;
                    dg = level_digit(d.rules, lvl, px, py, pz, lx, ly, lz, hx, hy, hz);
                    slot = 8 * node + dg;
                    v = ld_c(C + slot);
                }
                if (v == kEmpty) {
                    if (cas_c(C + slot, kEmpty, b)) {
                        CD[slot] = make_float4(px, py, pz, pm);  // the body's entry of the cell's child block
                        depth = max(depth, lvl);
                        active = false;
                    }
                } else if (v >= 0) {
                    if (cas_c(C + slot, v, kLocked)) {
                        q4 = node4[d.node_off + v];
                        float ax = lx, ay = ly, az = lz, bx = hx, by = hy, bz = hz;
                        float qx0 = lx, qy0 = ly, qz0 = lz, qx1 = hx, qy1 = hy, qz1 = hz;
                        int l = lvl, db, dq;
                        bool deep = false;
                        do {
                            ++l;
                            if (l > kMaxLevel) { deep = true; break; }
                            db = level_digit(d.rules, l, px, py, pz, ax, ay, az, bx, by, bz);
                            dq = level_digit(d.rules, l, q4.x, q4.y, q4.z, qx0, qy0, qz0, qx1, qy1, qz1);
                            ++k;
                        } while (db == dq);
                        if (deep) {
                            atomicOr(&S->status, kStDepthOverflow);
                            st_c(C + slot, v);  // unlock, leave the old body in place
                            k = 0;
                            active = false;
                        }
                    }
                } else {
                    __builtin_amdgcn_s_sleep(1);
                    if (ld_c(&S->status)) active = false;
                }
            }
        }
        // one allocation for the whole block: wave prefix sums, then across the waves in
        // LDS; in the queued phase the same step claims the next bodies of the lanes that
        // finish this round
        const bool want = queued && (!active || k > 0);
        const int incl = wave_incl_scan(k);
        if (lane == 63) s_wsum[wave] = incl;
        const int nwant = __popcll(__ballot(want));  // every lane takes part in the ballot
        if (queued && lane == 0) s_widle[wave] = nwant;
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0, u = 0;
            for (int q = 0; q < kWaves; ++q) t += s_wsum[q];
            if (queued)
                for (int q = 0; q < kWaves; ++q) u += s_widle[q];
            // the cell allocation and the queue claim in one 64-bit atomic add on the
            // adjacent counters (ncells low, bq high; neither can carry into the other),
            // so the round pays one atomic round trip
            const bool claim = queued && u > 0 && !drained;
            int base = 0, qb = nq;
            if (t > 0 || claim) {
                const unsigned long long add =
                    (unsigned long long)(unsigned)t | ((unsigned long long)(unsigned)(claim ? u : 0) << 32);
                const unsigned long long old =
                    atomicAdd(reinterpret_cast<unsigned long long*>(&S->ncells), add);
                base = (int)(unsigned)old;
                if (claim) qb = (int)(unsigned)(old >> 32);
            }
            s_base = base;
            if (queued) {
                s_qbase = qb;  // without a claim (no lane wanted a body): nq, so no entry is taken
                // only a claim that ran can find the queue empty: a round in which every lane
                // was still busy claims nothing and must not stop the workgroup's later claims
                if (claim) drained = qb >= nq;
            }
        }
        __syncthreads();
        const int qj = queued ? queue_entry(want) : nq;
        if (k > 0) {
            int woff = 0;
            for (int q = 0; q < wave; ++q) woff += s_wsum[q];
            const int base = s_base + woff + incl - k;
            if (base + k > d.cap) {
                atomicOr(&S->status, kStTreeOverflow);
                st_c(C + slot, v);  // unlock, leave the old body in place
            } else {
                float ax = lx, ay = ly, az = lz, bx = hx, by = hy, bz = hz;
                float qx0 = lx, qy0 = ly, qz0 = lz, qx1 = hx, qy1 = hy, qz1 = hz;
                int parent = node, pslot = dg, lb = 0, lq = 0;
                for (int i = 0; i < k; ++i) {
                    const int c = base + i;
                    const int l2 = lvl + 1 + i;
                    const int db2 = level_digit(d.rules, l2, px, py, pz, ax, ay, az, bx, by, bz);
                    const int dq2 = level_digit(d.rules, l2, q4.x, q4.y, q4.z, qx0, qy0, qz0, qx1, qy1, qz1);
                    par[c] = (parent << 3) | pslot;  // parent cell and the slot holding c
                    st_c(pend + c, (i < k - 1) ? 1 : 0);
                    // per-slot counts of a new cell start at zero (a later bucketed phase counts there)
                    int4* so = reinterpret_cast<int4*>(slotoff + 8 * ((size_t)d.cell_off + c));
                    so[0] = make_int4(0, 0, 0, 0);
                    so[1] = make_int4(0, 0, 0, 0);
                    // every slot of the new cell (a cell's creator writes all eight; nothing clears them)
#pragma unroll
                    for (int k8 = 0; k8 < 8; ++k8) {
                        int sv = kEmpty;
                        if (k8 == db2) sv = i < k - 1 ? n + c + 1 : b;
                        else if (i == k - 1 && k8 == dq2) sv = v;
                        st_c(C + 8 * c + k8, sv);
                    }
                    if (i == k - 1) {
                        lb = 8 * c + db2;
                        lq = 8 * c + dq2;
                    }
                    parent = c;
                    pslot = db2;
                }
                atomicAdd(pend + node, 1);  // the slot now holds a cell instead of a body
                // chain stores are complete (coherent) before the chain becomes reachable
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
                __builtin_amdgcn_s_waitcnt(0);
                st_c(C + slot, n + base);
                // the two bodies' entries of the last cell's child block (read by summarize, a later kernel)
                CD[lb] = make_float4(px, py, pz, pm);
                CD[lq] = q4;
                depth = max(depth, lvl + k);
            }
            active = false;
        }
        if (want && qj < nq) start(body_of(qj));
    }
    // deepest level: block maximum, then an atomic only if it raises the stored value
    depth = wave_max_i(depth);
    if (lane == 0) s_wsum[wave] = depth;
    __syncthreads();
    if (threadIdx.x == 0) {
        int bm = 1;
        for (int q = 0; q < kWaves; ++q) bm = max(bm, s_wsum[q]);
        if (bm > ld_c(&S->max_depth)) atomicMax(&S->max_depth, bm);
    }
}

typedef int v4i __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// The first insertion phase (every stride-th body) in LDS: one workgroup per system, one
// lane per body of the phase, builds that phase's tree level by level, then writes the
// cells (slots, parent links, counts of child cells), the cell count and the depth, and
// sets S->top so that k_build skips the phase.  Per level, each body not yet placed claims
// the slot of its current cell with a CAS and a loser marks the slot contested; after a
// barrier an uncontested body is placed (its child-block entry goes to cdata), and the
// body holding a contested slot allocates a new cell there; after another barrier the
// contested bodies move into it.  So a cell exists exactly for every path prefix that two
// or more of the bodies share, which is the topology k_build's insertion gives.  Cell
// numbers are those of the LDS tree.  More than kTopBodies bodies writes nothing (S->top stays
// 0, and k_build runs the phase in global memory: the host launches it for such systems); more
// than kTopCells cells writes no cell and appends the phase's bodies to the build's fallback
// list (S->top = 2), which the queued insertion after the last phase takes: the tree's
// topology does not depend on the order of insertion.

constexpr int kTopCells = 1280;  // 50 KiB of LDS: 8 slots, a parent link and contested bits per cell
static_assert(kTopBodies == 512, "k_build_top: 8 KiB of LDS for the phase's bodies, one lane each");

__global__ __launch_bounds__(kTopBodies) void k_build_top(int nsys, const int* __restrict__ slots,
                                                          const SysDesc* __restrict__ sd, SysState* st,
                                                          const float4* __restrict__ node4, int* child, int* cparent,
                                                          int* cpend, float4* cdata, int* slotoff, int stride,
                                                          int* fblist) {
    __shared__ int sC[8 * kTopCells];
    __shared__ int sPar[kTopCells];
    __shared__ int sCont[kTopCells];  // bit k: slot k was claimed by more than one body this level
    __shared__ float4 sP[kTopBodies];
    __shared__ int sN, sFail, sDepth;
    if ((int)blockIdx.x >= nsys) return;
    const int sys = slots ? slots[blockIdx.x] : blockIdx.x;
    const SysDesc d = sd[sys];
    SysState* S = st + sys;
    const int n = d.n;
    const int nb = (n + stride - 1) / stride;  // bodies j * stride, j < nb
    // the build starts from the root alone (as k_tree_init leaves it; a step's fused bounds
    // and clear launch leaves the re-seeding to this kernel)
    if (threadIdx.x == 0) {
        S->ncells = 1;
        S->max_depth = 1;  // build_tree.cl:28
        S->bq = 0;
        S->top = 0;
        S->nfb = 0;
        if (d.algo == 0 && d.cap > 0) {  // an empty root (written over below when this kernel builds the top)
            int4* so = reinterpret_cast<int4*>(slotoff + 8 * (size_t)d.cell_off);
            so[0] = make_int4(0, 0, 0, 0);
            so[1] = make_int4(0, 0, 0, 0);
            int4* c4 = reinterpret_cast<int4*>(child + 8 * (size_t)d.cell_off);
            c4[0] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
            c4[1] = make_int4(kEmpty, kEmpty, kEmpty, kEmpty);
            cpend[d.cell_off] = 0;
        }
    }
    if (d.algo != 0 || n == 0 || nb > kTopBodies || S->status) return;
    for (int i = threadIdx.x; i < 8 * kTopCells; i += blockDim.x) sC[i] = kEmpty;
    for (int i = threadIdx.x; i < kTopCells; i += blockDim.x) sCont[i] = 0;
    const int j = threadIdx.x;
    if (j < nb) sP[j] = node4[d.node_off + j * stride];
    if (threadIdx.x == 0) {
        sN = 1;
        sFail = 0;
        sDepth = 1;
    }
    float4* CD = cdata + 8 * (size_t)d.cell_off;
    float lx, ly, lz, hx, hy, hz;
    root_box(d.rules, S->box, lx, ly, lz, hx, hy, hz);
    const int cap = min(kTopCells, d.cap);
    const int b = j * stride;
    bool active = j < nb;
    int node = 0, lvl = 0, depth = 1;
    __syncthreads();
    const float4 p = active ? sP[j] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (;;) {
        int slot = 0, dg = 0;
        if (active) {  // claim the slot of this level
            dg = level_digit(d.rules, lvl, p.x, p.y, p.z, lx, ly, lz, hx, hy, hz);
            slot = 8 * node + dg;
            if (atomicCAS(sC + slot, kEmpty, b) != kEmpty) atomicOr(sCont + node, 1 << dg);
        }
        __syncthreads();
        if (active) {
            if (!((sCont[node] >> dg) & 1)) {  // alone in its slot: placed
                CD[slot] = p;
                depth = max(depth, lvl);
                active = false;
            } else if (sC[slot] == b) {  // the holder of a contested slot opens a cell there
                if (lvl + 1 > kMaxLevel) {
                    atomicOr(&S->status, kStDepthOverflow);
                    sFail = 1;
                } else {
                    const int c = atomicAdd(&sN, 1);
                    if (c >= cap) {
                        sFail = 1;
                    } else {
                        sPar[c] = (node << 3) | dg;
                        sC[slot] = n + c;
                    }
                }
            }
        }
        __syncthreads();
        if (sFail) break;  // written only between the two barriers above: every lane reads the same
        if (active) {      // into the new cell, one level down
            node = sC[slot] - n;
            ++lvl;
        }
        if (!__syncthreads_or(active)) break;
    }
    depth = wave_max_i(depth);
    if ((threadIdx.x & 63) == 0) atomicMax(&sDepth, depth);
    __syncthreads();
    if (sFail) {  // uniform; with a depth overflow the status stops the step
        if (threadIdx.x == 0) {
            sN = atomicAdd(&S->nfb, nb);  // the list's room: n entries (each body once per step)
            S->top = 2;
        }
        __syncthreads();
        if (j < nb && sN + j < n) fblist[d.body_off + sN + j] = b;
        return;
    }
    const int nc = sN;
    int* C = child + 8 * (size_t)d.cell_off;
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        const int4 a = make_int4(sC[8 * c], sC[8 * c + 1], sC[8 * c + 2], sC[8 * c + 3]);
        const int4 e = make_int4(sC[8 * c + 4], sC[8 * c + 5], sC[8 * c + 6], sC[8 * c + 7]);
        reinterpret_cast<int4*>(C + 8 * c)[0] = a;
        reinterpret_cast<int4*>(C + 8 * c)[1] = e;
        cpend[d.cell_off + c] = (a.x >= n) + (a.y >= n) + (a.z >= n) + (a.w >= n) + (e.x >= n) + (e.y >= n) +
                                (e.z >= n) + (e.w >= n);
        if (c > 0) cparent[d.cell_off + c] = sPar[c];
        int4* so = reinterpret_cast<int4*>(slotoff + 8 * ((size_t)d.cell_off + c));  // counts of a later phase
        so[0] = make_int4(0, 0, 0, 0);
        so[1] = make_int4(0, 0, 0, 0);
    }
    if (threadIdx.x == 0) {
        S->ncells = nc;
        S->max_depth = sDepth;
        S->top = 1;
    }
}

// ---------------------------------------------------------------------------
// summarize: bottom-up mass, centre of mass and body count (saturate_tree.cl:6-146)
// without spin-waits.  Each cell ends with its children's (x, y, z, m) in one
// contiguous 128-byte block (cdata), so the force traversal fetches a cell's slots
// and all its children's data in a single round trip, and the inclusive per-slot
// body counts (slotoff) for the sort.  The build already wrote every body's entry of
// its cell's block; a finished cell writes its own (COM, mass) and body count into its
// parent's blocks, so a cell reads three contiguous blocks (slots, child data, counts)
// in one batch of loads instead of gathering its children.  A lane starts at a cell whose
// children are all bodies; after finishing a cell it decrements the parent's count of
// unfinished child cells, and the lane that brings it to zero goes on with the parent.
// Children are summed in slot order (deterministic; the reference's late children are
// summed LIFO, saturate_tree.cl:96-120): mass += m_k, com = fmaf(x_k, m_k, com), then
// com * (1/mass).

__global__ __launch_bounds__(256) void k_summarize(const SysDesc* __restrict__ sd, const SysState* __restrict__ st,
                                                   const BlkWork* __restrict__ work, float4* node4,
                                                   const int* __restrict__ child, const int* __restrict__ cparent,
                                                   int* cpend, int* ccount, int* slotoff, float4* cdata) {
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    const SysState* S = st + w.sys;
    if (S->status) return;
    const int used = min(S->ncells, d.cap);
    const int ci = w.begin + threadIdx.x;
    if (ci >= w.end || ci >= used) return;
    const int n = d.n;
    const int* C = child + 8 * (size_t)d.cell_off;
    // a cell of a subtree the frame build summarized: done; that subtree's root only hands its
    // count down to the cell above (its (COM, mass) and count are already in that cell's blocks)
    const int pe = cpend[d.cell_off + ci];
    if (pe == kSumDone) return;
    const bool sroot = pe == kSumRoot;
    if (!sroot) {
        const int4* c4 = reinterpret_cast<const int4*>(C + 8 * ci);
        int4 a = c4[0], bq = c4[1];
        int mx = max(max(max(a.x, a.y), max(a.z, a.w)), max(max(bq.x, bq.y), max(bq.z, bq.w)));
        if (mx >= n) return;  // has child cells: its last finished child will handle it
    }
    float4* CD = cdata + 8 * (size_t)d.cell_off;
    int* SO = slotoff + 8 * (size_t)d.cell_off;
    int c = ci;
    for (bool first = true;; first = false) {
        if (first && sroot) {
            const int p = cparent[d.cell_off + c] >> 3;
            if (__hip_atomic_fetch_add(cpend + d.cell_off + p, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 1)
                break;
            c = p;
            continue;
        }
        // The cell's slots and parent link are final since the build (ordinary cached loads).
        // A lane's first cell has only bodies below it, whose child-block entries the build
        // wrote: cached loads, and no counts (a body counts 1).  A cell it goes on with also
        // holds entries that child cells wrote in this kernel: its child block (128 B) and
        // per-slot counts (32 B) come in ten 16-byte loads, coherent at agent scope (sc1, as
        // the compiler's agent-scope atomic loads), with one wait inside the block, so no load
        // is in flight when it ends.
        int vs[8];
        float4 q[8];
        int sc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        int pk;
        {
            const int4* c4 = reinterpret_cast<const int4*>(C + 8 * c);
            const int4 s0 = c4[0], s1 = c4[1];
            pk = c != 0 ? cparent[d.cell_off + c] : 0;
            vs[0] = s0.x; vs[1] = s0.y; vs[2] = s0.z; vs[3] = s0.w;
            vs[4] = s1.x; vs[5] = s1.y; vs[6] = s1.z; vs[7] = s1.w;
            if (first) {
#pragma unroll
                for (int k = 0; k < 8; ++k) q[k] = CD[8 * c + k];
            } else {
                v4i o0, o1;
                v4f d0, d1, d2, d3, d4, d5, d6, d7;
                asm volatile(
                    "global_load_dwordx4 %0, %10, off sc1\n\t"
                    "global_load_dwordx4 %1, %10, off offset:16 sc1\n\t"
                    "global_load_dwordx4 %2, %10, off offset:32 sc1\n\t"
                    "global_load_dwordx4 %3, %10, off offset:48 sc1\n\t"
                    "global_load_dwordx4 %4, %10, off offset:64 sc1\n\t"
                    "global_load_dwordx4 %5, %10, off offset:80 sc1\n\t"
                    "global_load_dwordx4 %6, %10, off offset:96 sc1\n\t"
                    "global_load_dwordx4 %7, %10, off offset:112 sc1\n\t"
                    "global_load_dwordx4 %8, %11, off sc1\n\t"
                    "global_load_dwordx4 %9, %11, off offset:16 sc1\n\t"
                    "s_waitcnt vmcnt(0)"
                    : "=&v"(d0), "=&v"(d1), "=&v"(d2), "=&v"(d3), "=&v"(d4), "=&v"(d5), "=&v"(d6), "=&v"(d7),
                      "=&v"(o0), "=&v"(o1)
                    : "v"(CD + 8 * c), "v"(SO + 8 * c)
                    : "memory");
                const v4f dd[8] = {d0, d1, d2, d3, d4, d5, d6, d7};
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    sc[k] = o0[k];
                    sc[4 + k] = o1[k];
                }
#pragma unroll
                for (int k = 0; k < 8; ++k) q[k] = make_float4(dd[k][0], dd[k][1], dd[k][2], dd[k][3]);
            }
        }
        float m = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
        int cnt = 0;
        int incl[8];  // bodies in slots 0..k
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int v = vs[k];
            if (v >= 0) {
                const float wm = q[k].w;
                cnt += v < n ? 1 : sc[k];
                m = m + wm;
                cx = fmaf(q[k].x, wm, cx);
                cy = fmaf(q[k].y, wm, cy);
                cz = fmaf(q[k].z, wm, cz);
            }
            incl[k] = cnt;
        }
        const float inv = 1.0f / m;
        const float4 com = make_float4(cx * inv, cy * inv, cz * inv, m);
        int4* so = reinterpret_cast<int4*>(SO + 8 * c);
        bool more = false;
        int p = 0;
        if (c != 0) {
            p = pk >> 3;
            const int ps = pk & 7;
            // (COM, mass) in one 16-byte store and the count, both coherent (sc1) like the loads
            const v4f cv = {com.x, com.y, com.z, com.w};
            asm volatile("global_store_dwordx4 %0, %1, off sc1\n\t"
                         "global_store_dword %2, %3, off sc1"
                         :
                         : "v"(CD + 8 * p + ps), "v"(cv), "v"(SO + 8 * p + ps), "v"(cnt)
                         : "memory");
            // the coherent stores into the parent's blocks complete before its counter
            // drops, so the lane that takes the counter to zero reads them (sc1 loads)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
            __builtin_amdgcn_s_waitcnt(0);
            more = __hip_atomic_fetch_add(cpend + d.cell_off + p, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1;
        }
        // the cell's own outputs: read by later kernels only
        so[0] = make_int4(incl[0], incl[1], incl[2], incl[3]);
        so[1] = make_int4(incl[4], incl[5], incl[6], incl[7]);
        if (c == 0) {  // another cell's (COM, mass) and count are in its parent's blocks (k_cell_export)
            node4[d.node_off + n] = com;
            ccount[d.cell_off] = cnt;
        }
        if (!more) break;
        c = p;
    }
}

// ---------------------------------------------------------------------------
// sort: tree-order placement (Burtscher & Pingali's in-order placement, which
// the reference leaves empty).  A body's rank in depth-first slot order is the
// sum, along its root-to-leaf path, of the bodies held by earlier slots of each
// cell on the path; summarize stored per-slot counts.  One lane per body
// re-derives its path (the same octant digits as the build) and adds them up:
// no waiting, no dependence on the order in which cells were allocated.

// The same walk gives each body's place in the force order: at every cell, the
// bodies of the children that come earlier on the cell's Hilbert order precede
// it.  Both ranks come from the cell's 8 inclusive per-slot counts (one 32-byte
// read per level).
__global__ __launch_bounds__(256) void k_sort(const SysDesc* __restrict__ sd, SysState* st,
                                              const BlkWork* __restrict__ work, const float4* __restrict__ node4,
                                              const int* __restrict__ child, const int* __restrict__ slotoff,
                                              int* sorted, int* forder, const int* __restrict__ sorder) {
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    if (S->status) return;
    const int t = w.begin + threadIdx.x;
    if (t >= w.end) return;
    const int n = d.n;
    // lanes take the bodies in the last force pass's order, so that the lanes of a wave
    // walk nearly the same path (the ranks do not depend on which lane takes which body)
    const int b = S->order_ok ? min(max(sorder[d.body_off + t], 0), n - 1) : t;
    const int* C = child + 8 * (size_t)d.cell_off;
    const int* O = slotoff + 8 * (size_t)d.cell_off;
    const float4 p = node4[d.node_off + b];
    float lx, ly, lz, hx, hy, hz;
    root_box(d.rules, S->box, lx, ly, lz, hx, hy, hz);
    const bool hilbert = d.rules != 0;  // REFERENCE cells are not boxes in y: force order = tree order
    int node = 0, rank = 0, hrank = 0, hs = 0;
    for (int lvl = 0; lvl <= kMaxLevel; ++lvl) {
        const int dg = level_digit(d.rules, lvl, p.x, p.y, p.z, lx, ly, lz, hx, hy, hz);
        const int v = C[8 * node + dg];
        if (hilbert) {
            const int4 i0 = reinterpret_cast<const int4*>(O + 8 * node)[0];
            const int4 i1 = reinterpret_cast<const int4*>(O + 8 * node)[1];
            const int incl[8] = {i0.x, i0.y, i0.z, i0.w, i1.x, i1.y, i1.z, i1.w};
            const unsigned hpos = kHPos[hs];
            const int mine = (hpos >> (3 * dg)) & 7;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int cnt = incl[k] - (k ? incl[k - 1] : 0);
                rank += k < dg ? cnt : 0;
                hrank += (int)((hpos >> (3 * k)) & 7) < mine ? cnt : 0;
            }
            hs = (int)(kHNext[hs] >> (5 * dg)) & 31;
        } else {
            rank += dg ? O[8 * node + dg - 1] : 0;
            hrank = rank;
        }
        if (v == b) {
            if ((unsigned)rank < (unsigned)n) sorted[d.body_off + rank] = b;
            if ((unsigned)hrank < (unsigned)n) forder[d.body_off + hrank] = b;
            return;
        }
        if (v < n) break;  // not this body's slot: inconsistent tree
        node = v - n;
    }
    atomicOr(&S->status, kStStall);
}

// ---------------------------------------------------------------------------
// force (Barnes-Hut): one wave per 64 consecutive bodies of the force order, one
// shared traversal per wave.  A body child is always taken; a cell child at
// parent level l is taken for the whole wave iff every lane has
// eps^2+|d|^2 >= dq[l] (the wave-wide form of work_group_all,
// calculate_forces.cl:122), otherwise pushed.  The stack lives in LDS; node
// reads are wave-uniform (scalar loads), and so is every per-child decision,
// which the build keeps as plain scalar branches.  acc = G * sum d*m*rsqrt(d2)^3.

constexpr int kStack = 8 * kDqLen;
// Paired walk: each iteration pops the two top entries A (top) and B and fetches
// both cells with one scalar-load round trip, then pushes B's opened children and A's on
// top of them.  Entry levels stay non-decreasing from the bottom of the stack up (B's
// level <= A's, and what lay below B is no deeper than B), so a level-L entry is pushed
// only when a level-(L-1) cell is one of the two popped, i.e. when at most one entry of
// level >= L is on the stack: at most 1 + 16 entries of any level, 17 * 65 in all.
// The same cells are accepted as by the one-cell walk (acceptance depends only on the
// cell and the group), in a different order.
constexpr int kStackPair = 17 * kDqLen;
// Which systems walk in pairs is a property of the system alone (its rule set), so that
// a system's results do not depend on the batch it is in: REFERENCE-rule passes are bound
// by the scalar loads their waves have in flight and walk in pairs (1M: 2.69 -> 2.33 ms);
// STANDARD passes are bound by VALU issue and keep the one-cell walk, in a kernel instance
// of their own (compiled into one kernel with the paired walk, the register and LDS
// footprint of the pairs slowed them 1.50 -> 1.76 ms).  A launch of one instance skips the
// workgroups of the other rule set; a batch that holds both runs both instances.

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

// sorder (full passes only, else NULL): receives the force order, the order in
// which the next sort visits the bodies.  keys (full passes with longest-first dispatch,
// else NULL): per force workgroup, the largest interaction count of its waves, the
// dispatch-order key of the next pass (a workgroup covers nearly the same bodies from
// one step to the next).
template <bool KAHAN, bool PAIR>
__global__ __launch_bounds__(64 * kForceWaves) void k_force_bh(const SysDesc* __restrict__ sd, SysState* st,
                                                  const BlkWork* __restrict__ work,
                                                  const float4* __restrict__ node4, const int* __restrict__ child,
                                                  const float4* __restrict__ cdata, const int* __restrict__ sorted,
                                                  float4* __restrict__ acc4, int* __restrict__ sorder,
                                                  int* __restrict__ keys) {
    constexpr int kS = PAIR ? kStackPair : kStack;
    __shared__ int stk[kForceWaves][kS];
    // the thresholds by level, over every value of a stack entry's 7-bit level field (the
    // levels past kMaxLevel, which a built tree does not have, repeat the last threshold)
    __shared__ float dqs[128];
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    const SysState* S = st + w.sys;
    if (S->status || (d.rules == 0) != PAIR) return;  // the other instance's system
    for (int i = threadIdx.x; i < 128; i += blockDim.x) dqs[i] = S->dq[min(i, kMaxLevel)];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int first = w.begin + wave * 64;
    if (first >= w.end) return;
    const int idx = first + lane;
    const bool valid = idx < w.end;
    const int b = min(max(sorted[d.body_off + (valid ? idx : w.end - 1)], 0), d.n - 1);
    float4 p = node4[d.node_off + b];
    // The empty asm consumes the position once, before the walk, so that the load's wait
    // happens there; else the compiler waits again in every child test of every cell.
    asm volatile("" : "+v"(p.x), "+v"(p.y), "+v"(p.z));
    const int* __restrict__ C = child + 8 * (size_t)d.cell_off;
    const float4* __restrict__ CD = cdata + 8 * (size_t)d.cell_off;
    const int n = d.n;
    const float e2 = d.eps * d.eps;
    float ax = 0.f, ay = 0.f, az = 0.f, kx = 0.f, ky = 0.f, kz = 0.f;
    int nint = 0;
    int* sk = stk[wave];
    // byte offsets of a stack entry's cell in child (8 ints a cell) and cdata (8 float4)
    auto data_off = [](int e) { return (unsigned)e & ~127u; };
    auto slot_off = [&](int e) { return data_off(e) >> 2; };
    int sp = 0;
    sk[0] = 0;
    sp = 1;
    while (sp > 0) {
        --sp;
        const int cur = __builtin_amdgcn_readfirstlane(sk[sp]);  // (cell << 7) | level
        // A popped cell's 8 slots (32 bytes) and its children's (x,y,z,m) (one 128-byte
        // block) come in one batch of wave-uniform scalar loads with a single wait, so
        // that a cell (or a pair of cells) costs one scalar-memory round trip.  Left to
        // itself the compiler sinks each child's load into the branch that uses it and
        // waits up to three times per cell, which lengthens the walk of the longest
        // waves.  The wait is inside the asm block: no load is in flight when it ends.
        // The loads address the system's two arrays by a fixed 64-bit base plus a 32-bit
        // byte offset in a scalar register (the cell pool is below 2^24 cells, so the
        // offsets stay below 2^31), which keeps 64-bit address arithmetic and its
        // registers out of the walk.
        // Cell 0 of the batch is visited first: with a pair, the second-popped cell.
        // The cells' opening thresholds are read from LDS before the batch; the empty asm
        // after it consumes them, so that their wait (already met by the batch's) comes
        // once, not before every child test.
        v8i sv[2];
        v16f q0[2], q1[2];
        float dqv[2];
        int lv[2];
        int cells = 1;
        if (PAIR && sp > 0) {
            --sp;
            const int cur2 = __builtin_amdgcn_readfirstlane(sk[sp]);
            lv[0] = cur2 & 127;
            lv[1] = cur & 127;
            dqv[0] = dqs[lv[0]];
            dqv[1] = dqs[lv[1]];
            cells = 2;
            asm volatile(
                "s_load_dwordx8 %[s0], %[cb], %[co0]\n\t"
                "s_load_dwordx16 %[a0], %[db], %[do0]\n\t"
                "s_load_dwordx16 %[b0], %[db], %[do0] offset:0x40\n\t"
                "s_load_dwordx8 %[s1], %[cb], %[co1]\n\t"
                "s_load_dwordx16 %[a1], %[db], %[do1]\n\t"
                "s_load_dwordx16 %[b1], %[db], %[do1] offset:0x40\n\t"
                "s_waitcnt lgkmcnt(0)"
                : [s0] "=&s"(sv[0]), [a0] "=&s"(q0[0]), [b0] "=&s"(q1[0]), [s1] "=&s"(sv[1]),
                  [a1] "=&s"(q0[1]), [b1] "=&s"(q1[1])
                : [cb] "s"(C), [db] "s"(CD), [co0] "s"(slot_off(cur2)), [do0] "s"(data_off(cur2)),
                  [co1] "s"(slot_off(cur)), [do1] "s"(data_off(cur))
                : "memory");
            asm volatile("" : "+v"(dqv[0]), "+v"(dqv[1]));
        } else {
            lv[0] = cur & 127;
            dqv[0] = dqs[lv[0]];
            asm volatile(
                "s_load_dwordx8 %[s0], %[cb], %[co0]\n\t"
                "s_load_dwordx16 %[a0], %[db], %[do0]\n\t"
                "s_load_dwordx16 %[b0], %[db], %[do0] offset:0x40\n\t"
                "s_waitcnt lgkmcnt(0)"
                : [s0] "=&s"(sv[0]), [a0] "=&s"(q0[0]), [b0] "=&s"(q1[0])
                : [cb] "s"(C), [db] "s"(CD), [co0] "s"(slot_off(cur)), [do0] "s"(data_off(cur))
                : "memory");
            asm volatile("" : "+v"(dqv[0]));
        }
#pragma unroll
        for (int c = 0; c < (PAIR ? 2 : 1); ++c) {
            if (c >= cells) break;
            // a pushed child's entry, ((v - n) << 7) | (level + 1), as (v << 7) + pb (mod 2^32)
            const unsigned pb = (unsigned)(lv[c] + 1) - ((unsigned)n << 7);
            const float dql = dqv[c];
            const int vs[8] = {sv[c][0], sv[c][1], sv[c][2], sv[c][3], sv[c][4], sv[c][5], sv[c][6], sv[c][7]};
            float4 qs[8];
            qs[0] = make_float4(q0[c][0], q0[c][1], q0[c][2], q0[c][3]);
            qs[1] = make_float4(q0[c][4], q0[c][5], q0[c][6], q0[c][7]);
            qs[2] = make_float4(q0[c][8], q0[c][9], q0[c][10], q0[c][11]);
            qs[3] = make_float4(q0[c][12], q0[c][13], q0[c][14], q0[c][15]);
            qs[4] = make_float4(q1[c][0], q1[c][1], q1[c][2], q1[c][3]);
            qs[5] = make_float4(q1[c][4], q1[c][5], q1[c][6], q1[c][7]);
            qs[6] = make_float4(q1[c][8], q1[c][9], q1[c][10], q1[c][11]);
            qs[7] = make_float4(q1[c][12], q1[c][13], q1[c][14], q1[c][15]);
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int v = vs[k];
                if (v < 0) continue;
                const float dx = qs[k].x - p.x, dy = qs[k].y - p.y, dz = qs[k].z - p.z;
                const float d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, e2)));
                // the vote: every lane far enough, written as "no lane near" (the same for
                // the finite d2 of a pass over finite positions), whose compare result
                // the branch tests directly, without a scalar compare against the exec mask
                bool take = v < n;
                if (!take) take = !__any(d2 < dql);
                if (take) {
                    const float ri = __builtin_amdgcn_rsqf(d2);
                    const float F = qs[k].w * (ri * ri * ri);
                    if (KAHAN) {
                        float y, t;
                        y = dx * F - kx; t = ax + y; kx = (t - ax) - y; ax = t;
                        y = dy * F - ky; t = ay + y; ky = (t - ay) - y; ay = t;
                        y = dz * F - kz; t = az + y; kz = (t - az) - y; az = t;
                    } else {
                        ax = fmaf(dx, F, ax);
                        ay = fmaf(dy, F, ay);
                        az = fmaf(dz, F, az);
                    }
                    ++nint;
                } else if (sp < kS) {  // the bounds above (one-cell walk: 7 * 64 + 8) keep sp below kS
                    sk[sp] = (int)(((unsigned)v << 7) + pb);
                    ++sp;
                }
            }
        }
    }
    // a target-range launch votes with whole groups but writes only its range
    if (valid && idx >= w.begin + (w.trim & 0xffff) && idx < w.end - (w.trim >> 16)) {
        if (KAHAN) { ax -= kx; ay -= ky; az -= kz; }
        acc4[d.body_off + b] = make_float4(ax * d.G, ay * d.G, az * d.G, (float)nint);
    }
    if (sorder && valid) {  // a full pass covers every body of the system once
        sorder[d.body_off + idx] = b;
        if (idx == 0) st[w.sys].order_ok = 1;
    }
    if (keys && lane == 0) atomicMax(keys + d.grp_first + w.begin / (kForceVote * kForceWaves), min(nint, 1 << 24));
}

// Dispatch order of the force workgroups, longest first (LPT list scheduling),
// for launches of at most two waves per SIMD slot, where the last workgroups to
// start decide when the kernel ends.  keys[i] (written by the previous force pass)
// estimates workgroup i's time; workgroups are binned by it (64 bins up to the largest) and
// written out largest bin first, and the keys are cleared for the next pass.  The
// order decides only when a workgroup starts, never what it computes.  One
// workgroup of 1024 threads; the keys stay in LDS.
constexpr int kOrderBins = 64;

__global__ __launch_bounds__(1024) void k_force_order(const BlkWork* __restrict__ in, int len, int* keys,
                                                      BlkWork* __restrict__ out) {
    __shared__ int key[kOrderMax];
    __shared__ int hist[kOrderBins];
    __shared__ int wmax[16];
    int m = 0;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const int k = min(max(keys[i], 0), 1 << 24);
        keys[i] = 0;
        key[i] = k;
        m = max(m, k);
    }
    m = wave_max_i(m);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    if (threadIdx.x < kOrderBins) hist[threadIdx.x] = 0;
    __syncthreads();
    int top = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) top = max(top, wmax[q]);
    const long long scale = (long long)top + 1;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        key[i] = kOrderBins - 1 - (int)((long long)key[i] * kOrderBins / scale);
        atomicAdd(&hist[key[i]], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int b = 0; b < kOrderBins; ++b) {
            const int c = hist[b];
            hist[b] = sum;
            sum += c;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < len; i += blockDim.x) out[atomicAdd(&hist[key[i]], 1)] = in[i];
}

// The same order computed by one 256-lane workgroup with the keys in registers (at most
// kOrderSmall workgroups), so that it can run as one workgroup of a step's first launch
// (k_bounds_clear), beside that launch's other workgroups.
__device__ __forceinline__ void force_order_small(const BlkWork* __restrict__ in, int len, int* keys,
                                                  BlkWork* __restrict__ out) {
    constexpr int K = kOrderSmall / 256;
    __shared__ int hist[kOrderBins];
    __shared__ int wmax[4];
    int kk[K];
    int m = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int i = (int)threadIdx.x + 256 * j;
        int k = 0;
        if (i < len) {
            k = min(max(keys[i], 0), 1 << 24);
            keys[i] = 0;
        }
        kk[j] = k;
        m = max(m, k);
    }
    m = wave_max_i(m);
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    if (threadIdx.x < kOrderBins) hist[threadIdx.x] = 0;
    __syncthreads();
    const long long scale = (long long)max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])) + 1;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        kk[j] = kOrderBins - 1 - (int)((long long)kk[j] * kOrderBins / scale);
        if ((int)threadIdx.x + 256 * j < len) atomicAdd(&hist[kk[j]], 1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int sum = 0;
        for (int b = 0; b < kOrderBins; ++b) {
            const int c = hist[b];
            hist[b] = sum;
            sum += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < K; ++j) {
        const int i = (int)threadIdx.x + 256 * j;
        if (i < len) out[atomicAdd(&hist[kk[j]], 1)] = in[i];
    }
}

// A step's bounds pass and the force pass's dispatch order in one launch: with an order
// table (olen > 0), workgroup 0 computes the longest-first order (force_order_small) beside
// the bounds workgroups (whose ticket counts only them).  No pass clears the tree: every
// cell's creator writes all its slots, and the root's re-seeding (k_tree_init's work) is done
// by k_build_top, the build's first launch.
__global__ __launch_bounds__(256) void k_bounds_order(const SysDesc* __restrict__ sd, SysState* st,
                                                      const BlkWork* __restrict__ wb,
                                                      const float4* __restrict__ node4, float* partial,
                                                      const BlkWork* __restrict__ og, int olen, int* keys,
                                                      BlkWork* __restrict__ oout) {
    int blk = blockIdx.x;
    if (olen > 0) {
        if (blk == 0) {
            force_order_small(og, olen, keys, oout);
            return;
        }
        --blk;
    }
    bounds_block(sd, st, wb[blk], blk, node4, partial);
}

// The per-cell (COM, mass) and body count of one system's tree for the exports, from the
// blocks the step writes: a cell's (COM, mass) is its entry in its parent's child block and its
// count its last inclusive per-slot count (the root's were written by k_summarize).  One lane
// per cell hands its child cells their entries (no parent links needed).
__global__ __launch_bounds__(256) void k_cell_export(int n, int ncells, const int* __restrict__ child,
                                                     const float4* __restrict__ cdata,
                                                     const int* __restrict__ slotoff, float4* cells, int* ccount) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int v = child[8 * (size_t)c + k] - n;
        if (v > 0 && v < ncells) {
            cells[v] = cdata[8 * (size_t)c + k];
            ccount[v] = slotoff[8 * (size_t)v + 7];
        }
    }
}

// Block table of one target range [lo, hi) of system `sys`: blocks of `per` items
// from `start` (<= lo) to `stop` (>= hi); trim marks the items outside [lo, hi).
__global__ __launch_bounds__(256) void k_range_table(BlkWork* __restrict__ out, int nblk, int sys, int start,
                                                     int stop, int lo, int hi, int per) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblk) return;
    const int b = start + i * per, e = min(stop, b + per);
    const int head = min(max(lo - b, 0), e - b), tail = min(max(e - hi, 0), e - b);
    out[i] = BlkWork{sys, b, e, head | (tail << 16)};
}

// ---------------------------------------------------------------------------
// integrator arithmetic of one body (calculate_forces.cl:150-182), shared by
// k_update and the brute-force kernel's fused form: a kick v += a*h, a drift
// x += v*dt (w, the mass, is kept).
__device__ __forceinline__ float4 kick(float4 v, float ax, float ay, float az, float h) {
    v.x = fmaf(ax, h, v.x);
    v.y = fmaf(ay, h, v.y);
    v.z = fmaf(az, h, v.z);
    return v;
}
__device__ __forceinline__ float4 drift(float4 x, float4 v, float dt) {
    x.x = fmaf(v.x, dt, x.x);
    x.y = fmaf(v.y, dt, x.y);
    x.z = fmaf(v.z, dt, x.z);
    return x;
}

// ---------------------------------------------------------------------------
// brute force: all pairs (brute_force_update.cl:3-105 `update`, :108-228 `update_local`).
// LAW 0: d*m*rsqrt(|d|^2+eps^2)^3.  LAW 1 (brute_force_update.cl:54-58,114-120):
// n = normalize(d), n*m/(|n|^2+ueps^2).  A normalised n has |n|^2 = 1 (the
// reference's recomputed |n|^2 differs from 1 by rounding only), so LAW 1 is
// evaluated as d * rsqrt(|d|^2) * (m/(1+ueps^2)), with m/(1+ueps^2) applied
// once per source when it is staged.
//
// Workgroup = 16 waves.  A wave holds 64*TB/W targets: W lanes per target
// (W = 2^bf_lanes_log2) and TB targets per lane (register blocking).  S =
// 2^bf_split_log2 waves share one target group; wave `part` sums the sources
// [n*part/S, n*(part+1)/S), staged 64 at a time through its own LDS tile, lane
// part p taking tile entries [64p/W, 64(p+1)/W) in order and reading them as
// broadcasts.  The W parts are added in part order across lanes, then the S
// waves' sums in wave order through LDS.  W and S depend on N only (small N
// splits the sources finer so that every CU gets work), and TB does not change
// any target's arithmetic, so a system's result does not depend on its batch.
// The self pair is not skipped: its displacement is exactly zero, so it adds
// exactly zero (eps > 0 keeps the PLUMMER denominator finite; LAW 1 maps a
// zero displacement to a zero unit vector).

template <int LAW>
__device__ __forceinline__ void bf_pair(const float4 q, float px, float py, float pz, float e2, float& fx, float& fy,
                                        float& fz, float& s) {
    const float dx = q.x - px, dy = q.y - py, dz = q.z - pz;
    if (LAW == 0) {
        const float d2 = fmaf(dz, dz, fmaf(dy, dy, fmaf(dx, dx, e2)));
        const float ri = __builtin_amdgcn_rsqf(d2);
        const float ri2 = ri * ri;
        const float ri3 = ri2 * ri;
        s = q.w * ri3;
        fx = dx; fy = dy; fz = dz;
    } else {
        // q.w = m/(1+ueps^2).  A zero displacement (the self pair) is clamped to the
        // smallest normal |d|^2: rsqrt stays finite and d * rsqrt = 0
        const float l2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
        s = q.w * __builtin_amdgcn_rsqf(fmaxf(l2, 1.17549435e-38f));
        fx = dx; fy = dy; fz = dz;
    }
}

__device__ __forceinline__ void kahan_add(float& a, float& c, float v) {
    const float y = v - c;
    const float t = a + y;
    c = (t - a) - y;
    a = t;
}

// One workgroup's share of a system: W lanes per target (parts of the wave's
// sources), TB targets per lane, S = 2^lg waves per target group.
template <int LAW, bool KAHAN, int TB, int W>
__device__ __forceinline__ void brute_body(const BlkWork& w, const SysDesc& d, const float4* __restrict__ node4,
                                           float4* __restrict__ acc4, float4 (*tile)[kBfTileStride],
                                           float (*red)[2 * 6][64], float4* __restrict__ vel4,
                                           float4* __restrict__ out4, int mode) {
    constexpr int NC = KAHAN ? 6 : 3;
    constexpr int L = 64 / W;  // lanes (targets) per part, sources per part and tile
    // the wave index is uniform over the wave: keep it (and the loop bounds) scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int pw = lane / L, tl = lane % L;
    const int lg = d.bf_split_log2;
    const int S = 1 << lg;
    const int grp = wave >> lg, part = wave & (S - 1);
    const int n = d.n;
    const float4* __restrict__ nodes = node4 + d.node_off;
    const float e = LAW == 0 ? d.eps : d.ueps;
    const float e2 = e * e;
    const float cu = 1.0f / (1.0f + e2);  // LAW 1: 1/(|n|^2 + ueps^2) with |n|^2 = 1
    const int gbase = w.begin + grp * L * TB;
    float px[TB], py[TB], pz[TB], pm[TB], acc[NC][TB];
#pragma unroll
    for (int t = 0; t < TB; ++t) {
        const int i = min(gbase + t * L + tl, w.end - 1);
        const float4 p = nodes[i];
        px[t] = p.x; py[t] = p.y; pz[t] = p.z; pm[t] = p.w;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[c][t] = 0.0f;
    }
    // part pw of a 64-source tile sits at row pw (rows padded by one entry: no bank conflicts)
    float4* tw = tile[wave];
    const int store_at = (lane / L) * (L + 1) + lane % L;
    const float4* mine = tw + pw * (L + 1);
    const int j0 = (int)((long long)n * part >> lg), j1 = (int)((long long)n * (part + 1) >> lg);
    for (int jt = j0; jt < j1; jt += 64) {
        const int cnt = min(64, j1 - jt);
        wave_sync();
        if (lane < cnt) {
            float4 q = nodes[jt + lane];
            if (LAW == 1) q.w *= cu;
            tw[store_at] = q;
        }
        wave_sync();
        const int my = min(max(cnt - pw * L, 0), L);
#pragma unroll 4
        for (int k = 0; k < my; ++k) {
            const float4 q = mine[k];
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                float fx, fy, fz, sc;
                bf_pair<LAW>(q, px[t], py[t], pz[t], e2, fx, fy, fz, sc);
                if (KAHAN) {
                    kahan_add(acc[0][t], acc[3][t], fx * sc);
                    kahan_add(acc[1][t], acc[4][t], fy * sc);
                    kahan_add(acc[2][t], acc[5][t], fz * sc);
                } else {
                    acc[0][t] = fmaf(fx, sc, acc[0][t]);
                    acc[1][t] = fmaf(fy, sc, acc[1][t]);
                    acc[2][t] = fmaf(fz, sc, acc[2][t]);
                }
            }
        }
    }
    if (W > 1) {  // parts of the wave, added in part order into the part-0 lanes
#pragma unroll
        for (int t = 0; t < TB; ++t) {
            float v[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) v[c] = KAHAN ? acc[c][t] - acc[c + 3][t] : acc[c][t];
            float o[W][3];
#pragma unroll
            for (int q = 1; q < W; ++q)
#pragma unroll
                for (int c = 0; c < 3; ++c) o[q][c] = __shfl(v[c], tl + q * L, 64);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                if (KAHAN) {
                    acc[c][t] = v[c];
                    acc[c + 3][t] = 0.0f;
#pragma unroll
                    for (int q = 1; q < W; ++q) kahan_add(acc[c][t], acc[c + 3][t], o[q][c]);
                } else {
#pragma unroll
                    for (int q = 1; q < W; ++q) acc[c][t] += o[q][c];
                }
            }
        }
    }
    if (S > 1) {  // uniform over the workgroup
        if (part != 0) {
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int t = 0; t < TB; ++t) red[wave][c * TB + t][lane] = acc[c][t];
        }
        __syncthreads();
        if (part != 0) return;
        for (int q = 1; q < S; ++q) {
#pragma unroll
            for (int t = 0; t < TB; ++t) {
                if (KAHAN) {
#pragma unroll
                    for (int c = 0; c < 3; ++c)
                        kahan_add(acc[c][t], acc[c + 3][t],
                                  red[wave + q][c * TB + t][lane] - red[wave + q][(c + 3) * TB + t][lane]);
                } else {
#pragma unroll
                    for (int c = 0; c < 3; ++c) acc[c][t] += red[wave + q][c * TB + t][lane];
                }
            }
        }
    }
    if (pw != 0) return;
#pragma unroll
    for (int t = 0; t < TB; ++t) {
        const int i = gbase + t * L + tl;
        if (i < w.end) {
            float ax = acc[0][t], ay = acc[1][t], az = acc[2][t];
            if (KAHAN) { ax -= acc[3][t]; ay -= acc[4][t]; az -= acc[5][t]; }
            const float4 a = make_float4(ax * d.G, ay * d.G, az * d.G, (float)(n - 1));
            acc4[d.body_off + i] = a;
            if (mode >= 0) {  // fused integrator, k_update's arithmetic
                const float4 v = kick(vel4[d.body_off + i], a.x, a.y, a.z, mode == 0 ? d.dt : 0.5f * d.dt);
                vel4[d.body_off + i] = v;
                if (mode == 0) out4[d.node_off + i] = drift(make_float4(px[t], py[t], pz[t], pm[t]), v, d.dt);
            }
        }
    }
}

// mode >= 0: the integrator of that mode (k_update's) follows in the lane that
// writes each target's acceleration.  Velocities are read by no other lane, so
// they are updated in place; mode 0 also moves the positions, which every
// workgroup still reads as sources, so the moved positions (with the masses) go
// to out4, a second position buffer, and the caller alternates the two.
template <int LAW, bool KAHAN, int TB>
__global__ __launch_bounds__(1024) void k_brute(const SysDesc* __restrict__ sd, const SysState* __restrict__ st,
                                                const BlkWork* __restrict__ work, const float4* __restrict__ node4,
                                                float4* __restrict__ acc4, float4* __restrict__ vel4,
                                                float4* __restrict__ out4, int mode) {
    __shared__ float4 tile[kBfWaves][kBfTileStride];
    __shared__ float red[kBfWaves][2 * 6][64];
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    if (st[w.sys].status) return;  // whole workgroup
    switch (d.bf_lanes_log2) {     // uniform over the workgroup
    case 0: brute_body<LAW, KAHAN, TB, 1>(w, d, node4, acc4, tile, red, vel4, out4, mode); break;
    case 1: brute_body<LAW, KAHAN, TB, 2>(w, d, node4, acc4, tile, red, vel4, out4, mode); break;
    default: brute_body<LAW, KAHAN, TB, 4>(w, d, node4, acc4, tile, red, vel4, out4, mode); break;
    }
}

// ---------------------------------------------------------------------------
// collision (update_collision_local, collision_update.cl:17-83, with
// calc_collision_traj :3-15): every body is a sphere of radius 0.015; for each
// other body j in index order, if |p_j - p_i| <= r_i + r_j the velocity of i is
// reflected about the unit separation n: v -= 2 (n.v) n, using the velocity as
// already reflected by earlier j.  Positions are read-only, each lane writes
// only its own velocity.  The reference tile loop re-reads tile 0 and skips the
// wrong partner (the brute-force tile-loop quirk); this is the intended
// all-pairs sweep.  A coincident pair has no separation direction and leaves v
// unchanged.  Sources are staged 64 at a time through a per-wave LDS tile.

constexpr float kCollisionRadius = 0.015f;

__global__ __launch_bounds__(256) void k_collision(const SysDesc* __restrict__ sd, const SysState* __restrict__ st,
                                                   const BlkWork* __restrict__ work,
                                                   const float4* __restrict__ node4, float4* __restrict__ vel4,
                                                   int all) {
    __shared__ float4 tile[4][64];
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    if (st[w.sys].status || !(all || d.collision)) return;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int i = min(w.begin + (int)threadIdx.x, w.end - 1);
    const float4* __restrict__ nodes = node4 + d.node_off;
    const float4 p = nodes[i];
    float4 v = vel4[d.body_off + i];
    const float rsum = kCollisionRadius + kCollisionRadius;
    // l2 above this bound has sqrt(l2) > rsum for certain: the exact test is skipped
    const float l2_skip = rsum * rsum * 1.0001f;
    const int n = d.n;
    for (int jt = 0; jt < n; jt += 64) {
        const int cnt = min(64, n - jt);
        wave_sync();
        if (lane < cnt) tile[wave][lane] = nodes[jt + lane];
        wave_sync();
        for (int k = 0; k < cnt; ++k) {
            const float4 q = tile[wave][k];
            const float dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const float l2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            if (l2 > l2_skip || jt + k == i || l2 == 0.0f) continue;
            const float dist = sqrtf(l2);
            if (dist > rsum) continue;
            const float nx = dx / dist, ny = dy / dist, nz = dz / dist;
            const float k2 = 2.0f * fmaf(nz, v.z, fmaf(ny, v.y, nx * v.x));
            v.x = fmaf(-k2, nx, v.x);
            v.y = fmaf(-k2, ny, v.y);
            v.z = fmaf(-k2, nz, v.z);
        }
    }
    if (w.begin + (int)threadIdx.x < w.end) vel4[d.body_off + i] = v;
}

// ---------------------------------------------------------------------------
// vertex export (copy_vertices, copy_vertices.cl:3-17): (x, y, z, 1) per body.

// Copies n float4: the fused brute-force step's copy of the moved positions back into
// node4 after an odd number of steps.  A kernel rather than a hipMemcpyAsync, so that every
// node of a captured step graph is a kernel node.
__global__ __launch_bounds__(256) void k_copy4(int n, const float4* __restrict__ src, float4* __restrict__ dst) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_vertices(int n, const float4* __restrict__ node4, float4* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float4 p = node4[i];
    out[i] = make_float4(p.x, p.y, p.z, 1.0f);
}

// ---------------------------------------------------------------------------
// FP64 conservation diagnostics of one system (no counterpart in the reference).
// Per body i: kinetic m v^2/2, momentum m v, angular momentum m x*v, mass, and
// half its pair potential sum_j!=i phi(r_ij) (PLUMMER: -G m_i m_j/sqrt(r^2+eps^2);
// REFERENCE_UNIT: +G m_i m_j r/(1+eps^2)), all in FP64.  Sources are staged
// through LDS 256 at a time.  Workgroup sums go to `partial` in a fixed tree
// order; k_energy_final adds them in block order, so the result is deterministic.

__device__ __forceinline__ void block_sum_f64(double* v, double (*sh)[256], int nv) {
    const int t = threadIdx.x;
    for (int c = 0; c < nv; ++c) sh[c][t] = v[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int c = 0; c < nv; ++c) sh[c][t] += sh[c][t + w];
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_energy_f64(int n, const float4* __restrict__ node4,
                                                    const float4* __restrict__ vel4, int law, double eps,
                                                    double G, double* __restrict__ partial) {
    __shared__ float4 tile[256];
    __shared__ double sh[kEnergyLen][256];
    const int i = blockIdx.x * 256 + threadIdx.x;
    const bool valid = i < n;
    const float4 p = node4[valid ? i : 0];
    const double px = p.x, py = p.y, pz = p.z;
    const double e2 = eps * eps, inv_unit = 1.0 / (1.0 + e2);
    double pot = 0.0;
    for (int jt = 0; jt < n; jt += 256) {
        __syncthreads();
        if (jt + threadIdx.x < n) tile[threadIdx.x] = node4[jt + threadIdx.x];
        __syncthreads();
        const int cnt = min(256, n - jt);
        for (int k = 0; k < cnt; ++k) {
            const float4 q = tile[k];
            const double dx = (double)q.x - px, dy = (double)q.y - py, dz = (double)q.z - pz;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double phi = law == 0 ? -1.0 / sqrt(r2 + e2) : sqrt(r2) * inv_unit;
            pot += (jt + k == i) ? 0.0 : (double)q.w * phi;
        }
    }
    double v[kEnergyLen] = {};
    if (valid) {
        const float4 u = vel4[i];
        const double m = p.w, vx = u.x, vy = u.y, vz = u.z;
        v[0] = 0.5 * m * (vx * vx + vy * vy + vz * vz);
        v[1] = 0.5 * G * m * pot;
        v[2] = m * vx; v[3] = m * vy; v[4] = m * vz;
        v[5] = m * (py * vz - pz * vy); v[6] = m * (pz * vx - px * vz); v[7] = m * (px * vy - py * vx);
        v[8] = m;
    }
    block_sum_f64(v, sh, kEnergyLen);
    if (threadIdx.x < kEnergyLen) partial[(size_t)blockIdx.x * kEnergyLen + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_energy_final(int nblk, const double* __restrict__ partial,
                                                      double* __restrict__ out) {
    __shared__ double sh[kEnergyLen][256];
    double v[kEnergyLen] = {};
    for (int b = threadIdx.x; b < nblk; b += 256)
        for (int c = 0; c < kEnergyLen; ++c) v[c] += partial[(size_t)b * kEnergyLen + c];
    block_sum_f64(v, sh, kEnergyLen);
    if (threadIdx.x < kEnergyLen) out[threadIdx.x] = sh[threadIdx.x][0];
}

// ---------------------------------------------------------------------------
// update (calculate_forces.cl:169-182 EULER; kick-drift-kick for LEAPFROG).
// mode 0: v += a*dt, x += v*dt.  mode 1: v += a*dt/2.  mode 2: v += a*dt/2, x += v*dt.

__global__ __launch_bounds__(256) void k_update(const SysDesc* __restrict__ sd, const SysState* __restrict__ st,
                                                const BlkWork* __restrict__ work, float4* node4, float4* vel4,
                                                const float4* __restrict__ acc4, int mode) {
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    if (st[w.sys].status) return;
    const int i = w.begin + threadIdx.x;
    if (i >= w.end) return;
    const float4 a = acc4[d.body_off + i];
    const float4 v = kick(vel4[d.body_off + i], a.x, a.y, a.z, mode == 0 ? d.dt : 0.5f * d.dt);
    vel4[d.body_off + i] = v;
    if (mode != 1) node4[d.node_off + i] = drift(node4[d.node_off + i], v, d.dt);
}

// ---------------------------------------------------------------------------
// host-side launchers

#define NB_LAUNCH(kern, grid, block, stream, ...)                                  \
    do {                                                                           \
        if ((grid) > 0) hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, stream, __VA_ARGS__); \
    } while (0)

void launch_bounds(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                   const float4* node4, float* partial) {
    NB_LAUNCH(k_bounds, nblk, 256, s, sd, st, work, node4, partial);
}
void launch_bounds_order(hipStream_t s, int nbounds, const BlkWork* wb, const SysDesc* sd, SysState* st,
                         const float4* node4, float* partial, const BlkWork* og, int olen, int* keys, BlkWork* oout) {
    if (olen > kOrderSmall) olen = 0;  // the caller orders with k_force_order then
    NB_LAUNCH(k_bounds_order, (olen > 0 ? 1 : 0) + nbounds, 256, s, sd, st, wb, node4, partial, og, olen, keys, oout);
}
void launch_cell_export(hipStream_t s, int n, int ncells, const int* child, const float4* cdata,
                        const int* slotoff, float4* cells, int* ccount) {
    NB_LAUNCH(k_cell_export, (ncells + 255) / 256, 256, s, n, ncells, child, cdata, slotoff, cells, ccount);
}
void launch_tree_init(hipStream_t s, int nsys, const int* slots, const SysDesc* sd, SysState* st, int* slotoff,
                      int* child, int* cpend) {
    if (nsys <= 0) return;
    NB_LAUNCH(k_tree_init, (nsys + 255) / 256, 256, s, nsys, slots, sd, st, slotoff, child, cpend);
}
void launch_build_top(hipStream_t s, int nsys, const int* slots, const SysDesc* sd, SysState* st,
                      const float4* node4, int* child, int* cparent, int* cpend, float4* cdata, int* slotoff,
                      int stride, int* fblist) {
    if (nsys <= 0) return;
    NB_LAUNCH(k_build_top, nsys, kTopBodies, s, nsys, slots, sd, st, node4, child, cparent, cpend, cdata, slotoff,
              stride, fblist);
}
void launch_build(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                  const float4* node4, int* child, int* cparent, int* cpend, float4* cdata, int* slotoff, int stride,
                  int exclude, const int* list) {
    if (build_block(stride) == kBuildBlockSparse)
        NB_LAUNCH(k_build<kBuildBlockSparse>, nblk, kBuildBlockSparse, s, sd, st, work, node4, child, cparent, cpend,
                  cdata, slotoff, stride, exclude, list);
    else if (build_block(stride) == kBuildBlockStrided)
        NB_LAUNCH(k_build<kBuildBlockStrided>, nblk, kBuildBlockStrided, s, sd, st, work, node4, child, cparent, cpend,
                  cdata, slotoff, stride, exclude, list);
    else
        NB_LAUNCH(k_build<kBuildBlock>, nblk, kBuildBlock, s, sd, st, work, node4, child, cparent, cpend, cdata,
                  slotoff, stride, exclude, list);
}
void launch_summarize(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                      float4* node4, const int* child, const int* cparent, int* cpend, int* ccount, int* slotoff,
                      float4* cdata) {
    NB_LAUNCH(k_summarize, nblk, 256, s, sd, st, work, node4, child, cparent, cpend, ccount, slotoff, cdata);
}
void launch_sort(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work, const float4* node4,
                 const int* child, const int* slotoff, int* sorted, int* forder, const int* sorder) {
    NB_LAUNCH(k_sort, nblk, 256, s, sd, st, work, node4, child, slotoff, sorted, forder, sorder);
}
void launch_force_order(hipStream_t s, const BlkWork* in, int len, int* keys, BlkWork* out) {
    if (len > 0 && len <= kOrderMax) NB_LAUNCH(k_force_order, 1, 1024, s, in, len, keys, out);
}
void launch_force_bh(hipStream_t s, int nblk, bool kahan, bool pair, const SysDesc* sd, SysState* st,
                     const BlkWork* work, const float4* node4, const int* child, const float4* cdata,
                     const int* order, float4* acc4, int* sorder, int* keys) {
#define NB_FBH(K, P)                                                                                          \
    if (kahan == K && pair == P)                                                                              \
    NB_LAUNCH((k_force_bh<K, P>), nblk, 64 * kForceWaves, s, sd, st, work, node4, child, cdata, order, acc4, sorder, \
              keys)
    NB_FBH(false, false); NB_FBH(true, false); NB_FBH(false, true); NB_FBH(true, true);
#undef NB_FBH
}
void launch_range_table(hipStream_t s, BlkWork* out, int nblk, int sys, int start, int stop, int lo, int hi,
                        int per) {
    NB_LAUNCH(k_range_table, (nblk + 255) / 256, 256, s, out, nblk, sys, start, stop, lo, hi, per);
}
void launch_brute(hipStream_t s, int nblk, int law, bool kahan, int tb, const SysDesc* sd, const SysState* st,
                  const BlkWork* work, const float4* node4, float4* acc4, float4* vel4, float4* out4, int mode) {
    const int B = kBfWaves * 64;
#define NB_BF(L, K, T)                                                                                   \
    if (law == L && kahan == K && tb == T)                                                               \
    NB_LAUNCH((k_brute<L, K, T>), nblk, B, s, sd, st, work, node4, acc4, vel4, out4, mode)
    NB_BF(0, false, 1); NB_BF(0, true, 1); NB_BF(1, false, 1); NB_BF(1, true, 1);
    NB_BF(0, false, 2); NB_BF(0, true, 2); NB_BF(1, false, 2); NB_BF(1, true, 2);
#undef NB_BF
}
void launch_collision(hipStream_t s, int nblk, bool all, const SysDesc* sd, const SysState* st, const BlkWork* work,
                      const float4* node4, float4* vel4) {
    NB_LAUNCH(k_collision, nblk, 256, s, sd, st, work, node4, vel4, all ? 1 : 0);
}
void launch_copy4(hipStream_t s, int n, const float4* src, float4* dst) {
    NB_LAUNCH(k_copy4, (n + 255) / 256, 256, s, n, src, dst);
}
void launch_vertices(hipStream_t s, int n, const float4* node4, float4* out) {
    NB_LAUNCH(k_vertices, (n + 255) / 256, 256, s, n, node4, out);
}
void launch_energy(hipStream_t s, int n, const float4* node4, const float4* vel4, int law, double eps, double G,
                   double* partial, double* out) {
    const int nblk = (n + 255) / 256;
    NB_LAUNCH(k_energy_f64, nblk, 256, s, n, node4, vel4, law, eps, G, partial);
    NB_LAUNCH(k_energy_final, 1, 256, s, nblk, partial, out);
}
void launch_update(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                   float4* node4, float4* vel4, const float4* acc4, int mode) {
    NB_LAUNCH(k_update, nblk, 256, s, sd, st, work, node4, vel4, acc4, mode);
}


}  // namespace nb
