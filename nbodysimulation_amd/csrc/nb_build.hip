// nb_build.hip -- the bucketed insertion phases of the tree build (build_tree.cl:42-161
// computes the same tree by lock-free insertion of every body from the root).
//
// After the sparse first phase the tree holds a skeleton T (final while a phase runs).
// A later phase inserts its bodies in four kernels, with no contention between lanes:
//   k_classify    each body of the phase descends T (read-only) to the slot where its
//                 insertion would start, the bucket key 8*cell + slot, and takes a rank
//                 among the bucket's bodies (one atomic add on the bucket's counter);
//   k_frames      per bucket of T's slots: the bucket's size (its bodies, plus the body
//                 already in the slot, if any); buckets of at most kFrame bodies are packed
//                 into frames of at most kFrame bodies, each bucket contiguous in the frame
//                 list; larger buckets are marked for the fallback;
//   k_scatter     each body goes to its bucket's place in the frame list (or to the
//                 fallback list);
//   k_frame_build one workgroup per frame builds every bucket's subtree in LDS, level by level
//                 (each body claims its slot by CAS, a loser marks the slot contested; an
//                 uncontested body is placed, the holder of a contested slot opens a cell
//                 there, the contested bodies move into it), so a cell exists exactly for
//                 every path prefix that two or more of the bucket's bodies share: the
//                 topology that insertion gives, whatever the order.  The workgroup then takes
//                 one range of cell numbers (one atomic add), writes the cells, the bodies'
//                 child-block entries and the link of each subtree into its slot of T.
// Buckets above kFrame bodies, and frames whose subtrees do not fit the LDS cell pool or
// reach the depth limit, go to k_build's queued insertion (list mode) afterwards.
//
// Scratch during the build (each array is written afresh by a later stage):
//   sorted / forder   per item of the phase's enumeration: bucket key and rank (the sort
//                     writes both permutations)
// and of its own: per item the body's position (ppos), the frame list (per entry the body's
// position; its id, bucket key and the
// level of the bucket's cell), the frames (list begin, size), the fallback list, and per slot
// of T the bucket's place in the frame list (kbase)
//   slotoff           per slot of T: the bucket's body count, cleared by k_frames (summarize
//                     writes the per-slot counts; cells are created with cleared counts)
//   ccount            per cell of T that ends a descent: its level (summarize writes counts)
#include <hip/hip_runtime.h>

#include "nb_device.h"
#include "nb_internal.h"

namespace nb {

constexpr int kFrameThreads = 128;             // one k_frame_build workgroup builds one frame
constexpr int kFrame = 4 * kFrameThreads;      // bodies per frame: four per lane
constexpr int kFrameCells = 384;               // cells of one frame's subtrees in LDS
// Bodies packed per frame: the subtrees of STANDARD-rule systems take about 0.48 cells per
// body, those of REFERENCE-rule systems (cells halved in y by the x comparison) about 0.72,
// so that a frame's cells stay within kFrameCells (above it the frame goes to the fallback).
__host__ __device__ constexpr int frame_budget(int rules) { return rules == 0 ? 448 : kFrame; }
constexpr int kCellTag = 1 << 16;              // LDS slot value >= kCellTag: local cell (value - kCellTag)

// The bodies of one insertion phase: b = bi * stride for the block's items bi, without the
// multiples of `exclude` (bodies of earlier phases).  Returns -1 for an item outside it.
// (32-bit arithmetic: the block tables keep bi * stride below n + stride.  Written with a
// 64-bit remainder, hipcc for gfx950 dropped the value on the path where the remainder is
// non-zero, so every strided item came out as -1.)
__device__ __forceinline__ int phase_body(const BlkWork& w, int n, int stride, int exclude) {
    const int bi = w.begin + (int)threadIdx.x;
    const int b = bi * stride;
    const bool in = bi < w.end && b < n && (exclude <= 0 || b % exclude != 0);
    return in ? b : -1;
}

// The body of a phase's item i = w.begin + thread: a phase over every body (stride 1) takes
// the bodies in the last force pass's order when there is one (as the sort does), so that
// the lanes of a wave descend the same cells and fall into few buckets; a strided phase
// takes every stride-th body.  -1: no body of the phase.
__device__ __forceinline__ int phase_item(const BlkWork& w, const SysState* S, const int* sorder, int n, int stride,
                                          int exclude) {
    if (stride != 1 || !S->order_ok) return phase_body(w, n, stride, exclude);
    const int i = w.begin + (int)threadIdx.x;
    const int b = i < w.end && i < n ? min(max(sorder[i], 0), n - 1) : -1;
    return b >= 0 && !(exclude > 0 && b % exclude == 0) ? b : -1;
}

// In the last force pass's order the lanes of a wave fall into few buckets: each bucket's
// count then takes one atomic add per wave (ranks within the wave by ballot), not one per
// body.  Key and rank are stored by the item's place in the enumeration (coalesced), and
// k_scatter enumerates the same way.
__global__ __launch_bounds__(256) void k_classify(const SysDesc* __restrict__ sd, SysState* st,
                                                  const BlkWork* __restrict__ work, const float4* __restrict__ node4,
                                                  const int* __restrict__ child, int* slotoff, int* ccount, int* bkey,
                                                  int* brank, float4* ppos, const int* __restrict__ sorder, int stride,
                                                  int exclude) {
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    if (S->status || d.algo != 0) return;
    // the phase's counters start at zero (the kernels that used them last have completed);
    // the fallback list (nfb) collects every phase's bodies until the insertion after the last
    if (w.begin == 0 && threadIdx.x == 0) {
        S->nframes = 0;
        S->listpos = 0;
        S->bq = 0;
        if (stride == 1) S->nfb1 = S->nfb;  // the last phase: bodies earlier phases left
    }
    const int n = d.n;
    const int i = w.begin + (int)threadIdx.x;  // the phase's enumeration: bkey/brank are indexed by i
    const int b = phase_item(w, S, sorder + d.body_off, n, stride, exclude);
    if (b < 0) return;
    const int* C = child + 8 * (size_t)d.cell_off;
    const float4 p = node4[d.node_off + b];
    float lx, ly, lz, hx, hy, hz;
    root_box(d.rules, S->box, lx, ly, lz, hx, hy, hz);
    int node = 0, lvl = 0, key = -1;
    for (; lvl <= kMaxLevel; ++lvl) {
        const int dg = level_digit(d.rules, lvl, p.x, p.y, p.z, lx, ly, lz, hx, hy, hz);
        const int v = C[8 * node + dg];
        if (v >= n && v - n < d.cap) {
            node = v - n;
            continue;
        }
        if (v >= kEmpty && v < n) key = 8 * node + dg;  // empty, or a body: no insertion runs, so no lock
        break;
    }
    if (key < 0) {  // an inconsistent or too deep tree; cannot happen on a tree the build made
        atomicOr(&S->status, kStStall);
        return;
    }
    // in force order, one count per distinct bucket of the wave: the first lane of each key
    // adds the lanes that share it, each lane's rank follows its place among them (in a
    // strided phase the lanes' bodies are far apart and count one by one)
    int* cnt = slotoff + 8 * (size_t)d.cell_off;
    const int lane = threadIdx.x & 63;
    int rank = 0;
    if (!(stride == 1 && S->order_ok)) {
        rank = atomicAdd(cnt + key, 1);
    } else {
        // groups of lanes by key (no memory traffic), then every group's first lane adds the
        // group's size in one vector atomic: one round trip per wave
        int leader = lane, gsize = 1, below = 0;
        for (unsigned long long todo = __ballot(true); todo;) {
            const int l0 = __ffsll((long long)todo) - 1;
            const int lk = __shfl(key, l0);
            const unsigned long long m = __ballot(key == lk);
            if (key == lk) {
                leader = l0;
                gsize = __popcll(m);
                below = __popcll(m & ((1ull << lane) - 1));
            }
            todo &= ~m;
        }
        int base = 0;
        if (lane == leader) base = atomicAdd(cnt + key, gsize);
        rank = __shfl(base, leader) + below;
    }
    brank[d.body_off + i] = rank;
    bkey[d.body_off + i] = key;
    ppos[d.body_off + i] = p;  // for k_scatter, in the same enumeration (a coalesced read there)
    ccount[d.cell_off + node] = lvl;  // the same value from every body that ends here
}

// One lane per slot of T (grid-stride: the block is w.begin of the system's w.end blocks and
// takes 256 consecutive slots at a time).  Each wave packs its 64 slots' buckets greedily in
// slot order: a frame takes buckets while its total stays within the budget.  With the
// buckets' sizes prefix-summed over the wave, the frame that starts at list offset `base`
// holds exactly the buckets whose inclusive sum exceeds `base` by at most the budget, so a
// frame costs one ballot, not one step per bucket, and a bucket's place in the wave's list
// region is its exclusive sum.  The block then claims its waves' frames and list space with
// one 64-bit atomic add (frames low, list entries high): same-address atomics cost about as
// much as a wave's whole packing, so there is one per block and round.
__global__ __launch_bounds__(256) void k_frames(const SysDesc* __restrict__ sd, SysState* st,
                                                const BlkWork* __restrict__ work, const float4* __restrict__ node4,
                                                const int* __restrict__ child, int* slotoff,
                                                const int* __restrict__ ccount, int* kbase, float4* lpos, int4* lmeta,
                                                int2* bframe) {
    __shared__ int s_nf[4], s_tot[4], s_fb, s_lb, s_ok;
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    if (S->status || d.algo != 0) return;
    const int n = d.n;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nkeys = 8 * min(S->ncells, d.cap);
    const int budget = frame_budget(d.rules);
    const int* C = child + 8 * (size_t)d.cell_off;
    int* cnt8 = slotoff + 8 * (size_t)d.cell_off;
    int* kb = kbase + 8 * (size_t)d.cell_off;
    float4* LP = lpos + d.body_off;
    int4* LM = lmeta + d.body_off;
    int2* frames = bframe + d.body_off;
    for (int k0 = w.begin * 256; k0 < nkeys; k0 += w.end * 256) {  // uniform over the block
        const int k = k0 + (int)threadIdx.x;
        // the slot's count, its content and its cell's level, loaded together
        int cnt = 0, v = -1, lvl = 0;
        if (k < nkeys) {
            cnt = cnt8[k];
            v = C[k];
            lvl = ccount[d.cell_off + (k >> 3)];
        }
        const int res = (cnt && v >= 0 && v < n) ? v : -1;  // the body already in the slot joins the bucket
        float4 rp = make_float4(0.f, 0.f, 0.f, 0.f);
        if (res >= 0) rp = node4[d.node_off + res];
        if (cnt) cnt8[k] = 0;  // cleared for the next phase's counts
        const int size = cnt ? cnt + (res >= 0 ? 1 : 0) : 0;
        const bool big = size > budget;
        const int sz = big ? 0 : size;
        int incl = sz;  // inclusive sum over the wave's lanes
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o);
            if (lane >= o) incl += y;
        }
        const int excl = incl - sz;
        const int tot = __shfl(incl, 63);
        // frames of the wave: lane q < nf receives frame q's offset and size
        int nf = 0, fbeg = 0, fsz = 0;
        for (int base = 0; base < tot; ++nf) {  // uniform: base, tot and nf are the same in every lane
            const bool in = sz > 0 && excl >= base && incl - base <= budget;
            const unsigned long long m = __ballot(in);
            const int last = 63 - __clzll((long long)m);  // in-lanes are contiguous (incl rises)
            const int end = __shfl(incl, last);
            if (lane == nf) {
                fbeg = base;
                fsz = end - base;
            }
            base = end;
        }
        if (lane == 0) {
            s_nf[wave] = nf;
            s_tot[wave] = tot;
        }
        __syncthreads();
        if (threadIdx.x == 0) {  // one claim for the block: frames low, list entries high
            int bnf = 0, btot = 0;
            for (int q = 0; q < 4; ++q) {
                bnf += s_nf[q];
                btot += s_tot[q];
            }
            unsigned long long old = 0;
            if (bnf > 0)
                old = atomicAdd(reinterpret_cast<unsigned long long*>(&S->nframes),
                                (unsigned long long)(unsigned)bnf | ((unsigned long long)(unsigned)btot << 32));
            s_fb = (int)(unsigned)old;
            s_lb = (int)(unsigned)(old >> 32);
            // frames <= buckets <= bodies of the phase; the list holds the phase's bodies and the
            // bodies of earlier phases already in the buckets' slots, so at most n entries
            s_ok = !(bnf > 0 && (s_fb + bnf > n || s_lb + btot > n));
            if (!s_ok) atomicOr(&S->status, kStStall);
        }
        __syncthreads();
        if (!s_ok) return;
        int fb = s_fb, lb = s_lb;
        for (int q = 0; q < wave; ++q) {
            fb += s_nf[q];
            lb += s_tot[q];
        }
        if (sz > 0) {
            const int pos = lb + excl;
            if (res >= 0) {
                LP[pos] = rp;
                LM[pos] = make_int4(res, k, lvl, 0);
            }
            kb[k] = pos + (res >= 0 ? 1 : 0);
        } else if (big) {
            kb[k] = -1;  // its bodies go to the fallback insertion; the body in the slot stays
        }
        if (lane < nf) frames[fb + lane] = make_int2(lb + fbeg, fsz);
        __syncthreads();  // the next round rewrites s_nf and s_tot
    }
}

// Appends the bodies of the lanes with `take` to the system's fallback list (one atomic
// per wave).
__device__ __forceinline__ void to_fallback(bool take, int b, SysState* S, int* fb) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    int j0 = 0;
    if (lane == leader) j0 = atomicAdd(&S->nfb, __popcll(m));
    j0 = __shfl(j0, leader);
    if (take) fb[j0 + __popcll(m & ((1ull << lane) - 1))] = b;
}

// Each entry of the frame list carries the body's position, id, bucket key and the level of
// the bucket's cell, so that a frame build reads its entries in one batch of loads.
__global__ __launch_bounds__(256) void k_scatter(const SysDesc* __restrict__ sd, SysState* st,
                                                 const BlkWork* __restrict__ work, const float4* __restrict__ ppos,
                                                 const int* __restrict__ kbase, const int* __restrict__ bkey,
                                                 const int* __restrict__ brank, const int* __restrict__ ccount,
                                                 const int* __restrict__ sorder, float4* lpos, int4* lmeta, int* fblist,
                                                 int stride, int exclude) {
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    if (S->status || d.algo != 0) return;
    const int n = d.n;
    const int i = w.begin + (int)threadIdx.x;  // the classify's enumeration
    const int b = phase_item(w, S, sorder + d.body_off, n, stride, exclude);
    if (b < 0) return;
    const int key = bkey[d.body_off + i];
    const int base = kbase[8 * (size_t)d.cell_off + key];
    if (base >= 0) {
        const int pos = base + brank[d.body_off + i];
        if (pos < n) {
            lpos[d.body_off + pos] = ppos[d.body_off + i];
            lmeta[d.body_off + pos] = make_int4(b, key, ccount[d.cell_off + (key >> 3)], 0);
        }
    }
    to_fallback(base < 0, b, S, fblist + d.body_off);
}

struct FrameLds {
    int c[8 * kFrameCells];  // slots: -1 empty, [0, kFrame) an entry of the frame, >= kCellTag a cell
    int par[kFrameCells];    // (parent cell << 3) | slot; a bucket's root: -1 - key (its slot of T)
    int cont[kFrameCells];   // bit k: slot k claimed by more than one body this level; bits 8..: the cell's level
    int body[kFrame];        // entry -> body
    int key[kFrame];         // entry -> bucket key; for a leader, later its bucket's root cell
    int wmax[kFrame / 64];  // per wave (up to one entry per lane)
    int wmin[kFrame / 64];
    int nc, fail, base;
};

// One workgroup per frame (grid-stride over the system's frames): entries e = EPT * thread + t
// in frame order, every bucket's subtree built at once, level by level.  At most 128 VGPRs
// (four waves per SIMD): left free the compiler takes 129, which allows three, fewer
// workgroups per CU than the LDS allows (eight).  The last phase (thousands of frames, each
// CU full) runs four entries per lane; the strided phase (a few dozen frames, one per CU)
// runs one, which shortens each level's chain of dependent steps.
template <int EPT>  // entries per lane
__global__ __launch_bounds__(kFrame / EPT) __attribute__((amdgpu_waves_per_eu(4))) void k_frame_build(
    const SysDesc* __restrict__ sd, SysState* st, const BlkWork* __restrict__ work, float4* node4, int* child,
    int* cparent, int* cpend, int* ccount, int* slotoff, float4* cdata, const float4* __restrict__ lpos,
    const int4* __restrict__ lmeta, const int2* __restrict__ bframe, int* fblist, int last) {
    constexpr int kThreads = kFrame / EPT;
    __shared__ FrameLds L;
    const BlkWork w = work[blockIdx.x];
    const SysDesc d = sd[w.sys];
    SysState* S = st + w.sys;
    if (S->status || d.algo != 0) return;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int n = d.n;
    const int nf = min(S->nframes, n);  // k_frames keeps it within n (else the status is set)
    int* C = child + 8 * (size_t)d.cell_off;
    float4* CD = cdata + 8 * (size_t)d.cell_off;
    float rlx, rly, rlz, rhx, rhy, rhz;
    root_box(d.rules, S->box, rlx, rly, rlz, rhx, rhy, rhz);
    int depth = 1;
    // The last phase's frames also summarize their subtrees, unless earlier phases left bodies
    // for the queued insertion (which could then land in them): nfb1 counts those.
    const bool summ = last && ld_c(&S->nfb1) == 0;
    for (int f = w.begin; f < nf; f += w.end) {  // uniform over the block
        const int2 fr = bframe[d.body_off + f];
        const int size = min(max(fr.y, 0), kFrame);
        int eb[EPT], ek[EPT], node[EPT], lvl[EPT], slot[EPT], dgs[EPT], lv0[EPT];
        float4 p[EPT];
        float bx[EPT][6];
        bool act[EPT], valid[EPT];
#pragma unroll
        for (int t = 0; t < EPT; ++t) {
            const int e = EPT * tid + t;
            valid[t] = e < size && fr.x + e < n;
            int4 mt = make_int4(-1, -1, 0, 0);
            p[t] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (valid[t]) {
                mt = lmeta[d.body_off + fr.x + e];
                p[t] = lpos[d.body_off + fr.x + e];
            }
            eb[t] = valid[t] ? min(max(mt.x, 0), n - 1) : -1;
            ek[t] = valid[t] ? min(max(mt.y, 0), 8 * d.cap - 1) : -1;
            lv0[t] = min(max(mt.z, 0), kMaxLevel);
            L.key[EPT * tid + t] = ek[t];
            if (valid[t]) L.body[EPT * tid + t] = eb[t];
        }
        if (tid == 0) {
            L.nc = 0;
            L.fail = 0;
        }
        __syncthreads();
        // bucket leaders (the first entry of each key) and each entry's leader: an inclusive
        // max-scan of (leader ? e : -1) over the frame (in the thread, the wave, the block)
        bool lead[EPT], multi[EPT];
        int li[EPT];
#pragma unroll
        for (int t = 0; t < EPT; ++t) {
            const int e = EPT * tid + t;
            const int prev = e > 0 ? L.key[e - 1] : -1;
            const int next = e + 1 < size ? L.key[e + 1] : -1;
            lead[t] = valid[t] && (e == 0 || prev != ek[t]);
            multi[t] = valid[t] && (next == ek[t] || !lead[t]);
            li[t] = lead[t] ? e : -1;
            if (t) li[t] = max(li[t], li[t - 1]);
        }
        int scan = li[EPT - 1];
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(scan, o);
            if (lane >= o) scan = max(scan, y);
        }
        if (lane == 63) L.wmax[wave] = scan;
        int before = __shfl_up(scan, 1);
        if (lane == 0) before = -1;
        __syncthreads();
        for (int q = 0; q < wave; ++q) before = max(before, L.wmax[q]);
#pragma unroll
        for (int t = 0; t < EPT; ++t) li[t] = max(li[t], before);
        // each entry's box down to its bucket's slot of T (the level of the slot's cell came
        // from k_classify); the bucket leaders open the bucket roots
#pragma unroll
        for (int t = 0; t < EPT; ++t) {
            act[t] = false;
            lvl[t] = lv0[t] + 1;
            if (!valid[t]) continue;
            float lx = rlx, ly = rly, lz = rlz, hx = rhx, hy = rhy, hz = rhz;
            int dg = 0;
            for (int l = 0; l <= lv0[t]; ++l) dg = level_digit(d.rules, l, p[t].x, p[t].y, p[t].z, lx, ly, lz, hx, hy, hz);
            bx[t][0] = lx; bx[t][1] = ly; bx[t][2] = lz; bx[t][3] = hx; bx[t][4] = hy; bx[t][5] = hz;
            if (dg != (ek[t] & 7)) L.fail = 2;  // not the slot the descent ended at: cannot happen
            if (lead[t] && multi[t]) {
                const int c = min(atomicAdd(&L.nc, 1), kFrameCells - 1);  // at most kFrame / 2 roots
                for (int k = 0; k < 8; ++k) L.c[8 * c + k] = kEmpty;
                L.par[c] = -1 - ek[t];
                L.cont[c] = (lv0[t] + 1) << 8;
                L.key[EPT * tid + t] = c;  // read only by this bucket's entries from here on
            }
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < EPT; ++t)
            if (multi[t]) {
                node[t] = min(max(L.key[min(max(li[t], 0), kFrame - 1)], 0), kFrameCells - 1);
                act[t] = true;
            }
        __syncthreads();
        // level by level, every bucket's subtree at once
        while (true) {
            bool any = false;
#pragma unroll
            for (int t = 0; t < EPT; ++t) any = any || act[t];
            if (!__syncthreads_or(any) || L.fail) break;
#pragma unroll
            for (int t = 0; t < EPT; ++t) {
                if (!act[t]) continue;
                if (lvl[t] > kMaxLevel) {
                    L.fail = 1;  // the depth limit: the insertion reports it
                    act[t] = false;
                    continue;
                }
                float lx = bx[t][0], ly = bx[t][1], lz = bx[t][2], hx = bx[t][3], hy = bx[t][4], hz = bx[t][5];
                const int dg = level_digit(d.rules, lvl[t], p[t].x, p[t].y, p[t].z, lx, ly, lz, hx, hy, hz);
                bx[t][0] = lx; bx[t][1] = ly; bx[t][2] = lz; bx[t][3] = hx; bx[t][4] = hy; bx[t][5] = hz;
                dgs[t] = dg;
                slot[t] = 8 * node[t] + dg;
                if (atomicCAS(&L.c[slot[t]], kEmpty, EPT * tid + t) != kEmpty) atomicOr(&L.cont[node[t]], 1 << dg);
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < EPT; ++t) {
                if (!act[t]) continue;
                if (!((L.cont[node[t]] >> dgs[t]) & 1)) {  // alone in its slot: placed
                    act[t] = false;
                    depth = max(depth, lvl[t]);
                } else if (L.c[slot[t]] == EPT * tid + t) {  // the holder of a contested slot opens a cell
                    const int c = atomicAdd(&L.nc, 1);
                    if (c >= kFrameCells) {
                        L.fail = 1;
                    } else {
                        for (int k = 0; k < 8; ++k) L.c[8 * c + k] = kEmpty;
                        L.par[c] = (node[t] << 3) | dgs[t];
                        L.cont[c] = (lvl[t] + 1) << 8;
                        L.c[slot[t]] = kCellTag + c;
                    }
                }
            }
            __syncthreads();
            if (L.fail) break;  // uniform: read between two barriers
#pragma unroll
            for (int t = 0; t < EPT; ++t)
                if (act[t]) {  // into the new cell, one level down
                    node[t] = min(max(L.c[slot[t]] - kCellTag, 0), kFrameCells - 1);
                    ++lvl[t];
                }
        }
        __syncthreads();
        const int fail = L.fail;
        if (fail == 2 && tid == 0) atomicOr(&S->status, kStStall);
        if (fail) {
            // nothing of this frame was written: its bodies go to the insertion, except the
            // ones already in their slots of T
#pragma unroll
            for (int t = 0; t < EPT; ++t) to_fallback(valid[t] && C[ek[t]] != eb[t], eb[t], S, fblist + d.body_off);
            __syncthreads();
            continue;
        }
        const int nc = L.nc;
        if (tid == 0) L.base = nc > 0 ? atomicAdd(&S->ncells, nc) : 0;
        __syncthreads();
        const int base = L.base;
        if (base + nc > d.cap) {
            if (tid == 0) atomicOr(&S->status, kStTreeOverflow);
            __syncthreads();
            continue;
        }
        // the cells: slots, parent links, counts of child cells, cleared per-slot counts
        for (int c = tid; c < nc; c += kThreads) {
            const int g = base + c;
            int v[8], np = 0;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int sv = L.c[8 * c + k];
                v[k] = sv == kEmpty ? kEmpty : (sv >= kCellTag ? n + base + (sv - kCellTag) : L.body[min(sv, kFrame - 1)]);
                np += sv >= kCellTag;
            }
            reinterpret_cast<int4*>(C + 8 * (size_t)g)[0] = make_int4(v[0], v[1], v[2], v[3]);
            reinterpret_cast<int4*>(C + 8 * (size_t)g)[1] = make_int4(v[4], v[5], v[6], v[7]);
            const int pv = L.par[c];
            cpend[d.cell_off + g] = summ ? (pv < 0 ? kSumRoot : kSumDone) : np;
            // parent link (k_summarize's; a summarized cell's only for a bucket root's count)
            if (!summ || pv < 0) cparent[d.cell_off + g] = pv < 0 ? (-1 - pv) : (((base + (pv >> 3)) << 3) | (pv & 7));
            if (!last) {  // a later phase counts its buckets here
                int4* so = reinterpret_cast<int4*>(slotoff + 8 * ((size_t)d.cell_off + g));
                so[0] = make_int4(0, 0, 0, 0);
                so[1] = make_int4(0, 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < EPT; ++t) {
            if (!valid[t]) continue;
            if (multi[t]) {  // placed in a new cell: the body's entry of that cell's child block
                CD[8 * (size_t)(base + node[t]) + dgs[t]] = p[t];
                if (lead[t]) {  // the subtree takes the bucket's slot of T
                    C[ek[t]] = n + base + L.key[EPT * tid + t];
                    atomicAdd(cpend + d.cell_off + (ek[t] >> 3), 1);
                }
            } else {  // alone in its bucket: straight into the slot
                C[ek[t]] = eb[t];
                CD[ek[t]] = p[t];
                depth = max(depth, lv0[t]);
            }
        }
        if (summ) {
            // The subtrees summarized bottom-up, one level per round, each cell's children in
            // slot order with k_summarize's arithmetic (so the same bits): a cell's (COM, mass)
            // and body count go to its parent's child block and per-slot counts (for a bucket's
            // root, its slot of T, where k_summarize reads them) and into the cell's own first
            // LDS slots, read by nothing else once the cell is done, for its parent's round.
            __syncthreads();  // every cell's slots are written out (above)
            int lmax = 0, lmin = kMaxLevel + 1;
            for (int c = tid; c < nc; c += kThreads) {
                const int lv = L.cont[c] >> 8;
                lmax = max(lmax, lv);
                lmin = min(lmin, lv);
            }
            lmax = wave_max_i(lmax);
            lmin = -wave_max_i(-lmin);
            if (lane == 0) {
                L.wmax[wave] = lmax;
                L.wmin[wave] = lmin;
            }
            __syncthreads();
            for (int q = 0; q < kThreads / 64; ++q) {
                lmax = max(lmax, L.wmax[q]);
                lmin = min(lmin, L.wmin[q]);
            }
            const float4* FP = lpos + d.body_off + fr.x;
            int* SO = slotoff + 8 * (size_t)d.cell_off;
            for (int lv = lmax; lv >= lmin; --lv) {  // uniform
                for (int c = tid; c < nc; c += kThreads) {
                    if ((L.cont[c] >> 8) != lv) continue;
                    int sv[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) sv[k] = L.c[8 * c + k];
                    float m = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
                    int cnt = 0, incl[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const int v = sv[k];
                        if (v != kEmpty) {
                            float4 q;
                            int cc = 1;
                            if (v >= kCellTag) {  // a child cell, summarized in an earlier round
                                const int* s2 = L.c + 8 * min(v - kCellTag, kFrameCells - 1);
                                q = make_float4(__int_as_float(s2[0]), __int_as_float(s2[1]), __int_as_float(s2[2]),
                                                __int_as_float(s2[3]));
                                cc = s2[4];
                            } else {
                                q = FP[min(v, size - 1)];  // a valid entry of this frame
                            }
                            const float wm = q.w;
                            cnt += cc;
                            m = m + wm;
                            cx = fmaf(q.x, wm, cx);
                            cy = fmaf(q.y, wm, cy);
                            cz = fmaf(q.z, wm, cz);
                        }
                        incl[k] = cnt;
                    }
                    const float inv = 1.0f / m;
                    const float4 com = make_float4(cx * inv, cy * inv, cz * inv, m);
                    const int g = base + c;
                    int4* so = reinterpret_cast<int4*>(SO + 8 * (size_t)g);
                    so[0] = make_int4(incl[0], incl[1], incl[2], incl[3]);
                    so[1] = make_int4(incl[4], incl[5], incl[6], incl[7]);
                    const int pv = L.par[c];
                    // the parent's child block entry; the count only for a bucket's root (a frame
                    // cell's per-slot counts are its own inclusive ones, written above)
                    const size_t e = pv >= 0 ? 8 * (size_t)(base + (pv >> 3)) + (pv & 7) : (size_t)(-1 - pv);
                    CD[e] = com;
                    if (pv < 0) SO[e] = cnt;
                    L.c[8 * c] = __float_as_int(com.x);
                    L.c[8 * c + 1] = __float_as_int(com.y);
                    L.c[8 * c + 2] = __float_as_int(com.z);
                    L.c[8 * c + 3] = __float_as_int(com.w);
                    L.c[8 * c + 4] = cnt;
                }
                __syncthreads();
            }
        }
        __syncthreads();  // the next frame reuses the LDS
    }
    depth = wave_max_i(depth);
    if (lane == 0 && depth > ld_c(&S->max_depth)) atomicMax(&S->max_depth, depth);
}

void launch_classify(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                     const float4* node4, const int* child, int* slotoff, int* ccount, int* bkey, int* brank,
                     float4* ppos, const int* sorder, int stride, int exclude) {
    if (nblk > 0)
        hipLaunchKernelGGL(k_classify, dim3(nblk), dim3(256), 0, s, sd, st, work, node4, child, slotoff, ccount, bkey,
                           brank, ppos, sorder, stride, exclude);
}
void launch_frames(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                   const float4* node4, const int* child, int* slotoff, const int* ccount, int* kbase, float4* lpos,
                   int4* lmeta, int2* bframe) {
    if (nblk > 0)
        hipLaunchKernelGGL(k_frames, dim3(nblk), dim3(256), 0, s, sd, st, work, node4, child, slotoff, ccount, kbase,
                           lpos, lmeta, bframe);
}
void launch_scatter(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                    const float4* ppos, const int* kbase, const int* bkey, const int* brank, const int* ccount,
                    const int* sorder, float4* lpos, int4* lmeta, int* fblist, int stride, int exclude) {
    if (nblk > 0)
        hipLaunchKernelGGL(k_scatter, dim3(nblk), dim3(256), 0, s, sd, st, work, ppos, kbase, bkey, brank, ccount,
                           sorder, lpos, lmeta, fblist, stride, exclude);
}
void launch_frame_build(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                        float4* node4, int* child, int* cparent, int* cpend, int* ccount, int* slotoff,
                        float4* cdata, const float4* lpos, const int4* lmeta, const int2* bframe, int* fblist,
                        bool last) {
    if (nblk <= 0) return;
    if (last)
        hipLaunchKernelGGL(k_frame_build<4>, dim3(nblk), dim3(kFrame / 4), 0, s, sd, st, work, node4, child, cparent,
                           cpend, ccount, slotoff, cdata, lpos, lmeta, bframe, fblist, 1);
    else
        hipLaunchKernelGGL(k_frame_build<1>, dim3(nblk), dim3(kFrame), 0, s, sd, st, work, node4, child, cparent,
                           cpend, ccount, slotoff, cdata, lpos, lmeta, bframe, fblist, 0);
}

}  // namespace nb
