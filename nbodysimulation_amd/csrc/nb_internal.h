// nb_internal.h -- device data layout shared by the HIP kernels and the host engine.
//
// Every kernel is written for a batch of independent systems ("ensemble"); a
// single simulation is a batch of one.  Per-system constants live in SysDesc,
// per-system mutable device state in SysState, and launches are described by
// block tables (one BlkWork per workgroup) built on the host.
//
// HBM layout (all systems concatenated):
//   node4  float4[sum(n_s + cap_s)]  bodies [node_off, node_off+n) as (x,y,z,m),
//                                    cells  [node_off+n, node_off+n+cap) as (com, mass)
//   vel4   float4[sum n_s]           velocity (w unused)
//   acc4   float4[sum n_s]           acceleration (w = interactions of the last force pass)
//   sorted int[sum n_s]              local body ids in tree order (children in slot order)
//   forder int[sum n_s]              local body ids in force order (children along a Hilbert
//                                    curve); the force pass's vote groups are its 64-runs
//   sorder int[sum n_s]              the order in which the sort visits the bodies: the force
//                                    order of the last full force pass (identity before one)
//   child  int[8 * sum cap_s]        per cell 8 slots: -1 empty, -2 locked,
//                                    [0,n) body, n+ci cell ci (local numbering)
//   cparent, cpend, ccount           int[sum cap_s]   parent, unfinished child cells, bodies below
//   slotoff                          int[8 * sum cap_s] bodies in slots 0..k of the same cell
//   cdata                            float4[8 * sum cap_s] children (x,y,z,m) in slot order
//   kbase                            int[8 * sum cap_s] build scratch: a bucket's place in the frame list
//   lpos float4[sum n_s], lmeta int4[sum n_s], bframe int2[sum n_s], fblist int[sum n_s]
//                                    build scratch: frame list (position; body, key, level),
//                                    frames (list begin, size), fallback list
#pragma once
#include <stdint.h>

namespace nb {

constexpr int kMaxLevel = 64;          // deepest cell level (root = 0)
constexpr int kTopBodies = 512;        // most bodies of a first insertion phase built in LDS (k_build_top)
constexpr int kDqLen = kMaxLevel + 1;  // opening thresholds per level
constexpr int kBfWaves = 16;    // brute-force workgroup: 16 waves of 64
// Build workgroups: the queued last phase shares one allocation atomic per round among
// 1024 lanes; the strided skeleton phases (few bodies, latency-bound rounds) spread
// over more CUs in smaller workgroups: 256 lanes, and one wave for the sparsest phase.
constexpr int kBuildBlock = 1024;
constexpr int kBuildBlockStrided = 256;
constexpr int kBuildBlockSparse = 64;
constexpr int build_block(int stride) {
    return stride > 64 ? kBuildBlockSparse : stride > 1 ? kBuildBlockStrided : kBuildBlock;
}
constexpr int kBfTileStride = 4 * 17;  // a wave's LDS tile: up to 4 parts of 16 sources, rows padded
#ifndef NB_FORCE_WAVES
#define NB_FORCE_WAVES 4
#endif
constexpr int kForceWaves = NB_FORCE_WAVES;  // Barnes-Hut force workgroup: waves of 64
constexpr int kForceVote = 64;  // bodies per opening decision: one wave, one lane per body
constexpr int kOrderSmall = 8192;  // force workgroups that one 256-lane workgroup orders (keys in registers)
constexpr int kOrderMax = 32768;  // force workgroups that k_force_order can order (keys in LDS)
constexpr int kBfFuseMax = 16384;  // single systems up to this size run the integrator inside k_brute
constexpr int kEnergyLen = 9;   // kinetic, potential, p[3], L[3], mass
constexpr int kEmpty = -1;
constexpr int kLocked = -2;

// cpend markers of cells that the last phase's frame build summarized (their (COM, mass),
// counts and child blocks are final): summarize skips kSumDone; from kSumRoot (a bucket's
// root) it only hands the count down to the cell above
constexpr int kSumDone = -(1 << 30);
constexpr int kSumRoot = -(1 << 30) + 1;

// status bits (SysState::status)
constexpr int kStTreeOverflow = 1;
constexpr int kStDepthOverflow = 2;
constexpr int kStNonfinite = 4;
constexpr int kStStall = 8;  // a bounded wait ran out (inconsistent tree); never expected

struct SysDesc {
    int n;          // bodies
    int cap;        // cell capacity
    int node_off;   // offset into node4
    int body_off;   // offset into vel4/acc4/sorted
    int cell_off;   // offset into cell arrays (child uses 8*cell_off)
    int rules;      // 0 reference, 1 standard
    int algo;       // 0 BH, 1 BF global, 2 BF local
    int law;        // 0 plummer, 1 reference unit
    int accum;      // 0 plain, 1 kahan
    int method;     // 1 leapfrog, 2 euler
    float G, theta, eps, dt, ueps;
    int bbox_first;   // first global block of this system in the bounds table
    int bbox_blocks;  // number of bounds blocks of this system
    int bf_split_log2;  // brute force: log2 of the waves sharing one target group
    int bf_lanes_log2;  // brute force: log2 of the lanes sharing one target
    int collision;      // 1: the batch's collision stage includes this system
    int grp_first;      // first workgroup of this system in the force block table
    int pad[3];
};

struct SysState {
    float box[6];
    int status;
    int max_depth;
    // ncells and bq are the two halves of one 8-byte-aligned word: the build allocates
    // cells and claims queued bodies with a single 64-bit atomic add
    int ncells;
    int bq;         // next entry of the build's last-phase body queue
    int ticket;     // bounds last-block ticket (self-resetting)
    int step;
    float dq[kDqLen];
    int order_ok;   // 1: sorder holds a permutation of the bodies (the last full force pass's order)
    int top;        // this step's first insertion phase: 1 built in LDS (k_build_top); 2 its bodies
                    // went to the fallback list (the LDS cell pool overflowed); 0 neither
    int nfb1;       // the fallback list's length when the last bucketed phase began
    // a bucketed build phase: frames and frame-list entries claimed (the two halves of one
    // 8-byte-aligned word, claimed together), and bodies sent to the fallback insertion
    int nframes;
    int listpos;
    int nfb;
    int pad2;
};

// One workgroup's share of one system: items [begin, end).  `trim` is only set
// by target-range force launches: its low and high 16 bits count the leading and
// trailing items of the share that take part in the work (a vote group keeps all
// its members) but whose results are not written.
struct BlkWork {
    int sys;
    int begin;
    int end;
    int trim;
};

}  // namespace nb
