// nb_device.h -- device helpers shared by the kernel files: coherent loads and
// stores, wave reductions, the octant rule and the Hilbert orientation tables.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "nb_internal.h"

namespace nb {

// ---------------------------------------------------------------------------
// helpers

__device__ __forceinline__ int ld_c(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float ld_cf(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_c64(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_c(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_cf(float* p, float v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_rel(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ bool cas_c(int* p, int expect, int desired) {
    return __hip_atomic_compare_exchange_strong(p, &expect, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
}

// DPP row reduction (16 lanes) followed by a read of the four row results:
// every lane returns the wave-wide min / max.
template <int CTRL>
__device__ __forceinline__ float dppc(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
__device__ __forceinline__ float wave_min(float v) {
    v = fminf(v, dppc<0xB1>(v));   // quad_perm [1,0,3,2]
    v = fminf(v, dppc<0x4E>(v));   // quad_perm [2,3,0,1]
    v = fminf(v, dppc<0x124>(v));  // row_ror:4
    v = fminf(v, dppc<0x128>(v));  // row_ror:8
    float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fminf(fminf(r0, r1), fminf(r2, r3));
}
__device__ __forceinline__ float wave_max(float v) {
    v = fmaxf(v, dppc<0xB1>(v));
    v = fmaxf(v, dppc<0x4E>(v));
    v = fmaxf(v, dppc<0x124>(v));
    v = fmaxf(v, dppc<0x128>(v));
    float r0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 0));
    float r1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 16));
    float r2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 32));
    float r3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 48));
    return fmaxf(fmaxf(r0, r1), fmaxf(r2, r3));
}
__device__ __forceinline__ int wave_max_i(int v) {
    for (int o = 32; o >= 1; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Root cell box.  REFERENCE: the bounding box (build_tree.cl:53-54).  STANDARD:
// cube of half-edge r = half the largest extent about the box centre.
__device__ __forceinline__ void root_box(int rules, const float* box, float& lx, float& ly, float& lz,
                                         float& hx, float& hy, float& hz) {
    if (rules == 0) {
        lx = box[0]; ly = box[1]; lz = box[2]; hx = box[3]; hy = box[4]; hz = box[5];
        return;
    }
    float r = 0.5f * fmaxf(fmaxf(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
    float cx = (box[0] + box[3]) * 0.5f, cy = (box[1] + box[4]) * 0.5f, cz = (box[2] + box[5]) * 0.5f;
    lx = cx - r; hx = cx + r;
    ly = cy - r; hy = cy + r;
    lz = cz - r; hz = cz + r;
}

// Octant 4*[y<=cy] + 2*[x<=cx] + [z<=cz] (calculateOctantIdx, build_tree.hcl:14-29,
// ties to the lower half) and the halving that follows (adjustBoundaryValues,
// build_tree.hcl:31-54; REFERENCE rules halve y by the x comparison, :42).
// Under REFERENCE rules level 0 is the root's single hop into slot 0.
__device__ __forceinline__ int level_digit(int rules, int lvl, float px, float py, float pz, float& lx,
                                           float& ly, float& lz, float& hx, float& hy, float& hz) {
    if (rules == 0 && lvl == 0) return 0;
    float cx = (lx + hx) * 0.5f, cy = (ly + hy) * 0.5f, cz = (lz + hz) * 0.5f;
    bool gx = px > cx, gy = py > cy, gz = pz > cz;
    int d = (gy ? 0 : 4) + (gx ? 0 : 2) + (gz ? 0 : 1);
    if (gx) lx = cx; else hx = cx;
    bool ys = rules == 0 ? gx : gy;
    if (ys) ly = cy; else hy = cy;
    if (gz) lz = cz; else hz = cz;
    return d;
}

// Order LDS accesses of one wave (same-wave LDS operations complete in order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bool finite3(float4 p) {
    return isfinite(p.x) && isfinite(p.y) && isfinite(p.z);
}


// Force order (the vote groups' order): depth-first with each cell's children
// visited along a 3-D Hilbert curve, so consecutive bodies stay in face-sharing
// cells.  24 orientations, the root's is 0.  kHPos[s] packs, 3 bits per octant,
// the octant's place in the visiting order of orientation s; kHNext[s] packs,
// 5 bits per octant, the child's orientation.  The tables follow the Hilbert
// index of Skilling ("Programming the Hilbert curve", AIP Conf. Proc. 707, 2004)
// on the octree's octants 4[y<=c]+2[x<=c]+[z<=c].
static __constant__ unsigned kHPos[24] = {
    0x07e6a5, 0x1ce715, 0xc78aa3, 0x97e681, 0x8f8ab1, 0x05afa5, 0x1dc395, 0x9d8b91,
    0x0fc2b5, 0xdec213, 0x9ee611, 0x0caf35, 0x95af81, 0x8eae31, 0xc6ae23, 0xdc8b13,
    0x458ba7, 0xd7c283, 0x4c8b37, 0x55c387, 0x54e707, 0xd4af03, 0x4ec237, 0x46e627};
static __constant__ unsigned long long kHNext[24] = {
    0x3148310020ull, 0x41d6a78021ull, 0x8c107109e0ull, 0xa0c4020e6cull, 0x90e3021107ull, 0x081c56b0c5ull,
    0x5a92541cc6ull, 0x82552398e4ull, 0x2c46b32088ull, 0x7852eb3526ull, 0xbbd413d94dull, 0x0542b621abull,
    0x9b1ac71683ull, 0xb31adadd6aull, 0xaddce5a925ull, 0x49a44789e1ull, 0x80aef84cf6ull, 0x144159c588ull,
    0x951e2912d3ull, 0x9cecc8c2b7ull, 0xa52b790e30ull, 0x756952c46bull, 0xb4ecd4ba54ull, 0xbbaf482552ull};


// One level of the Hilbert walk: the place of octant `dg` in the visiting order of
// orientation `hs` (0..7), and the child's orientation.
__device__ __forceinline__ int hilbert_pos(int hs, int dg) { return (int)(kHPos[hs] >> (3 * dg)) & 7; }
__device__ __forceinline__ int hilbert_next(int hs, int dg) { return (int)(kHNext[hs] >> (5 * dg)) & 31; }

}  // namespace nb
