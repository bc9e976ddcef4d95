// nb_engine.h -- host runtime of the N-body engine: device arena, block tables,
// stage launches for a batch of independent systems.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "nb_internal.h"
#include "nbody/nbody.h"

namespace nb {

// kernel launchers (nb_kernels.hip)
void launch_bounds(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                   const float4* node4, float* partial);
void launch_tree_clear(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                       int* child, int* cpend);
void launch_tree_init(hipStream_t s, int nsys, SysState* st);
void launch_build(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work,
                  const float4* node4, int* child, int* cparent, int* cpend, int stride, int exclude);
void launch_summarize(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                      float4* node4, const int* child, const int* cparent, int* cpend, int* ccount, int* slotoff,
                      float4* cdata);
void launch_sort(hipStream_t s, int nblk, const SysDesc* sd, SysState* st, const BlkWork* work, const float4* node4,
                 const int* child, const int* slotoff, int* sorted);
void launch_force_bh(hipStream_t s, int nblk, bool kahan, const SysDesc* sd, const SysState* st,
                     const BlkWork* work, const float4* node4, const int* child, const float4* cdata,
                     const int* sorted, float4* acc4);
void launch_brute(hipStream_t s, int nblk, int law, bool kahan, const SysDesc* sd, const SysState* st,
                  const BlkWork* work, const float4* node4, float4* acc4);
void launch_update(hipStream_t s, int nblk, const SysDesc* sd, const SysState* st, const BlkWork* work,
                   float4* node4, float4* vel4, const float4* acc4, int mode);

void set_error(const std::string& msg);

// One system of a batch, as requested by the caller.
struct SysSpec {
    int n = 0;
    int cap = 0;  // cells (BH only)
    int rules = 0, algo = 0, law = 1, accum = 0, method = 2;
    float G = 1.f, theta = 0.5f, eps = 0.01f, dt = 1e-3f;
};

// Device-side batch of systems with its block tables.
struct Engine {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;

    std::vector<SysSpec> spec;
    std::vector<SysDesc> hdesc;
    int nsys = 0;
    long long tot_nodes = 0, tot_bodies = 0, tot_cells = 0;
    bool any_bh = false, any_bf = false;
    int law = 1;
    bool kahan = false;

    void* arena = nullptr;
    size_t arena_bytes = 0;
    float4 *node4 = nullptr, *vel4 = nullptr, *acc4 = nullptr, *cdata = nullptr;
    int *sorted = nullptr, *child = nullptr, *cparent = nullptr, *cpend = nullptr, *ccount = nullptr,
        *slotoff = nullptr;
    SysDesc* d_desc = nullptr;
    SysState* d_state = nullptr;
    float* d_partial = nullptr;
    static constexpr int kBuildPhases = 3;
    BlkWork* t_phase[kBuildPhases] = {};
    int nb_phase[kBuildPhases] = {};
    BlkWork *t_bounds = nullptr, *t_build = nullptr, *t_cells = nullptr, *t_groups = nullptr, *t_brute = nullptr,
            *t_update = nullptr;
    int nb_bounds = 0, nb_build = 0, nb_cells = 0, nb_groups = 0, nb_brute = 0, nb_update = 0;

    ~Engine();
    nb_status init(const std::vector<SysSpec>& systems, void* user_stream);
    void release();

    // whole-batch stage launches (asynchronous on `stream`)
    void bounds();
    void tree_reset();
    void build();
    void summarize();
    void sort();
    void force();
    void brute();
    void update(int mode);

    nb_status download_state(std::vector<SysState>& out);
};

size_t align_up(size_t x, size_t a);

}  // namespace nb
