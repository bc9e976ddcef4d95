// nb_noalloc.hip -- nb_energy_f64_noalloc: the FP64 conservation diagnostics of
// nb_sim_energy over caller-owned device arrays, for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops).
//
// nb_sim_energy's kernels (k_energy_f64 + k_energy_final) need one scratch buffer:
// kEnergyLen partial sums per 256-body workgroup.  nb_sim_energy allocates it on
// its first call; here it is the file-scope device array `scratch`, sized for
// NB_NOALLOC_MAX_BODIES bodies.  Each device holds its own copy of the array at
// its own address, looked up once per device with hipGetSymbolAddress (the
// array's name in host code is only a host-side handle, never a device address).
// Every workgroup overwrites its own partials before k_energy_final reads them,
// so nothing in `scratch` has to start at zero, and calls on one device share it:
// calls that overlap in time on one device race.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <string>

#include "nb_engine.h"
#include "nbody/nbody.h"

static constexpr int kMaxBlocks = NB_NOALLOC_MAX_BODIES / 256;
static_assert(NB_NOALLOC_MAX_BODIES % 256 == 0, "whole 256-body workgroups");

static __device__ __align__(256) double scratch[(size_t)nb::kEnergyLen * kMaxBlocks];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};  // per device: the device address of `scratch`

nb_status fail(nb_status s, const std::string& msg) {
    nb::set_error(msg);
    return s;
}

// The device address of the current device's `scratch`, looked up on the first call
// on that device (with that device current).
nb_status scratch_ptr(double** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(NB_E_NO_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail(NB_E_INVALID_ARG, "device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NB_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return NB_OK;
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

}  // namespace

extern "C" nb_status nb_energy_f64_noalloc(const float* bodies_xyzm, const float* vel_xyzw, int32_t n, int32_t law,
                                           double eps, double G, double* out, void* stream) {
    if (n < 0 || n > NB_NOALLOC_MAX_BODIES)
        return fail(NB_E_INVALID_ARG, "n = " + std::to_string(n) + " outside [0, NB_NOALLOC_MAX_BODIES = " +
                                          std::to_string(NB_NOALLOC_MAX_BODIES) + "]");
    if (!out || (n > 0 && (!bodies_xyzm || !vel_xyzw))) return fail(NB_E_INVALID_ARG, "NULL argument");
    if (!aligned16(bodies_xyzm) || !aligned16(vel_xyzw))
        return fail(NB_E_INVALID_ARG, "body and velocity arrays must be 16-byte aligned (float4 per body)");
    if (reinterpret_cast<uintptr_t>(out) & 7u) return fail(NB_E_INVALID_ARG, "out must be 8-byte aligned");
    if (law != NB_LAW_PLUMMER && law != NB_LAW_REFERENCE_UNIT) return fail(NB_E_INVALID_ARG, "bad law");
    if (!std::isfinite(eps) || !std::isfinite(G) || eps < 0.0 || (law == NB_LAW_PLUMMER && !(eps > 0.0)))
        return fail(NB_E_INVALID_ARG, "eps must be finite and > 0 (>= 0 for the unit law); G finite");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(NB_E_NO_DEVICE, "no HIP device");
    double* partial = nullptr;
    nb_status st = scratch_ptr(&partial);
    if (st) return st;
    // n == 0: only k_energy_final runs, over no partials, and writes zeros
    nb::launch_energy(static_cast<hipStream_t>(stream), n, reinterpret_cast<const float4*>(bodies_xyzm),
                      reinterpret_cast<const float4*>(vel_xyzw), law, eps, G, partial, out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(NB_E_HIP, std::string("energy launch: ") + hipGetErrorString(e));
    return NB_OK;
}
