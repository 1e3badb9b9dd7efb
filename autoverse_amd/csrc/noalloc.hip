// noalloc.hip -- reintegration on caller device buffers without any device allocation.
//
// hp_advect_mass_2d_noalloc is hp_advect_mass_2d (advect_mass_into, reintegration.rs:60-88) for
// callers that cannot allocate device memory (code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc): inputs and output are the caller's device buffers,
// and the per-source destination boxes and the flags the two kernels exchange live in one
// file-scope device array.  The kernels are the ones hp_advect_mass_2d runs, so the results
// are the same bit for bit.
//
// The array exists once per GPU for the life of the process.  Its address on a device is
// looked up with hipGetSymbolAddress, with that device current, on the first call there and
// kept; the array's name itself is never handed to a kernel or a copy (in host code it is not
// a device address).  Every call on a GPU shares that GPU's array: calls that overlap on one
// GPU (from several host threads, or on several streams) race.  The flags are cleared by a
// kernel at the start of every call; the boxes are fully rewritten by each call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "../../include/flowlenia_hip.h"
#include "hp_common.hpp"
#include "host_internal.hpp"

namespace {

// Largest grid the scratch covers: 2^20 cells (1024 x 1024).  Layout: box_lo[kMaxCells],
// box_hi[kMaxCells], then one float4 whose first two ints are the flags (needed gather
// radius, non-finite bit).  32 MiB + 16 bytes per GPU.
constexpr long long kMaxCells = 1ll << 20;
constexpr long long kScratchVec = 2 * kMaxCells + 1;

static __device__ __align__(256) float4 scratch[kScratchVec];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float4* g_scratch_ptr[kMaxDevices] = {};

// Device address of this GPU's scratch (the current device must be `device`).
hp_status scratch_on(int device, float4** out) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            hp::set_global_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return HP_ERR_HIP;
        }
        g_scratch_ptr[device] = (float4*)p;
    }
    *out = g_scratch_ptr[device];
    return HP_OK;
}

#define NA_TRY(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            hp::set_global_error(std::string(#expr) + ": " + hipGetErrorString(e_));    \
            return HP_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

hp_status advect_noalloc(int device, const float* cur, const float* fx, const float* fy, float* next, uint32_t W,
                         uint32_t H, float dt, float s) {
    if (!cur || !fx || !fy || !next) return HP_ERR_NULL;
    if (W == 0 || H == 0) return HP_ERR_INVALID_DIMENSIONS;
    if ((long long)W * (long long)H > kMaxCells) {
        hp::set_global_error("grid of " + std::to_string((long long)W * H) +
                             " cells exceeds the no-allocation scratch limit of " + std::to_string(kMaxCells) +
                             " cells; use hp_advect_mass_2d");
        return HP_ERR_UNSUPPORTED_SIZE;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev ||
        device >= kMaxDevices) {
        hp::set_global_error("no HIP device");
        return HP_ERR_NO_DEVICE;
    }
    NA_TRY(hipSetDevice(device));
    float4* base = nullptr;
    hp_status st = scratch_on(device, &base);
    if (st != HP_OK) return st;
    float4* box_lo = base;
    float4* box_hi = base + kMaxCells;
    int* flags = (int*)(base + 2 * kMaxCells);
    hp::Geo g{};
    g.W = (int)W;
    g.H = (int)H;
    g.D = 1;
    g.N = (long long)W * H;
    // the stream is the null stream, as hp_advect_mass_2d's; flags are cleared by a kernel
    hp::launch_clear_ints(nullptr, flags, 4);
    hp::launch_boxes_from_flow(nullptr, cur, fx, fy, nullptr, g, dt, s, box_lo, box_hi, flags);
    NA_TRY(hipGetLastError());
    int h[2] = {0, 0};
    NA_TRY(hipMemcpy(h, flags, sizeof(h), hipMemcpyDeviceToHost));
    if (h[1]) {
        hp::set_global_error("flow or mass holds NaN/Inf");
        return HP_ERR_NONFINITE;
    }
    const int R = std::max(h[0], 1);
    auto win = [](int R_, int n_) { return (double)std::min(2 * R_ + 1, n_); };
    if (win(R, g.W) * win(R, g.H) * (double)g.N > 2e10) {
        hp::set_global_error("flow displacement beyond the general gather's work bound");
        return HP_ERR_UNSUPPORTED_SIZE;
    }
    hp::launch_gather_general(nullptr, box_lo, box_hi, next, 1, g, s, R);
    NA_TRY(hipGetLastError());
    NA_TRY(hipStreamSynchronize(nullptr));
    return HP_OK;
}

}  // namespace

extern "C" hp_status hp_advect_mass_2d_noalloc(int device, const float* current, const float* flow_x,
                                               const float* flow_y, float* next, uint32_t width, uint32_t height,
                                               float dt, float distribution_size) {
    try {
        return advect_noalloc(device, current, flow_x, flow_y, next, width, height, dt, distribution_size);
    } catch (...) {
        return HP_ERR_HIP;
    }
}
