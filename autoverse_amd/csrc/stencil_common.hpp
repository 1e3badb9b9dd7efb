// stencil_common.hpp -- the reference's f32 stencil arithmetic shared by the forward
// stencil kernels and their adjoints (both compiled with -ffp-contract=off, so that the
// adjoint recomputes exactly the forward's flow and destinations).
#pragma once

#include <math.h>

#include "hp_common.hpp"

namespace hp {

__device__ __forceinline__ int wrapi(int c, int n) {
    int r = c % n;
    return r < 0 ? r + n : r;
}

// Wrap for coordinates a few grid lengths out at most (no integer division).
__device__ __forceinline__ int wrap_near(int c, int n) {
    while (c < 0) c += n;
    while (c >= n) c -= n;
    return c;
}

// Wrap c in [-n, 3n) into [0, n): at most two steps, no loop and no division.
__device__ __forceinline__ int wrap3(int c, int n) {
    c = c < 0 ? c + n : c;
    c = c >= n ? c - n : c;
    return c >= n ? c - n : c;
}

// Wrap c into [0, n).  `multi` is uniform over the workgroup: false when the caller knows
// c lies in [-n, 2n) (one conditional step, no loop and no division), true for grids so
// small that a tile window spans them more than once.
__device__ __forceinline__ int wrap_tile(int c, int n, bool multi) {
    if (multi) return wrapi(c, n);
    c = c < 0 ? c + n : c;
    return c >= n ? c - n : c;
}

// The reference's source test (reintegration.rs:76, reintegration3d.rs:77): a mass with
// |m| < 1e-10 is skipped; NaN is not.
__device__ __forceinline__ bool active_mass(float m) { return !(fabsf(m) < 1e-10f); }

// Workgroup barrier that orders LDS only: this wave's LDS accesses complete before it, and the
// workgroup's LDS writes are visible after it.  Unlike __syncthreads() it does not wait for the
// wave's outstanding global loads and stores (a release fence at workgroup scope would).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// OR of v over the 64 lanes of the wave.
__device__ __forceinline__ unsigned wave_or(unsigned v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v |= (unsigned)__shfl_xor((int)v, o);
    return v;
}

// Rust `f as i32`: saturating, NaN -> 0.  Branch-free: the clamp keeps the conversion in range
// (2147483520 is the largest float below 2^31; v >= 2147483647.0f means v >= 2^31), a NaN
// clamps to -2^31 and is then replaced by 0.
__device__ __forceinline__ int sat_i32(float v) {
    const int t = (int)fminf(fmaxf(v, -2147483648.0f), 2147483520.0f);
    const int r = v >= 2147483647.0f ? 2147483647 : t;
    return v != v ? 0 : r;
}

// compute_alpha (flow.rs:14-32): integer fast paths for n = 1..4, else powf; clamp keeps NaN.
__device__ __forceinline__ float alpha_of(float mass, float beta_a, float n) {
    const float ratio = mass / beta_a;
    float p;
    if (n == 1.0f) p = ratio;
    else if (n == 2.0f) p = ratio * ratio;
    else if (n == 3.0f) p = ratio * ratio * ratio;
    else if (n == 4.0f) {
        const float r2 = ratio * ratio;
        p = r2 * r2;
    } else p = powf(ratio, n);
    if (p < 0.0f) p = 0.0f;
    if (p > 1.0f) p = 1.0f;
    return p;
}

// gradient.rs:99-102
__device__ __forceinline__ void sobel9(float tl, float tc, float tr, float ml, float mr, float bl, float bc,
                                       float br, float& gx, float& gy) {
    gx = (-tl + tr - 2.0f * ml + 2.0f * mr - bl + br) * 0.125f;
    gy = (-tl - 2.0f * tc - tr + bl + 2.0f * bc + br) * 0.125f;
}

// Overlap of cell [c0, c0 + 1) (c1 = c0 + 1 as f32) with [lo, hi] (reintegration.rs:137-144):
// max(min(c1, hi) - max(c0, lo), 0) with Rust max/min (a NaN operand is ignored).
__device__ __forceinline__ float overlap1(float c0, float c1, float lo, float hi) {
    return fmaxf(fminf(c1, hi) - fmaxf(c0, lo), 0.0f);
}

// The same overlap as clamp(hi, c0, c1) - clamp(lo, c0, c1), two v_med3_f32 and a subtract.
// For finite lo <= hi and c0 < c1 it equals overlap1 bit for bit: inside the cell both forms
// subtract the same operands, and a box that misses the cell clamps both ends to the same
// bound (+0) where overlap1 clamps a negative difference to +0.  A non-finite box may differ;
// the fused stencils flag such a source and the host discards the step.
__device__ __forceinline__ float overlap_clamped(float c0, float c1, float lo, float hi) {
    return __builtin_amdgcn_fmed3f(hi, c0, c1) - __builtin_amdgcn_fmed3f(lo, c0, c1);
}

}  // namespace hp
