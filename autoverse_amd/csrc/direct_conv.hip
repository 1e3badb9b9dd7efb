// direct_conv.hip -- the convolution stage for grid sides the FFT passes do not cover.
//
// The reference convolves by FFT for any side (rustfft plans every length, fft.rs:231-237).
// The LDS FFT passes here cover sides 2^a * 3^b up to 4096; for any other side the periodic
// convolution of each kernel's source channel with its ring kernel (kernel.rs:32-117,
// kernel3d.rs:32-134) is summed directly, tap by tap, which is the same circular convolution
// (the kernel is point-symmetric, so correlation and convolution agree) to rounding.  The growth
// and the weighted accumulation in kernel order (propagator.rs:270-279, growth.rs:44-59) follow
// in the same kernel.
//
// Layout.  One workgroup owns a tile of destinations (2D: 32 x 32, 3D: 16 x 16 x 4, or 8 x 8 x 4
// when the wider tile's halo does not fit in LDS) and, per
// kernel, stages its source channel over the tile plus the kernel's halo in LDS.  Each thread
// computes a column of 4 destinations along the last axis (y in 2D, z in 3D): for every tap
// position in the other axes it slides a 4 + 1 value window along that axis, one LDS read per
// tap, and only over the range where the kernel's column is nonzero (the ring kernel vanishes
// outside its disk; the taps stay in kernel raster order inside), summing each column first.
// Every (tile, kernel) pair is a workgroup of its own, writing the kernel's convolution to
// conv[k]; k_direct_growth then forms U[t] = sum of w G(conv_k) in kernel order.  The embedded-
// parameter step takes the convolutions alone.
#include <math.h>

#include <algorithm>

#include "hp_common.hpp"

namespace hp {

namespace {

constexpr int kDcOut = 4;  // destinations per thread along the last axis

// One kernel's growth term w G(c) (growth.rs:51-57), as the LDS FFT path computes it.
__device__ __forceinline__ float dc_growth(float c, const DirectKernel& k) {
#pragma clang fp contract(off)
    const float d = c - k.mu;
    const float g = 2.0f * expf(-d * d * k.inv) - 1.0f;
    return k.weight * g;
}

__device__ __forceinline__ int dc_wrap(int c, int n) {
    int r = c % n;
    return r < 0 ? r + n : r;
}

// DIM 2: tile TX x (TY * kDcOut) with threads TX x TY; DIM 3: tile TX x TY x kDcOut.  Workgroup
// (tile, kernel k = blockIdx.y) writes kernel k's convolution over its tile into conv + k*N.
template <int DIM, int TX, int TY>
__global__ __launch_bounds__(TX * TY) void k_direct(const float* __restrict__ A, float* __restrict__ conv,
                                                    const DirectKernel* __restrict__ kd,
                                                    const float* __restrict__ taps, const int2* __restrict__ cols,
                                                    int W, int H, int D, int tiles_x, int tiles_y) {
    extern __shared__ float lds[];
    constexpr int OX = TX, OY = DIM == 2 ? TY * kDcOut : TY, OZ = DIM == 2 ? 1 : kDcOut;
    const long long N = (long long)W * H * D;
    const int bx = blockIdx.x % tiles_x, rest = blockIdx.x / tiles_x;
    const int by = rest % tiles_y, bz = rest / tiles_y;
    const int x0 = bx * OX, y0 = by * OY, z0 = bz * OZ;
    const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
    // this thread's destinations: (x, y, z + j) in 3D, (x, y + j) in 2D, j < kDcOut
    const int x = x0 + tx;
    const int y = DIM == 2 ? y0 + ty * kDcOut : y0 + ty;
    const int z = z0;
    const int k = blockIdx.y;
    const DirectKernel kk = kd[k];
    const int R = kk.R, S = 2 * R + 1;
    const int LX = OX + 2 * R, LY = OY + 2 * R, LZ = DIM == 2 ? 1 : OZ + 2 * R;
    const float* src = A + (long long)kk.src * N;
    const int total = LX * LY * LZ;
    for (int i = threadIdx.x; i < total; i += TX * TY) {
        const int lx = i % LX, r2 = i / LX, ly = r2 % LY, lz = r2 / LY;
        const int gx = dc_wrap(x0 - R + lx, W), gy = dc_wrap(y0 - R + ly, H);
        const int gz = DIM == 2 ? 0 : dc_wrap(z0 - R + lz, D);
        lds[i] = src[((long long)gz * H + gy) * W + gx];
    }
    __syncthreads();
    float acc[kDcOut];
#pragma unroll
    for (int j = 0; j < kDcOut; ++j) acc[j] = 0.0f;
    const float* kt = taps + kk.tap_off;
    const int2* kc = cols + kk.col_off;
    // the sliding axis: y in 2D (stride LX), z in 3D (stride LX * LY)
    const int sstride = DIM == 2 ? LX : LX * LY;
    const int ncol = DIM == 2 ? S : S * S;  // tap columns: dx (2D) or (dy, dx) (3D)
    for (int cidx = 0; cidx < ncol; ++cidx) {
        const int2 rg = kc[cidx];  // nonzero range [lo, hi] along the sliding axis
        if (rg.x > rg.y) continue;
        const int dxo = DIM == 2 ? cidx : cidx % S, dyo = DIM == 2 ? 0 : cidx / S;
        // LDS index of (tx + dxo, ty + dyo [2D: ty * kDcOut], slide = rg.x)
        const int base = DIM == 2 ? (ty * kDcOut + rg.x) * LX + tx + dxo : rg.x * sstride + (ty + dyo) * LX + tx + dxo;
        // taps of this column along the sliding axis: kt[(slide * S + dy) * S + dx] in 3D raster
        // order, kt[slide * S + dx] in 2D
        const int tstep = DIM == 2 ? S : S * S;
        const float* tp = kt + (DIM == 2 ? rg.x * S + cidx : rg.x * S * S + cidx);
        float win[kDcOut], cacc[kDcOut];
#pragma unroll
        for (int j = 0; j < kDcOut - 1; ++j) win[j] = lds[base + j * sstride];
#pragma unroll
        for (int j = 0; j < kDcOut; ++j) cacc[j] = 0.0f;
        for (int sl = rg.x; sl <= rg.y; ++sl) {
            win[kDcOut - 1] = lds[base + (sl - rg.x + kDcOut - 1) * sstride];
            const float t = *tp;
            tp += tstep;
#pragma unroll
            for (int j = 0; j < kDcOut; ++j) cacc[j] = fmaf(win[j], t, cacc[j]);
#pragma unroll
            for (int j = 0; j < kDcOut - 1; ++j) win[j] = win[j + 1];
        }
        // column sums first, then their total: the rounding error grows with the square roots of
        // the column length and of the column count, not of the tap count (9261 taps in 3D at R 10)
#pragma unroll
        for (int j = 0; j < kDcOut; ++j) acc[j] += cacc[j];
    }
    if (x >= W) return;
    for (int j = 0; j < kDcOut; ++j) {
        const int yy = DIM == 2 ? y + j : y, zz = DIM == 2 ? 0 : z + j;
        if (yy < H && zz < D) conv[(long long)k * N + ((long long)zz * H + yy) * W + x] = acc[j];
    }
}

// U[t] = sum over the kernels whose target is t, in kernel order, of w G(conv_k) (0 for a channel
// no kernel targets).
__global__ __launch_bounds__(256) void k_direct_growth(const float* __restrict__ conv,
                                                       const DirectKernel* __restrict__ kd, float* __restrict__ U,
                                                       int K, int C, long long N) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        for (int t = 0; t < C; ++t) {
            float aff = 0.0f;
            for (int k = 0; k < K; ++k)
                if (kd[k].tgt == t) aff += dc_growth(conv[(long long)k * N + i], kd[k]);
            U[(long long)t * N + i] = aff;
        }
    }
}

template <int DIM, int TX, int TY>
size_t dc_lds_bytes(int Rmax) {
    constexpr int OX = TX, OY = DIM == 2 ? TY * kDcOut : TY, OZ = DIM == 2 ? 1 : kDcOut;
    const size_t lz = DIM == 2 ? 1 : (size_t)(OZ + 2 * Rmax);
    return (size_t)(OX + 2 * Rmax) * (OY + 2 * Rmax) * lz * sizeof(float);
}

constexpr size_t kDcMaxLds = 160 * 1024;

template <int DIM, int TX, int TY>
void dc_launch(hipStream_t st, size_t lds, const float* A, float* conv, const DirectKernel* kd, const float* taps,
               const int2* cols, int K, const Geo& g) {
    constexpr int OY = DIM == 2 ? TY * kDcOut : TY, OZ = DIM == 2 ? 1 : kDcOut;
    const int tx = (g.W + TX - 1) / TX, ty = (g.H + OY - 1) / OY, tz = DIM == 2 ? 1 : (g.D + OZ - 1) / OZ;
    (void)hipFuncSetAttribute((const void*)k_direct<DIM, TX, TY>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL((k_direct<DIM, TX, TY>), dim3(tx * ty * tz, K), dim3(TX * TY), lds, st, A, conv, kd, taps,
                       cols, g.W, g.H, DIM == 2 ? 1 : g.D, tx, ty);
}

}  // namespace

// The largest kernel radius the direct path takes for a grid of dimension dim.
int direct_max_radius(int dim) {
    int r = 0;
    while ((dim == 2 ? dc_lds_bytes<2, 32, 8>(r + 1) : dc_lds_bytes<3, 8, 8>(r + 1)) <= kDcMaxLds) ++r;
    return r;
}

bool launch_direct_affinity(hipStream_t st, const float* A, float* U, float* conv, const DirectKernel* kd,
                            const float* taps, const int2* cols, int K, int C, const Geo& g, int Rmax) {
    if (K > 0) {
        if (g.D > 1) {
            // 16 x 16 x 4 destinations per workgroup; 8 x 8 x 4 when the wider tile's halo does not fit
            const size_t big = dc_lds_bytes<3, 16, 16>(Rmax), small = dc_lds_bytes<3, 8, 8>(Rmax);
            if (big <= kDcMaxLds) dc_launch<3, 16, 16>(st, big, A, conv, kd, taps, cols, K, g);
            else if (small <= kDcMaxLds) dc_launch<3, 8, 8>(st, small, A, conv, kd, taps, cols, K, g);
            else return false;
        } else {
            const size_t lds = dc_lds_bytes<2, 32, 8>(Rmax);
            if (lds > kDcMaxLds) return false;
            dc_launch<2, 32, 8>(st, lds, A, conv, kd, taps, cols, K, g);
        }
    }
    if (U) {
        const int blocks = (int)std::min<long long>((g.N + 255) / 256, 8192);
        hipLaunchKernelGGL(k_direct_growth, dim3(blocks), dim3(256), 0, st, conv, kd, U, K, C, g.N);
    }
    return true;
}

}  // namespace hp
