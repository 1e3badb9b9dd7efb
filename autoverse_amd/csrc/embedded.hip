// embedded.hip -- the embedded-parameter Flow Lenia step on the device
// (EmbeddedPropagator::step, src/compute/embedded_propagator.rs:222-331).
//
// The per-cell parameters (CellParams: mu, sigma, weight, beta_a, n) live as five planes per
// channel, P[c][f][N], so every pass reads them coalesced.  The convolutions come from the FFT
// passes (the convolution of a channel with a kernel does not depend on the parameters; they
// enter the growth), and the kernels here do the rest:
//   k_emb_growth : affinity from every kernel's convolution with the target's per-cell growth
//                  (direct_convolution.rs:129-174);
//   k_emb_flow   : Sobel of the channel sum and of U, per-cell alpha and the flow
//                  (embedded_propagator.rs:251-302, 343-362), and the largest |F| per channel;
//   k_emb_advect_tile / k_emb_advect : the reference's destination gather with parameter
//                  mixing (param_advection.rs:83-166, embedding.rs:60-167), tiled through LDS
//                  for the usual radii, straight from global memory otherwise.
// Compiled with -ffp-contract=off: every f32 operation rounds as the reference's does.  The
// growth and softmax exponentials are evaluated in FP64 and rounded once (the reference calls
// libm's expf; the two agree to the last bit except in about 7 of 10000 arguments, by one ulp).
#include <math.h>

#include <algorithm>

#include "hp_common.hpp"
#include "stencil_common.hpp"

namespace hp {

namespace {

constexpr int kEmbThreads = 256;

__device__ __forceinline__ float exp_rn(float x) { return (float)exp((double)x); }

// f32 sum of the channels at one cell, in channel order ((0 + A_0) + A_1) + ...
__device__ __forceinline__ float emb_asum(const float* A, int C, long long N, long long i) {
    float s = 0.0f;
    for (int c = 0; c < C; ++c) s += A[(long long)c * N + i];
    return s;
}

// compute_distribution_overlap (param_advection.rs:318-348): no periodic images.
__device__ __forceinline__ float emb_overlap(float dest_x, float dest_y, float s, float cell_x, float cell_y) {
    const float xmin = dest_x - s, xmax = dest_x + s;
    const float ymin = dest_y - s, ymax = dest_y + s;
    const float cxm = cell_x + 1.0f, cym = cell_y + 1.0f;
    const float ow = fmaxf(fminf(xmax, cxm) - fmaxf(xmin, cell_x), 0.0f);
    const float oh = fmaxf(fminf(ymax, cym) - fmaxf(ymin, cell_y), 0.0f);
    const float area = ow * oh;
    const float total = (2.0f * s) * (2.0f * s);
    if (total > 1e-10f) return area / total;
    return (cell_x <= dest_x && dest_x < cxm && cell_y <= dest_y && dest_y < cym) ? 1.0f : 0.0f;
}

// AoS cells (5 floats each, the reference's Vec<CellParams>) of C channels <-> planes.
__global__ __launch_bounds__(kEmbThreads) void k_emb_split(const float* __restrict__ aos, float* __restrict__ planes,
                                                           long long N, long long total) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / N, i = t - c * N;
        const float* src = aos + t * 5;
#pragma unroll
        for (int f = 0; f < 5; ++f) planes[(c * 5 + f) * N + i] = src[f];
    }
}

__global__ __launch_bounds__(kEmbThreads) void k_emb_join(const float* __restrict__ planes, float* __restrict__ aos,
                                                          long long N, long long total) {
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const long long c = t / N, i = t - c * N;
        float* dst = aos + t * 5;
#pragma unroll
        for (int f = 0; f < 5; ++f) dst[f] = planes[(c * 5 + f) * N + i];
    }
}

// Affinity: per target channel t, the kernels whose target is t in config order, each adding
// weight * (2 exp(-diff diff / (2 sigma sigma)) - 1) with the cell's own parameters.
__global__ __launch_bounds__(kEmbThreads) void k_emb_growth(const float* __restrict__ conv, const float* __restrict__ P,
                                                            const int* __restrict__ ktarget, float* __restrict__ U,
                                                            int C, int K, long long N) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        for (int t = 0; t < C; ++t) {
            const float* pt = P + (long long)t * 5 * N;
            const float mu = pt[i], sg = pt[N + i], w = pt[2 * N + i];
            float aff = 0.0f;
            for (int k = 0; k < K; ++k) {
                if (ktarget[k] != t) continue;
                const float diff = conv[(long long)k * N + i] - mu;
                const float s2 = 2.0f * sg * sg;
                const float g = 2.0f * exp_rn(-diff * diff / s2) - 1.0f;
                aff += w * g;
            }
            U[(long long)t * N + i] = aff;
        }
    }
}

// Flows of channels c0 .. c0 + kFlowChunk - 1 from Sobel(channel sum) (once per cell), Sobel(U_c)
// and the cell's alpha; the largest |F| of each channel goes to maxbits[c] (non-negative floats
// order as their bits), one atomic per workgroup and channel.
constexpr int kFlowChunk = 8;
__global__ __launch_bounds__(kEmbThreads) void k_emb_flow(const float* __restrict__ A, const float* __restrict__ U,
                                                          const float* __restrict__ P, float* __restrict__ F,
                                                          unsigned* __restrict__ maxbits, int C, int c0, int W, int H,
                                                          long long N) {
    __shared__ float red[kEmbThreads / 64][kFlowChunk];
    const int nc = min(C - c0, kFlowChunk);
    float lmax[kFlowChunk];
#pragma unroll
    for (int k = 0; k < kFlowChunk; ++k) lmax[k] = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), y = (int)(i / W);
        const int xp = x == 0 ? W - 1 : x - 1, xn = x + 1 == W ? 0 : x + 1;
        const int yp = y == 0 ? H - 1 : y - 1, yn = y + 1 == H ? 0 : y + 1;
        const long long rp = (long long)yp * W, rc = (long long)y * W, rn = (long long)yn * W;
        float gax, gay;
        sobel9(emb_asum(A, C, N, rp + xp), emb_asum(A, C, N, rp + x), emb_asum(A, C, N, rp + xn),
               emb_asum(A, C, N, rc + xp), emb_asum(A, C, N, rc + xn), emb_asum(A, C, N, rn + xp),
               emb_asum(A, C, N, rn + x), emb_asum(A, C, N, rn + xn), gax, gay);
        const float asum = emb_asum(A, C, N, i);
#pragma unroll
        for (int k = 0; k < kFlowChunk; ++k) {
            if (k >= nc) break;
            const int c = c0 + k;
            const float* Uc = U + (long long)c * N;
            float gux, guy;
            sobel9(Uc[rp + xp], Uc[rp + x], Uc[rp + xn], Uc[rc + xp], Uc[rc + xn], Uc[rn + xp], Uc[rn + x],
                   Uc[rn + xn], gux, guy);
            const float al = alpha_of(asum, P[((long long)c * 5 + 3) * N + i], P[((long long)c * 5 + 4) * N + i]);
            const float oma = 1.0f - al;
            const float fx = oma * gux - al * gax;
            const float fy = oma * guy - al * gay;
            F[(long long)c * 2 * N + i] = fx;
            F[((long long)c * 2 + 1) * N + i] = fy;
            lmax[k] = fmaxf(lmax[k], sqrtf(fx * fx + fy * fy));  // f32::max: a NaN magnitude is passed over
        }
    }
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
#pragma unroll
    for (int k = 0; k < kFlowChunk; ++k) {
        float v = lmax[k];
        for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if (lane == 0) red[wave][k] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nc) {
        float v = red[0][threadIdx.x];
        for (int w = 1; w < kEmbThreads / 64; ++w) v = fmaxf(v, red[w][threadIdx.x]);
        atomicMax(maxbits + c0 + threadIdx.x, __float_as_uint(v));
    }
}

// compute_max_flow_magnitude (param_advection.rs:351-357) of given flows F[c][2][N].
__global__ __launch_bounds__(kEmbThreads) void k_emb_maxmag(const float* __restrict__ F, unsigned* __restrict__ maxbits,
                                                            long long N) {
    __shared__ float red[kEmbThreads];
    const int c = blockIdx.y;
    const float* fxv = F + (long long)c * 2 * N;
    const float* fyv = fxv + N;
    float local = 0.0f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x)
        local = fmaxf(local, sqrtf(fxv[i] * fxv[i] + fyv[i] * fyv[i]));
    red[threadIdx.x] = local;
    __syncthreads();
    for (int w = kEmbThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(maxbits + c, __float_as_uint(red[0]));
}

// The tiled gather stages a destination tile's source window in LDS: radii up to kTileR, sides of
// at least kTile cells, and boxes of area above the point threshold.  k_emb_advect covers the
// rest; both kernels are launched and one of them returns at once.
constexpr int kTile = 16, kTileR = 4, kRegion = kTile + 2 * kTileR, kCap = 10;
__device__ __forceinline__ bool tiled_gather(int R, int W, int H, float s) {
    return R <= kTileR && W >= kTile && H >= kTile && (2.0f * s) * (2.0f * s) > 1e-10f;
}

// Destination gather of channel blockIdx.z over one 16 x 16 tile (one destination per thread),
// the same arithmetic as k_emb_advect: each window source's box [x + dt F - s, x + dt F + s] and
// mass are staged once (the values every destination would compute), and each destination keeps
// its contributions in LDS (up to kCap; more are recomputed) so the softmax weights are formed
// once per contribution.  A tile whose window holds no active source writes mass 0.
__global__ __launch_bounds__(kTile * kTile) void k_emb_advect_tile(
    const float* __restrict__ A, const float* __restrict__ P, const float* __restrict__ F,
    const unsigned* __restrict__ maxbits, float* __restrict__ Aout, float* __restrict__ Pout, int W, int H,
    long long N, float dt, float s, int linear, float temperature, int rmax) {
    __shared__ float4 sbox[kRegion * kRegion];
    __shared__ float smass[kRegion * kRegion];
    __shared__ float spar[5][kRegion * kRegion];
    __shared__ float lc[kCap][kTile * kTile];
    __shared__ short li[kCap][kTile * kTile];
    const int c = blockIdx.z;
    const float mag = __uint_as_float(maxbits[c]);
    const int r0 = sat_i32(ceilf(mag * dt + s));
    if (r0 >= rmax) return;  // k_emb_advect raises the flag
    const int R = r0 + 1;
    if (!tiled_gather(R, W, H, s)) return;
    const float* Ac = A + (long long)c * N;
    const float* fxv = F + (long long)c * 2 * N;
    const float* fyv = fxv + N;
    const float* Pc = P + (long long)c * 5 * N;
    const int tx0 = blockIdx.x * kTile, ty0 = blockIdx.y * kTile;
    const int RW = kTile + 2 * R;
    const int tid = threadIdx.x;
    bool any = false;
    for (int e = tid; e < RW * RW; e += kTile * kTile) {
        const int ry = e / RW, rx = e - ry * RW;
        const int sx = wrapi(tx0 - R + rx, W), sy = wrapi(ty0 - R + ry, H);
        const long long si = (long long)sy * W + sx;
        const float m = Ac[si];
        float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float mm = 0.0f;  // 0 marks a skipped source (an active mass has |m| >= 1e-10)
        if (!(fabsf(m) < 1e-10f)) {
            const float dfx = (float)sx + dt * fxv[si];
            const float dfy = (float)sy + dt * fyv[si];
            b = make_float4(dfx - s, dfx + s, dfy - s, dfy + s);
            mm = m;
            any = true;
#pragma unroll
            for (int f = 0; f < 5; ++f) spar[f][e] = Pc[f * N + si];
        }
        sbox[e] = b;
        smass[e] = mm;
    }
    any = __syncthreads_or(any);
    const int lx = tid % kTile, ly = tid / kTile;
    const int x = tx0 + lx, y = ty0 + ly;
    if (x >= W || y >= H) return;
    const long long di = (long long)y * W + x;
    float* Ao = Aout + (long long)c * N;
    if (!any) {
        Ao[di] = 0.0f;
        return;
    }
    float* Po = Pout + (long long)c * 5 * N;
    const float cx = (float)x, cy = (float)y, cxm = cx + 1.0f, cym = cy + 1.0f;
    const float ta = (2.0f * s) * (2.0f * s);
    // a power-of-two area divides exactly as a multiplication by its reciprocal (both round the
    // same exact quotient)
    int ex;
    const bool pow2 = frexpf(ta, &ex) == 0.5f;
    const float inv_ta = 1.0f / ta;
    // the contribution of window entry e to this destination; false when it contributes nothing.
    // An empty overlap along either axis makes the reference's area / total zero, so those
    // entries are passed over before the division.
    auto contribution = [&](int e, float& cm) -> bool {
        const float m = smass[e];
        if (m == 0.0f) return false;
        const float4 b = sbox[e];
        const float oh = fmaxf(fminf(b.w, cym) - fmaxf(b.z, cy), 0.0f);
        if (!(oh > 0.0f)) return false;
        const float ow = fmaxf(fminf(b.y, cxm) - fmaxf(b.x, cx), 0.0f);
        if (!(ow > 0.0f)) return false;
        const float area = ow * oh;
        const float ov = pow2 ? area * inv_ta : area / ta;
        if (!(ov > 0.0f)) return false;
        cm = m * ov;
        return true;
    };
    float total = 0.0f, mx = -INFINITY;
    int count = 0, first = 0;
    for (int dy = -R; dy <= R; ++dy) {
        const int row = (ly + R + dy) * RW + lx + R;
        for (int dx = -R; dx <= R; ++dx) {
            float cm;
            if (!contribution(row + dx, cm)) continue;
            total += cm;
            mx = fmaxf(mx, cm);
            if (count < kCap) {
                lc[count][tid] = cm;
                li[count][tid] = (short)(row + dx);
            }
            if (count == 0) first = row + dx;
            ++count;
        }
    }
    Ao[di] = total;
    if (count == 0 || !(total > 1e-10f)) return;
    float acc[5];
    if (count == 1) {
#pragma unroll
        for (int f = 0; f < 5; ++f) Po[f * N + di] = spar[f][first];
        return;
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) acc[f] = 0.0f;
    const bool cached = count <= kCap;
    if (linear) {
        if (cached) {
            for (int k = 0; k < count; ++k) {
                const float w = lc[k][tid] / total;
                const int e = li[k][tid];
#pragma unroll
                for (int f = 0; f < 5; ++f) acc[f] += spar[f][e] * w;
            }
        } else {
            for (int dy = -R; dy <= R; ++dy) {
                const int row = (ly + R + dy) * RW + lx + R;
                for (int dx = -R; dx <= R; ++dx) {
                    float cm;
                    if (!contribution(row + dx, cm)) continue;
                    const float w = cm / total;
#pragma unroll
                    for (int f = 0; f < 5; ++f) acc[f] += spar[f][row + dx] * w;
                }
            }
        }
    } else {
        float wsum = 0.0f;
        if (cached) {
            for (int k = 0; k < count; ++k) {
                const float e = exp_rn((lc[k][tid] - mx) / temperature);
                lc[k][tid] = e;
                wsum += e;
            }
        } else {
            for (int dy = -R; dy <= R; ++dy) {
                const int row = (ly + R + dy) * RW + lx + R;
                for (int dx = -R; dx <= R; ++dx) {
                    float cm;
                    if (!contribution(row + dx, cm)) continue;
                    wsum += exp_rn((cm - mx) / temperature);
                }
            }
        }
        const bool uniform = !(wsum > 0.0f);
        const float u = 1.0f / (float)count;
        if (cached) {
            for (int k = 0; k < count; ++k) {
                const float w = uniform ? u : lc[k][tid] / wsum;
                const int e = li[k][tid];
#pragma unroll
                for (int f = 0; f < 5; ++f) acc[f] += spar[f][e] * w;
            }
        } else {
            for (int dy = -R; dy <= R; ++dy) {
                const int row = (ly + R + dy) * RW + lx + R;
                for (int dx = -R; dx <= R; ++dx) {
                    float cm;
                    if (!contribution(row + dx, cm)) continue;
                    const float w = uniform ? u : exp_rn((cm - mx) / temperature) / wsum;
#pragma unroll
                    for (int f = 0; f < 5; ++f) acc[f] += spar[f][row + dx] * w;
                }
            }
        }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) Po[f * N + di] = acc[f];
}

// Destination gather of channel blockIdx.y with parameter mixing.  R = ceil(max|F| dt + s) + 1
// from maxbits[c]; a radius above rmax sets err bit 0 and writes nothing (the host refuses the
// step).  Pout keeps its content at destinations that receive no mix.
__global__ __launch_bounds__(kEmbThreads) void k_emb_advect(const float* __restrict__ A, const float* __restrict__ P,
                                                            const float* __restrict__ F,
                                                            const unsigned* __restrict__ maxbits,
                                                            float* __restrict__ Aout, float* __restrict__ Pout,
                                                            int W, int H, long long N, float dt, float s, int linear,
                                                            float temperature, int rmax, int* __restrict__ err) {
    const int c = blockIdx.y;
    const float mag = __uint_as_float(maxbits[c]);
    const int r0 = sat_i32(ceilf(mag * dt + s));
    if (r0 >= rmax) {
        if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(err, 1);
        return;
    }
    const int R = r0 + 1;
    if (tiled_gather(R, W, H, s)) return;  // k_emb_advect_tile covers this step
    const float* Ac = A + (long long)c * N;
    const float* fxv = F + (long long)c * 2 * N;
    const float* fyv = fxv + N;
    const float* Pc = P + (long long)c * 5 * N;
    float* Po = Pout + (long long)c * 5 * N;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), y = (int)(i / W);
        const float cx = (float)x, cy = (float)y;
        // the contribution of the source at offset (dx, dy); false when it contributes nothing
        auto contribution = [&](int dx, int dy, long long& si, float& cm) -> bool {
            const int sx = wrapi(x + dx, W), sy = wrapi(y + dy, H);
            si = (long long)sy * W + sx;
            const float m = Ac[si];
            if (fabsf(m) < 1e-10f) return false;
            const float ov = emb_overlap((float)sx + dt * fxv[si], (float)sy + dt * fyv[si], s, cx, cy);
            if (!(ov > 0.0f)) return false;
            cm = m * ov;
            return true;
        };
        float total = 0.0f, mx = -INFINITY;
        int count = 0;
        long long first = 0;
        for (int dy = -R; dy <= R; ++dy)
            for (int dx = -R; dx <= R; ++dx) {
                long long si;
                float cm;
                if (!contribution(dx, dy, si, cm)) continue;
                total += cm;
                mx = fmaxf(mx, cm);
                if (count == 0) first = si;
                ++count;
            }
        Aout[(long long)c * N + i] = total;
        if (count == 0 || !(total > 1e-10f)) continue;
        float acc[5];
        if (count == 1) {
#pragma unroll
            for (int f = 0; f < 5; ++f) acc[f] = Pc[f * N + first];
        } else if (linear) {
            // mix_linear (embedding.rs:131-167): w = m / total, the same total the loop above formed
#pragma unroll
            for (int f = 0; f < 5; ++f) acc[f] = 0.0f;
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    long long si;
                    float cm;
                    if (!contribution(dx, dy, si, cm)) continue;
                    const float w = cm / total;
#pragma unroll
                    for (int f = 0; f < 5; ++f) acc[f] += Pc[f * N + si] * w;
                }
        } else {
            // mix_softmax (embedding.rs:60-126): weights exp((m - max) / T), normalised
            float wsum = 0.0f;
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    long long si;
                    float cm;
                    if (!contribution(dx, dy, si, cm)) continue;
                    wsum += exp_rn((cm - mx) / temperature);
                }
            const bool uniform = !(wsum > 0.0f);
            const float u = 1.0f / (float)count;
#pragma unroll
            for (int f = 0; f < 5; ++f) acc[f] = 0.0f;
            for (int dy = -R; dy <= R; ++dy)
                for (int dx = -R; dx <= R; ++dx) {
                    long long si;
                    float cm;
                    if (!contribution(dx, dy, si, cm)) continue;
                    const float w = uniform ? u : exp_rn((cm - mx) / temperature) / wsum;
#pragma unroll
                    for (int f = 0; f < 5; ++f) acc[f] += Pc[f * N + si] * w;
                }
        }
#pragma unroll
        for (int f = 0; f < 5; ++f) Po[f * N + i] = acc[f];
    }
}

int grid_for(long long work) { return (int)std::min<long long>((work + kEmbThreads - 1) / kEmbThreads, 8192); }

}  // namespace

void launch_emb_split(hipStream_t st, const float* aos, float* planes, int C, long long N) {
    const long long total = (long long)C * N;
    hipLaunchKernelGGL(k_emb_split, dim3(grid_for(total)), dim3(kEmbThreads), 0, st, aos, planes, N, total);
}

void launch_emb_join(hipStream_t st, const float* planes, float* aos, int C, long long N) {
    const long long total = (long long)C * N;
    hipLaunchKernelGGL(k_emb_join, dim3(grid_for(total)), dim3(kEmbThreads), 0, st, planes, aos, N, total);
}

void launch_emb_growth(hipStream_t st, const float* conv, const float* P, const int* ktarget, float* U, int C, int K,
                       long long N) {
    hipLaunchKernelGGL(k_emb_growth, dim3(grid_for(N)), dim3(kEmbThreads), 0, st, conv, P, ktarget, U, C, K, N);
}

void launch_emb_flow(hipStream_t st, const float* A, const float* U, const float* P, float* F, unsigned* maxbits,
                     int C, int W, int H) {
    const long long N = (long long)W * H;
    const int blocks = (int)std::min<long long>((N + kEmbThreads - 1) / kEmbThreads, 2048);
    for (int c0 = 0; c0 < C; c0 += kFlowChunk)
        hipLaunchKernelGGL(k_emb_flow, dim3(blocks), dim3(kEmbThreads), 0, st, A, U, P, F, maxbits, C, c0, W, H, N);
}

void launch_emb_maxmag(hipStream_t st, const float* F, unsigned* maxbits, int C, long long N) {
    hipLaunchKernelGGL(k_emb_maxmag, dim3(grid_for(N), C), dim3(kEmbThreads), 0, st, F, maxbits, N);
}

void launch_emb_advect(hipStream_t st, const float* A, const float* P, const float* F, const unsigned* maxbits,
                       float* Aout, float* Pout, int C, int W, int H, float dt, float s, bool linear,
                       float temperature, int rmax, int* err) {
    const long long N = (long long)W * H;
    hipLaunchKernelGGL(k_emb_advect_tile, dim3((W + kTile - 1) / kTile, (H + kTile - 1) / kTile, C),
                       dim3(kTile * kTile), 0, st, A, P, F, maxbits, Aout, Pout, W, H, N, dt, s, linear ? 1 : 0,
                       temperature, rmax);
    hipLaunchKernelGGL(k_emb_advect, dim3(grid_for(N), C), dim3(kEmbThreads), 0, st, A, P, F, maxbits, Aout, Pout, W,
                       H, N, dt, s, linear ? 1 : 0, temperature, rmax, err);
}

}  // namespace hp
