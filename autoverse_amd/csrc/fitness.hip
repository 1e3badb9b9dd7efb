// fitness.hip -- per-sample statistics of a population's channel sums, on the device.
//
// FitnessEvaluator::evaluate (src/compute/evolution/fitness.rs:26-86) records, at every sample,
// the channel sum's centre of mass and radius over cells with m > 0 (fitness.rs:196-244), its
// mass, its largest value, its count of cells above 1e-6 (fitness.rs:146-160), and keeps the
// channel sum itself for the similarity metrics (fitness.rs:353-366); the state metrics read the
// final state's mean, variance, neighbour differences, peak and negative cells
// (fitness.rs:369-413, 486-511).  Here every one of those sums is one FP64 reduction over the
// member's cells, computed in a fixed order (per work item a block tree, then the items of a
// member in index order), so results are reproducible run to run.  The host turns the sums into
// the reference's f32 metric values.
//
// Compiled with -ffp-contract=off: the channel sum is formed as the reference forms it
// (0 + A_0 + A_1 + ...).
#include <math.h>

#include "hp_common.hpp"

namespace hp {

namespace {

constexpr int kFitThreads = 256;

// Block reduction of one double; every thread gets the result.  Two barriers.
__device__ double block_sum(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = kFitThreads / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double* red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int w = kFitThreads / 2; w > 0; w >>= 1) {
        if (t < w) red[t] = fmax(red[t], red[t + w]);
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ float chan_sum(const float* A, int C, long long N, long long i) {
    float v = 0.0f;
    for (int c = 0; c < C; ++c) v += A[(long long)c * N + i];
    return v;
}

// item = (member, part, nparts, -): cells [part * chunk, (part + 1) * chunk) of the member.
// part_out[item * kFitStats + k] receives the item's sums; snap (if the member's base is not
// null) receives the channel sum at slot `slot`.
__global__ __launch_bounds__(kFitThreads) void kb_fit_stats(const MemberDesc* __restrict__ mem,
                                                            const int4* __restrict__ items, int cur, int with_grad,
                                                            float* const* __restrict__ snap_base, int slot,
                                                            double* __restrict__ part_out) {
    __shared__ double red[kFitThreads];
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const float* A = m.A[cur];
    const int W = m.g.W, H = m.g.H;
    const long long N = (long long)W * H;
    const long long chunk = (N + it.z - 1) / it.z;
    const long long i0 = (long long)it.y * chunk, i1 = i0 + chunk < N ? i0 + chunk : N;
    float* snap = snap_base ? snap_base[it.x] : nullptr;
    if (snap) snap += (long long)slot * N;
    double s_all = 0.0, s_pos = 0.0, mx = 0.0, my = 0.0, m2 = 0.0, vmax = 0.0, act = 0.0, neg = 0.0, q = 0.0,
           grad = 0.0;
    for (long long i = i0 + threadIdx.x; i < i1; i += kFitThreads) {
        const float v = chan_sum(A, m.C, m.g.N, i);
        if (snap) snap[i] = v;
        const int x = (int)(i % W), y = (int)(i / W);
        const double d = (double)v;
        s_all += d;
        q += d * d;
        if (v > 0.0f) {
            s_pos += d;
            mx += (double)x * d;
            my += (double)y * d;
            m2 += ((double)x * (double)x + (double)y * (double)y) * d;
        }
        if (v == v && d > vmax) vmax = d;  // fold(0.0, f32::max): NaN is ignored
        if (v > 1e-6f) act += 1.0;
        if (v < 0.0f) neg += 1.0;
        if (with_grad) {  // fitness.rs:380-393: left and upper neighbours, no wrap
            if (x > 0) grad += (double)fabsf(v - chan_sum(A, m.C, m.g.N, i - 1));
            if (y > 0) grad += (double)fabsf(v - chan_sum(A, m.C, m.g.N, i - W));
        }
    }
    double* o = part_out + (long long)blockIdx.x * kFitStats;
    const double r0 = block_sum(s_all, red), r1 = block_sum(s_pos, red), r2 = block_sum(mx, red);
    const double r3 = block_sum(my, red), r4 = block_sum(m2, red), r5 = block_max(vmax, red);
    const double r6 = block_sum(act, red), r7 = block_sum(neg, red), r8 = block_sum(q, red);
    const double r9 = block_sum(grad, red);
    if (threadIdx.x == 0) {
        o[0] = r0; o[1] = r1; o[2] = r2; o[3] = r3; o[4] = r4;
        o[5] = r5; o[6] = r6; o[7] = r7; o[8] = r8; o[9] = r9;
    }
}

// One thread per (member, statistic): the member's items in index order.  first[m] is the
// index of member m's first item; its count is first[m + 1] - first[m].
__global__ void kb_fit_reduce(const int* __restrict__ first, int nmem, const double* __restrict__ part,
                              double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nmem * kFitStats) return;
    const int mm = t / kFitStats, k = t - mm * kFitStats;
    double r = 0.0;
    for (int j = first[mm]; j < first[mm + 1]; ++j) {
        const double v = part[(long long)j * kFitStats + k];
        r = k == 5 ? fmax(r, v) : r + v;
    }
    out[(long long)mm * kFitStats + k] = r;
}

// pair = (member, slot a, slot b, -): FP64 dot product of two snapshots of the member.
__global__ __launch_bounds__(kFitThreads) void kb_fit_dots(const MemberDesc* __restrict__ mem,
                                                           const int4* __restrict__ pairs,
                                                           float* const* __restrict__ snap_base,
                                                           double* __restrict__ out) {
    __shared__ double red[kFitThreads];
    const int4 p = pairs[blockIdx.x];
    const long long N = (long long)mem[p.x].g.W * mem[p.x].g.H;
    const float* a = snap_base[p.x] + (long long)p.y * N;
    const float* b = snap_base[p.x] + (long long)p.z * N;
    double s = 0.0;
    for (long long i = threadIdx.x; i < N; i += kFitThreads) s += (double)a[i] * (double)b[i];
    const double r = block_sum(s, red);
    if (threadIdx.x == 0) out[blockIdx.x] = r;
}

}  // namespace

void launch_fit_stats(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur, bool with_grad,
                      float* const* snap_base, int slot, double* part) {
    hipLaunchKernelGGL(kb_fit_stats, dim3(nitems), dim3(kFitThreads), 0, st, mem, items, cur, with_grad ? 1 : 0,
                       snap_base, slot, part);
}

void launch_fit_reduce(hipStream_t st, const int* first, int nmem, const double* part, double* out) {
    const int n = nmem * kFitStats;
    hipLaunchKernelGGL(kb_fit_reduce, dim3((n + 127) / 128), dim3(128), 0, st, first, nmem, part, out);
}

void launch_fit_dots(hipStream_t st, const MemberDesc* mem, const int4* pairs, int npairs, float* const* snap_base,
                     double* out) {
    if (npairs <= 0) return;
    hipLaunchKernelGGL(kb_fit_dots, dim3(npairs), dim3(kFitThreads), 0, st, mem, pairs, snap_base, out);
}

}  // namespace hp
