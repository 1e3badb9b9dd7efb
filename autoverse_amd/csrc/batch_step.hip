// batch_step.hip -- a population's whole Flow Lenia step as one persistent launch.
//
// The multi-launch step (host_batch.cpp) runs, per step, one launch per FFT pass and size
// class and one stencil launch per channel-count group, each over a work-item list.  Here one
// launch walks four phases over merged item lists -- every member's forward row pass, then
// every spectral pass, every inverse row pass, every stencil tile -- with a grid-wide barrier
// between phases, so that the classes' passes share the device instead of following one
// another.  The grid holds as many workgroups as the device keeps resident at once; each walks
// the items of a phase at a stride of the grid size.
//
// The per-item code is the multi-launch step's: the same pass bodies (fft_passes.hpp, with 256
// threads, which changes the thread mapping but not one operation of the arithmetic) and the
// same tile core (stencil2d.hpp).  The FFT passes are compiled with contraction as in their own
// files, the stencil core without, so a member's trajectory is bit-identical to the
// multi-launch step's and to a lone HipPropagator's with radix order 0.
//
// Grid barrier.  Each workgroup adds one to a device counter and waits until it reaches
// k * gridDim.x at the k-th barrier.  The wait is bounded: a workgroup that polls past the limit
// sets bit 2 of the step flags and goes on, and a launch that finds that bit set returns at
// once, so a grid that does not fit on the device ends instead of hanging; the host then
// discards the steps and recomputes them through the checked path.  The last workgroup out
// resets the counters for the next launch.
#include <atomic>

#include "fft_passes.hpp"

#pragma clang fp contract(off)
#include "stencil2d.hpp"

namespace hp {

namespace {
constexpr int kPNT = 256;       // threads of the persistent step
using PTile = StTile2;          // 32 x 16 stencil tile of 256 threads
static_assert(PTile::NT == kPNT, "the stencil tile runs with the step's threads");

// Spectral column block per length: about four elements per thread.
template <int L>
struct PSpec {
    static constexpr int CB = 1024 / L < 1 ? 1 : 1024 / L;
    static constexpr size_t LDS = sizeof(float2) * ((size_t)CB * (L + 1) + L);
};

constexpr size_t cmax(size_t a, size_t b) { return a > b ? a : b; }
constexpr size_t kPLds =
    cmax(sizeof(St2Lds<PTile>),
         cmax(cmax(PSpec<64>::LDS, PSpec<128>::LDS), cmax(PSpec<192>::LDS, cmax(PSpec<256>::LDS, 2 * 256 * 8))));

// Bounded grid barrier (see the file comment).  target: k * gridDim.x at the k-th barrier.
__device__ __forceinline__ void grid_barrier(unsigned* __restrict__ count, unsigned target, int* __restrict__ flags,
                                             int spin_limit) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // this workgroup's writes are visible device-wide before it arrives
        atomicAdd(count, 1u);
        int n = 0;
        while (__hip_atomic_load(count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++n > spin_limit) {
                atomicOr(flags, 4);
                break;
            }
            __builtin_amdgcn_s_sleep(4);
        }
        __threadfence();  // and sees every other workgroup's
    }
    __syncthreads();
}

template <int L>
__device__ __forceinline__ void p_rows_fwd(unsigned char* lds, const MemberDesc& m, int4 it, int cur) {
    float2* buf = reinterpret_cast<float2*>(lds);
    rows_fwd_core<L, kPNT, 0>(buf, buf + L, m.A[cur], m.X, m.src_chan, m.g.H * m.g.D, m.g.Wp, m.g.N, m.g.spec, m.tw_w,
                              it.y, it.z);
}

template <int L>
__device__ __forceinline__ void p_spectral(unsigned char* lds, const MemberDesc& m, int4 it) {
    constexpr int CB = PSpec<L>::CB;
    float2* buf = reinterpret_cast<float2*>(lds);
    spectral_core<L, CB, kPNT, 2, 0>(buf, buf + CB * (L + 1), m.X, m.Y, m.Kq, m.spec_work, m.klist_spec, m.g.Wh,
                                     m.g.Wp, m.g.H, m.g.Hq, m.g.spec, (long long)m.g.Dq * m.g.Hq * m.g.Wh, m.tw_h,
                                     it.y, it.z);
}

template <int L>
__device__ __forceinline__ void p_rows_inv(unsigned char* lds, const MemberDesc& m, int4 it) {
    float2* buf = reinterpret_cast<float2*>(lds);
    rows_inv_core<L, kPNT, false, 0>(buf, buf + L, m.Y, m.U, m.tgt_work, m.klist_tgt, m.gp, m.g.H * m.g.D, m.g.Wp,
                                     m.g.N, m.g.spec, 1.0f / (float)m.g.N, m.tw_w, it.y, it.z);
}

template <int CC, bool POW2>
__device__ __forceinline__ void p_stencil_cc(unsigned char* lds, const MemberDesc& m, int4 it, int cur, int* flags) {
    St2Lds<PTile>& S = *reinterpret_cast<St2Lds<PTile>*>(lds);
    const float ta = (2.0f * m.s) * (2.0f * m.s);
    const int tiles_x = (m.g.W + PTile::TX - 1) / PTile::TX;
    stencil2d_core<CC, POW2, false, PTile, gptr<const float>, gptr<float>>(
        S.sU, S.sM, S.sOX, S.sOY, S.sAct, gpc(m.A[cur]), gpc(m.U), gp(m.A[1 - cur]), m.C, m.g.W, m.g.H, m.g.N, tiles_x,
        it.y, m.dt, m.beta_a, m.nexp, m.s, ta, 1.0f / ta, flags);
}

}  // namespace

// The persistent step's item segments: each phase's items sorted by class, class by class --
// the FFT phases by length (64, 128, 192, 256), the stencil by channel count (1, 2, 3, other) and
// power-of-two box area (yes, no) -- so that each class runs in a loop of its own and the pass
// bodies never share live ranges (one dispatch over all of them spills).
constexpr int kSegFft = 4, kSegSt = 8, kSegs = 3 * kSegFft + kSegSt;
struct SegStarts {
    int s[kSegs + 1];
};

__global__ __launch_bounds__(kPNT, 4) void kb_step_persistent(const MemberDesc* __restrict__ mem,
                                                              const int4* __restrict__ items, SegStarts seg, int cur,
                                                              int* __restrict__ flags, unsigned* __restrict__ bar,
                                                              int spin_limit) {
    extern __shared__ __align__(16) unsigned char lds[];
    // an earlier step of this run timed out at a barrier: its results are discarded, so are this
    // one's, and the workgroup only counts itself out (uniform over the workgroup)
    __shared__ int s_skip;
    if (threadIdx.x == 0) s_skip = (__hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & 4) != 0;
    __syncthreads();
    const int G = (int)gridDim.x;
    // workgroup b takes the items of a phase at b, b + G, ...: its first index in segment k
    auto first = [&](int p0, int k) {
        int r = (p0 + (int)blockIdx.x - seg.s[k]) % G;
        return seg.s[k] + (r < 0 ? r + G : r);
    };
    // LDS is reused by the next item only after every thread is done with it
#define HP_SEGMENT(k, CALL)                                                      \
    for (int i = first(p0, k); i < seg.s[(k) + 1]; i += G) {                   \
        const int4 it = items[i];                                               \
        const MemberDesc& m = mem[it.x];                                        \
        CALL;                                                                   \
        __syncthreads();                                                        \
    }
    if (s_skip == 0) {
        int p0 = seg.s[0];
        HP_SEGMENT(0, p_rows_fwd<64>(lds, m, it, cur))
        HP_SEGMENT(1, p_rows_fwd<128>(lds, m, it, cur))
        HP_SEGMENT(2, p_rows_fwd<192>(lds, m, it, cur))
        HP_SEGMENT(3, p_rows_fwd<256>(lds, m, it, cur))
        grid_barrier(bar, (unsigned)G, flags, spin_limit);
        p0 = seg.s[4];
        HP_SEGMENT(4, p_spectral<64>(lds, m, it))
        HP_SEGMENT(5, p_spectral<128>(lds, m, it))
        HP_SEGMENT(6, p_spectral<192>(lds, m, it))
        HP_SEGMENT(7, p_spectral<256>(lds, m, it))
        grid_barrier(bar, 2u * (unsigned)G, flags, spin_limit);
        p0 = seg.s[8];
        HP_SEGMENT(8, p_rows_inv<64>(lds, m, it))
        HP_SEGMENT(9, p_rows_inv<128>(lds, m, it))
        HP_SEGMENT(10, p_rows_inv<192>(lds, m, it))
        HP_SEGMENT(11, p_rows_inv<256>(lds, m, it))
        grid_barrier(bar, 3u * (unsigned)G, flags, spin_limit);
        p0 = seg.s[12];
        HP_SEGMENT(12, (p_stencil_cc<1, true>(lds, m, it, cur, flags)))
        HP_SEGMENT(13, (p_stencil_cc<1, false>(lds, m, it, cur, flags)))
        HP_SEGMENT(14, (p_stencil_cc<2, true>(lds, m, it, cur, flags)))
        HP_SEGMENT(15, (p_stencil_cc<2, false>(lds, m, it, cur, flags)))
        HP_SEGMENT(16, (p_stencil_cc<3, true>(lds, m, it, cur, flags)))
        HP_SEGMENT(17, (p_stencil_cc<3, false>(lds, m, it, cur, flags)))
        HP_SEGMENT(18, (p_stencil_cc<0, true>(lds, m, it, cur, flags)))
        HP_SEGMENT(19, (p_stencil_cc<0, false>(lds, m, it, cur, flags)))
    }
#undef HP_SEGMENT
    if (threadIdx.x == 0 && atomicAdd(bar + 1, 1u) == gridDim.x - 1) {  // the last workgroup out
        atomicExch(bar, 0u);
        atomicExch(bar + 1, 0u);
    }
}

// Segment of an item: FFT phases by length, the stencil by channel count and box area.
int batch_step_persistent_segment(int phase, int L, int C, bool pow2) {
    if (phase < 3) return phase * kSegFft + (L == 64 ? 0 : L == 128 ? 1 : L == 192 ? 2 : 3);
    return 3 * kSegFft + 2 * (C >= 1 && C <= 3 ? C - 1 : 3) + (pow2 ? 0 : 1);
}
int batch_step_persistent_segments() { return kSegs; }

bool batch_step_persistent_size(int L) { return L == 64 || L == 128 || L == 192 || L == 256; }
int batch_step_persistent_tile_w() { return PTile::TX; }
int batch_step_persistent_tile_h() { return PTile::TY; }
int batch_step_persistent_block_lines(int L) {
    switch (L) {
    case 64: return PSpec<64>::CB;
    case 128: return PSpec<128>::CB;
    case 192: return PSpec<192>::CB;
    case 256: return PSpec<256>::CB;
    default: return 0;
    }
}

int batch_step_persistent_grid() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) dev = 0;
    int n = cached[dev].load(std::memory_order_relaxed);
    if (n > 0) return n;
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kb_step_persistent, kPNT, kPLds) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) cus = 1;
    n = per_cu * cus;
    cached[dev].store(n, std::memory_order_relaxed);
    return n;
}

void launch_batch_step_persistent(hipStream_t st, const MemberDesc* mem, const int4* items, const int* starts,
                                  int cur, int* flags, unsigned* bar, int grid, int spin_limit) {
    SegStarts seg;
    for (int i = 0; i <= kSegs; i++) seg.s[i] = starts[i];
    hipLaunchKernelGGL(kb_step_persistent, dim3(grid), dim3(kPNT), kPLds, st, mem, items, seg, cur, flags, bar,
                       spin_limit);
}

}  // namespace hp
