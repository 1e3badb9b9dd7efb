// stencil_kernels.hip -- stages 2-5 of the Flow Lenia step on gfx950, and the FP64 mass.
//
// Compiled with -ffp-contract=off: every product and sum rounds on its own, as in
// the reference's Rust f32 code, so that given the same affinity U these kernels
// reproduce CpuPropagator's next state bit for bit.
//
// Reference stages restated (paths relative to the reference repository):
//   channel sum            src/compute/propagator.rs:304-310  (0 + A_0 + A_1 + ...)
//   Sobel 2D / 3D          src/compute/gradient.rs:68-108, src/compute/gradient3d.rs:41-143
//   alpha, flow            src/compute/flow.rs:14-32, 68-87; src/compute/flow3d.rs:50-73
//   reintegration          src/compute/reintegration.rs:60-174, src/compute/reintegration3d.rs:56-198
//
// The reference reintegrates by scattering each source's mass over the cells its
// box [dest - s, dest + s] overlaps, adding into next[] in source raster order.
// Here every destination gathers instead: it visits the sources that can reach it
// in ascending (wrapped) raster order and adds exactly the same products, so each
// destination's sum is formed in the same order with the same roundings --
// deterministic and atomic-free.  The fused kernel assumes every active source
// lands within RG cells; a source that does not raises a flag and the host reruns
// the step through the general-radius path (boxes + gather_general), which walks
// every periodic image the scatter could produce.
#include <math.h>

#include <algorithm>

#include "stencil_common.hpp"
#include "stencil2d.hpp"

namespace hp {

template <int CC, bool POW2, bool CONS, class TL, class PC = const float*, class PW = float*>
__device__ __forceinline__ void stencil2d_body(PC A, PC U, PW Aout, int Cdyn, int W, int H, long long N,
                                               int tiles_x, int tile, float dt, float beta_a, float nexp, float s,
                                               float ta, float inv_ta, int* __restrict__ flag) {
    __shared__ float sU[2][TL::NLOAD];
    __shared__ float sM[TL::NSRC];
    __shared__ float sOX[4][TL::NSRC];
    __shared__ float sOY[4][TL::NSRC];
    __shared__ unsigned sAct[TL::NT / 64];
    stencil2d_core<CC, POW2, CONS, TL, PC, PW>(sU, sM, sOX, sOY, sAct, A, U, Aout, Cdyn, W, H, N, tiles_x, tile, dt,
                                               beta_a, nexp, s, ta, inv_ta, flag);
}

// Tile rows per XCD block of the lone 2D stencil.  At 1024^2 x 3 (64 x 16 tiles) blocks of 4
// rows took 27.8 us against 29.0 us for single rows and 28.6 us for pairs; blocks of 8 (one
// contiguous band per XCD) varied from run to run (27.7 and 29.8 us).  The 3D stencil gained
// nothing from blocks of 2 or 4 and deals single rows.
constexpr int kSt2RowBlock = 4;

template <int CC, bool POW2, bool CONS, class TL>
__global__ __launch_bounds__(TL::NT, TL::WAVES_PER_EU) void k_stencil2d(const float* __restrict__ A, const float* __restrict__ U,
                                                      float* __restrict__ Aout, int Cdyn, int W, int H, long long N,
                                                      int tiles_x, int ntiles, float dt, float beta_a, float nexp,
                                                      float s, float ta, float inv_ta, int* __restrict__ flag) {
    const int tile = xcd_rows_tile<kSt2RowBlock>(blockIdx.x, tiles_x, ntiles / tiles_x);
    if (tile >= ntiles) return;
    stencil2d_body<CC, POW2, CONS, TL>(A, U, Aout, Cdyn, W, H, N, tiles_x, tile, dt, beta_a, nexp, s, ta, inv_ta,
                                       flag);
}

// batched: item = (member, tile); every member of one launch has the same channel count class
template <int CC, bool POW2, class TL>
__global__ __launch_bounds__(TL::NT, TL::WAVES_PER_EU) void kb_stencil2d(const MemberDesc* __restrict__ mem,
                                                       const int4* __restrict__ items, int cur,
                                                       int* __restrict__ flag) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const float ta = (2.0f * m.s) * (2.0f * m.s);
    const int tiles_x = (m.g.W + TL::TX - 1) / TL::TX;
    stencil2d_body<CC, POW2, false, TL>(gpc(m.A[cur]), gpc(m.U), gp(m.A[1 - cur]), m.C, m.g.W, m.g.H, m.g.N, tiles_x,
                                        it.y, m.dt, m.beta_a, m.nexp, m.s, ta, 1.0f / ta, flag);
}

// HP_TILE binds the run-time tile variant to the compile-time tile type TL of the call.
#define HP_TILE(t, ...)                          \
    do {                                         \
        switch (t) {                             \
        case 1: {                                \
            using TL = StTile1;                  \
            __VA_ARGS__;                         \
        } break;                                 \
        case 2: {                                \
            using TL = StTile2;                  \
            __VA_ARGS__;                         \
        } break;                                 \
        case 3: {                                \
            using TL = StTile3;                  \
            __VA_ARGS__;                         \
        } break;                                 \
        case 4: {                                \
            using TL = StTile4;                  \
            __VA_ARGS__;                         \
        } break;                                 \
        case 5: {                                \
            using TL = StTile5;                  \
            __VA_ARGS__;                         \
        } break;                                 \
        default: {                               \
            using TL = StTile0;                  \
            __VA_ARGS__;                         \
        }                                        \
        }                                        \
    } while (0)

template <class TL>
static void launch_kb_st2(hipStream_t st, int C, bool pow2, const MemberDesc* mem, const int4* items, int nitems,
                          int cur, int* flag) {
#define B_(cc)                                                                                                   \
    if (pow2)                                                                                                    \
        hipLaunchKernelGGL((kb_stencil2d<cc, true, TL>), dim3(nitems), dim3(TL::NT), 0, st, mem, items, cur, flag); \
    else                                                                                                         \
        hipLaunchKernelGGL((kb_stencil2d<cc, false, TL>), dim3(nitems), dim3(TL::NT), 0, st, mem, items, cur, flag);
    switch (C) {
    case 1: B_(1) break;
    case 2: B_(2) break;
    case 3: B_(3) break;
    case 4: B_(4) break;
    default: B_(0)
    }
#undef B_
}

bool launch_batch_stencil2d(hipStream_t st, int C, bool pow2, const MemberDesc* mem, const int4* items, int nitems,
                            int cur, int* flag, int tile) {
    HP_TILE(tile, launch_kb_st2<TL>(st, C, pow2, mem, items, nitems, cur, flag));
    return true;
}

int stencil2d_tile_w(int tile) {
    int w = 0;
    HP_TILE(tile, w = TL::TX);
    return w;
}
int stencil2d_tile_h(int tile) {
    int h = 0;
    HP_TILE(tile, h = TL::TY);
    return h;
}
int stencil2d_tile_count() { return kStencilTiles; }

template <int CC, class TL>
static void launch_st2(hipStream_t st, bool pow2, bool cons, dim3 grid, const float* A, const float* U, float* Aout,
                       int C, const Geo& g, int tiles_x, int ntiles, float dt, float beta_a, float nexp, float s,
                       float ta, float inv_ta, int* flag) {
    if (cons)
        hipLaunchKernelGGL((k_stencil2d<CC, false, true, TL>), grid, dim3(TL::NT), 0, st, A, U, Aout, C, g.W, g.H,
                           g.N, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag);
    else if (pow2)
        hipLaunchKernelGGL((k_stencil2d<CC, true, false, TL>), grid, dim3(TL::NT), 0, st, A, U, Aout, C, g.W, g.H,
                           g.N, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag);
    else
        hipLaunchKernelGGL((k_stencil2d<CC, false, false, TL>), grid, dim3(TL::NT), 0, st, A, U, Aout, C, g.W, g.H,
                           g.N, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag);
}

bool stencil2d_supported(const Geo& g, float s, int tile) {
    // the rotated visiting orders need every strip's source rows / columns to be distinct cells;
    // the 16-source window assumes s <= 1; the point-distribution fallback (total area < 1e-10)
    // lives in the general path
    const float ta = (2.0f * s) * (2.0f * s);
    int srows = 0;
    HP_TILE(tile, srows = TL::SROWS);
    return g.D == 1 && g.W >= 4 && g.H >= srows && s > 0.0f && s <= 1.0f && ta >= 1e-10f;
}

template <class TL>
static void launch_k_st2(hipStream_t st, const float* A, const float* U, float* Aout, int C, const Geo& g, float dt,
                         float beta_a, float nexp, float s, int* flag, bool cons) {
    const float ta = (2.0f * s) * (2.0f * s);
    int e;
    const float fr = frexpf(ta, &e);
    const bool pow2 = fr == 0.5f && ta >= 1.17549435e-38f && ta < INFINITY;
    const float inv_ta = 1.0f / ta;
    const int tiles_x = (g.W + TL::TX - 1) / TL::TX, tiles_y = (g.H + TL::TY - 1) / TL::TY;
    const int ntiles = tiles_x * tiles_y;
    const dim3 grid(xcd_rows_grid<kSt2RowBlock>(tiles_x, ntiles / tiles_x));
    switch (C) {
    case 1: launch_st2<1, TL>(st, pow2, cons, grid, A, U, Aout, C, g, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag); break;
    case 2: launch_st2<2, TL>(st, pow2, cons, grid, A, U, Aout, C, g, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag); break;
    case 3: launch_st2<3, TL>(st, pow2, cons, grid, A, U, Aout, C, g, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag); break;
    case 4: launch_st2<4, TL>(st, pow2, cons, grid, A, U, Aout, C, g, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag); break;
    default: launch_st2<0, TL>(st, pow2, cons, grid, A, U, Aout, C, g, tiles_x, ntiles, dt, beta_a, nexp, s, ta, inv_ta, flag);
    }
}

bool launch_stencil2d(hipStream_t st, const float* A, const float* U, float* Aout, int C, const Geo& g, float dt,
                      float beta_a, float nexp, float s, int* flag, int tile, bool exact) {
    if (!stencil2d_supported(g, s, tile)) return false;
    HP_TILE(tile, launch_k_st2<TL>(st, A, U, Aout, C, g, dt, beta_a, nexp, s, flag, exact));
    return true;
}

// ---------------------------------------------------------------------------------------------
// Fused 3D stencil: channel sum, 27-point Sobel of the sum and of U (gradient3d.rs:75-138,
// same grouping, / 32), alpha, flow (flow3d.rs:50-73), per-source box and its four overlaps
// per axis (reintegration3d.rs:104-176), and the 64-source gather.
//
// A source at x whose box has |dt F| < 1 and s <= 1 reaches cells x-2 .. x+1 on every axis,
// so a destination receives from offsets o = source - dest in {-1, 0, 1, 2}^3.  Every
// destination visits them in ascending wrapped raster order (z, then y, then x; rotated at
// the periodic edges), the scatter's order, and adds m * (((ow*oh)*od) / V) exactly as the
// reference forms cell_mass; zero-volume pairs add +-0 and change nothing.  Any owned source
// whose box reaches further sets flag bit 0 (bit 1 for non-finite values) and the host reruns
// the step through the general path.
//
// Work split.  A workgroup owns a 32 x 8 x 4 destination tile (its 35 x 11 x 7 source region is
// 2.6 sources per destination; a half-wave's gather reads one contiguous source row, which keeps
// LDS bank conflicts lower than 16-wide tiles do).  LDS holds one affinity tile and, per source,
// its mass and four overlaps per axis; the channel-sum tile shares memory with the z overlaps,
// which are first written after the alpha phase.  That is 154 KiB: one workgroup of 1024
// threads per CU, four waves per SIMD.  Each thread gathers a z strip of SR destinations.
//
// Empty regions.  The reference skips every source with |m| < 1e-10; a channel none of whose
// sources in the tile's region passes that test sends no mass into the tile, whose outputs are
// then +0, as the reference's.  Such a channel skips the source and gather phases (the test is
// uniform over the workgroup), and a tile with no active channel skips everything after its
// loads.
// ---------------------------------------------------------------------------------------------
namespace st3 {
template <int TX_, int TY_, int TZ_, int SR_>
struct Tile {
    static constexpr int TX = TX_, TY = TY_, TZ = TZ_, SR = SR_, NT = TX * TY * TZ / SR;
    static constexpr int SX = TX + 3, SY = TY + 3, SZ = TZ + 3;  // sources: offsets -1 .. +2 around the tile
    static constexpr int LX = SX + 2, LY = SY + 2, LZ = SZ + 2;  // + Sobel halo
    static constexpr int NSRC = SX * SY * SZ;
    static constexpr int NLOAD = LX * LY * LZ;
    static constexpr int LPT = (NLOAD + NT - 1) / NT;  // load-region cells per thread
    static constexpr int SPT = (NSRC + NT - 1) / NT;   // sources per thread
    static constexpr int SROWS = SR + 3;               // source planes of one strip
    static constexpr int LDS = 4 * (NLOAD + 13 * NSRC);
    static_assert(TZ % SR == 0 && NLOAD <= 4 * NSRC && LDS <= 163840, "strip, channel-sum aliasing, LDS");
};
using Tile3 = Tile<32, 8, 4, 1>;
}  // namespace st3

// gradient3d.rs:75-138 on an LDS box: P(dz, dy, dx) = L[i + dz*sz + dy*sy + dx].
__device__ __forceinline__ void sobel27_lds(const float* L, int i, int sy, int sz, float& vx, float& vy, float& vz) {
    auto P = [&](int dz, int dy, int dx) -> float { return L[i + dz * sz + dy * sy + dx]; };
    const float p_mmm = P(-1, -1, -1), p_mm0 = P(-1, -1, 0), p_mmp = P(-1, -1, 1);
    const float p_m0m = P(-1, 0, -1), p_m00 = P(-1, 0, 0), p_m0p = P(-1, 0, 1);
    const float p_mpm = P(-1, 1, -1), p_mp0 = P(-1, 1, 0), p_mpp = P(-1, 1, 1);
    const float p_0mm = P(0, -1, -1), p_0m0 = P(0, -1, 0), p_0mp = P(0, -1, 1);
    const float p_00m = P(0, 0, -1), p_00p = P(0, 0, 1);
    const float p_0pm = P(0, 1, -1), p_0p0 = P(0, 1, 0), p_0pp = P(0, 1, 1);
    const float p_pmm = P(1, -1, -1), p_pm0 = P(1, -1, 0), p_pmp = P(1, -1, 1);
    const float p_p0m = P(1, 0, -1), p_p00 = P(1, 0, 0), p_p0p = P(1, 0, 1);
    const float p_ppm = P(1, 1, -1), p_pp0 = P(1, 1, 0), p_ppp = P(1, 1, 1);
    const float gx_zm = -1.0f * (1.0f * p_mmm + 2.0f * p_m0m + 1.0f * p_mpm) + 1.0f * (1.0f * p_mmp + 2.0f * p_m0p + 1.0f * p_mpp);
    const float gx_z0 = -1.0f * (1.0f * p_0mm + 2.0f * p_00m + 1.0f * p_0pm) + 1.0f * (1.0f * p_0mp + 2.0f * p_00p + 1.0f * p_0pp);
    const float gx_zp = -1.0f * (1.0f * p_pmm + 2.0f * p_p0m + 1.0f * p_ppm) + 1.0f * (1.0f * p_pmp + 2.0f * p_p0p + 1.0f * p_ppp);
    vx = (1.0f * gx_zm + 2.0f * gx_z0 + 1.0f * gx_zp) / 32.0f;
    const float gy_zm = -1.0f * (1.0f * p_mmm + 2.0f * p_mm0 + 1.0f * p_mmp) + 1.0f * (1.0f * p_mpm + 2.0f * p_mp0 + 1.0f * p_mpp);
    const float gy_z0 = -1.0f * (1.0f * p_0mm + 2.0f * p_0m0 + 1.0f * p_0mp) + 1.0f * (1.0f * p_0pm + 2.0f * p_0p0 + 1.0f * p_0pp);
    const float gy_zp = -1.0f * (1.0f * p_pmm + 2.0f * p_pm0 + 1.0f * p_pmp) + 1.0f * (1.0f * p_ppm + 2.0f * p_pp0 + 1.0f * p_ppp);
    vy = (1.0f * gy_zm + 2.0f * gy_z0 + 1.0f * gy_zp) / 32.0f;
    const float gz_xm = -1.0f * (1.0f * p_mmm + 2.0f * p_m0m + 1.0f * p_mpm) + 1.0f * (1.0f * p_pmm + 2.0f * p_p0m + 1.0f * p_ppm);
    const float gz_x0 = -1.0f * (1.0f * p_mm0 + 2.0f * p_m00 + 1.0f * p_mp0) + 1.0f * (1.0f * p_pm0 + 2.0f * p_p00 + 1.0f * p_pp0);
    const float gz_xp = -1.0f * (1.0f * p_mmp + 2.0f * p_m0p + 1.0f * p_mpp) + 1.0f * (1.0f * p_pmp + 2.0f * p_p0p + 1.0f * p_ppp);
    vz = (1.0f * gz_xm + 2.0f * gz_x0 + 1.0f * gz_xp) / 32.0f;
}

// first visited offset index of the ascending wrapped order of g + {-1, 0, 1, 2} on n cells
__device__ __forceinline__ int wrap_rotation(int g, int n) {
    return g < 1 ? 1 - g : (g + 2 >= n ? 4 - (g + 2 - (n - 1)) : 0);
}

// CONS: the exact-conservation mode of the 2D kernel, with volumes ((ow * oh) * od) in double.
template <bool POW2, bool CONS, class TL>
__global__ __launch_bounds__(TL::NT) void k_stencil3d(const float* __restrict__ A, const float* __restrict__ U,
                                                    float* __restrict__ Aout, int C, int W, int H, int D, int N,
                                                    int tiles_x, int tiles_y, int ntiles, float dt, float beta_a,
                                                    float nexp, float s, float tv, float inv_tv,
                                                    int* __restrict__ flag) {
    constexpr int TX = TL::TX, TY = TL::TY, TZ = TL::TZ, SR = TL::SR, NT = TL::NT;
    constexpr int SX = TL::SX, SY = TL::SY, LX = TL::LX, LY = TL::LY;
    constexpr int NSRC = TL::NSRC, NLOAD = TL::NLOAD, LPT = TL::LPT, SPT = TL::SPT, SROWS = TL::SROWS;
    __shared__ float sU[NLOAD];
    __shared__ float sM[NSRC];
    __shared__ float sOX[4 * NSRC], sOY[4 * NSRC], sOZ[4 * NSRC];  // overlap j of source i at [j * NSRC + i]
    __shared__ unsigned sAct[NT / 64];
    float* const sSum = sOZ;

    const int tile = xcd_row_tile(blockIdx.x, tiles_x, ntiles / tiles_x);
    if (tile >= ntiles) return;
    const int x0 = (tile % tiles_x) * TX, y0 = ((tile / tiles_x) % tiles_y) * TY, z0 = (tile / (tiles_x * tiles_y)) * TZ;
    const int tid = threadIdx.x;
    const int HW = W * H;
    // strip geometry: destination (gdx, gdy), planes gdz0 .. gdz0 + SR - 1
    const int lx = tid % TX, ly = (tid / TX) % TY, lzs = tid / (TX * TY);
    const int gdx = x0 + lx, gdy = y0 + ly, gdz0 = z0 + lzs * SR;
    const bool inside = gdx < W && gdy < H && gdz0 < D;
    // load-region coordinates lie in [-2, n + T + 1): one wrap step suffices unless the grid is
    // smaller than the tile plus its halo
    const bool multi = W < TX + 2 || H < TY + 2 || D < TZ + 2;

    // ---- every global read of the tile, issued together ----
    int goff[LPT];
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
        const int i = tid + q * NT;
        const int lz = i / (LX * LY), r = i - lz * (LX * LY), ly = r / LX, lx = r - ly * LX;
        goff[q] = i < NLOAD ? wrap_tile(z0 - 2 + lz, D, multi) * HW + wrap_tile(y0 - 2 + ly, H, multi) * W +
                                  wrap_tile(x0 - 2 + lx, W, multi)
                            : 0;
    }
    int soff[SPT];  // grid offsets of the owned sources
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int i = tid + q * NT;
        const int sz = i / (SX * SY), r = i - sz * (SX * SY), sy = r / SX, sx = r - sy * SX;
        soff[q] = i < NSRC ? wrap_tile(z0 - 1 + sz, D, multi) * HW + wrap_tile(y0 - 1 + sy, H, multi) * W +
                                 wrap_tile(x0 - 1 + sx, W, multi)
                           : 0;
    }
    // the owned sources' masses, the first affinity and channel 0's state over the load region
    // are all requested before any of them is used (used as loaded, each group cost a memory
    // latency, and the channel sum waited for every load in turn)
    float msrc[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) msrc[q] = A[soff[q]];
    float upre[LPT], a0[LPT];
#pragma unroll
    for (int q = 0; q < LPT; ++q) upre[q] = U[goff[q]];
#pragma unroll
    for (int q = 0; q < LPT; ++q) a0[q] = A[goff[q]];
    // bit c: channel c has a source in the region that the reference does not skip (channels
    // past 31 count as active).  A channel without one receives no mass anywhere in the tile.
    unsigned act_t = 0;
#pragma unroll
    for (int q = 0; q < SPT; ++q)
        if (tid + q * NT < NSRC && active_mass(msrc[q])) act_t = 1u;
    for (int c = 1; c < C && c < 32; ++c) {
        const float* Ac = A + (long long)c * N;
#pragma unroll
        for (int q = 0; q < SPT; ++q)
            if (tid + q * NT < NSRC && active_mass(Ac[soff[q]])) act_t |= 1u << c;
    }
    act_t = wave_or(act_t);
    if ((tid & 63) == 0) sAct[tid >> 6] = act_t;
#pragma unroll
    for (int q = 0; q < LPT; ++q) {  // channel sum (propagator3d.rs: 0 + A_0 + A_1 + ...)
        float sum = 0.0f + a0[q];
        for (int c = 1; c < C; ++c) sum += A[(long long)c * N + goff[q]];
        const int i = tid + q * NT;
        if (i < NLOAD) {
            sSum[i] = sum;
            sU[i] = upre[q];
        }
    }
    lds_barrier();
    unsigned act = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) act |= sAct[w];
    if (act == 0 && C <= 32) {  // no mass reaches the tile: every output is +0, as the reference's
        if (inside) {
            for (int c = 0; c < C; ++c) {
                float* Oc = Aout + (long long)c * N + gdy * W + gdx;
#pragma unroll
                for (int r = 0; r < SR; ++r)
                    if (gdz0 + r < D) Oc[(gdz0 + r) * HW] = 0.0f;
            }
        }
        return;
    }
    // alpha and the mass gradient of the owned sources.  The loop stays rolled (alpha_of may
    // expand to a powf); each thread parks its own results in the x overlap planes, which are
    // free until the first source phase, and reads them back into registers -- the same thread
    // writes and reads every entry, so no barrier is needed.
#pragma unroll 1
    for (int i = tid; i < NSRC; i += NT) {
        const int sz = i / (SX * SY), r = i - sz * (SX * SY), sy = r / SX, sx = r - sy * SX;
        const int li = ((sz + 1) * LY + (sy + 1)) * LX + (sx + 1);
        float gx, gy, gz;
        sobel27_lds(sSum, li, LX, LX * LY, gx, gy, gz);
        const float al = alpha_of(sSum[li], beta_a, nexp);
        sOX[i] = 1.0f - al;
        sOX[NSRC + i] = al * gx;
        sOX[2 * NSRC + i] = al * gy;
        sOX[3 * NSRC + i] = al * gz;
    }
    float aOm[SPT], aGx[SPT], aGy[SPT], aGz[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int i = tid + q * NT;
        const bool ok = i < NSRC;
        aOm[q] = ok ? sOX[i] : 0.0f;
        aGx[q] = ok ? sOX[NSRC + i] : 0.0f;
        aGy[q] = ok ? sOX[2 * NSRC + i] : 0.0f;
        aGz[q] = ok ? sOX[3 * NSRC + i] : 0.0f;
    }
    lds_barrier();  // the channel sum (in the z overlap planes) has been read

    const int rx = wrap_rotation(gdx, W), ry = wrap_rotation(gdy, H);
    // ascending wrapped order of the strip's planes: start with the first that does not wrap low
    const int rz = gdz0 < 1 ? 1 - gdz0 : (gdz0 + SR + 1 >= D ? SROWS - (gdz0 + SR + 1 - (D - 1)) : 0);
    const int base = (lzs * SR * SY + ly) * SX + lx;  // source (plane 0, row 0, column 0) of the strip

    int local_flag = 0;
#pragma unroll 1
    for (int c = 0; c < C; ++c) {
        const bool live = c >= 32 || ((act >> c) & 1u);  // uniform over the workgroup
        // ---- sources: flow, box, overlaps ----
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int i = tid + q * NT;
            if (i >= NSRC || !live) break;
            const float m = msrc[q];
            if (fabsf(m) < 1e-10f) {  // reintegration3d.rs:77: a skipped source adds nothing
                sM[i] = 0.0f;
#pragma unroll
                for (int j = 0; j < 4; ++j) sOX[j * NSRC + i] = sOY[j * NSRC + i] = sOZ[j * NSRC + i] = 0.0f;
                continue;
            }
            const int sz = i / (SX * SY), r = i - sz * (SX * SY), sy = r / SX, sx = r - sy * SX;
            const int li = ((sz + 1) * LY + (sy + 1)) * LX + (sx + 1);
            float gux, guy, guz;
            sobel27_lds(sU, li, LX, LX * LY, gux, guy, guz);
            const float fx = aOm[q] * gux - aGx[q];
            const float fy = aOm[q] * guy - aGy[q];
            const float fz = aOm[q] * guz - aGz[q];
            // wrapped grid coordinates: integer-valued floats below 2^24
            const int gxs = wrap_tile(x0 - 1 + sx, W, multi), gys = wrap_tile(y0 - 1 + sy, H, multi),
                      gzs = wrap_tile(z0 - 1 + sz, D, multi);
            const float cx = (float)gxs, cy = (float)gys, cz = (float)gzs;
            const float dxp = cx + dt * fx;  // reintegration3d.rs:83-85
            const float dyp = cy + dt * fy;
            const float dzp = cz + dt * fz;
            const float xm = dxp - s, xM = dxp + s, ym = dyp - s, yM = dyp + s, zm = dzp - s, zM = dzp + s;
            // overlaps with destination cells g + 1 - j, j = o + 1 (o = source - dest)
            double wx = 0.0, wy = 0.0, wz = 0.0;  // width sums, cells c-2 .. c+1 ascending (CONS)
#pragma unroll
            for (int j = 3; j >= 0; --j) {
                const float ox = overlap_clamped(cx + (float)(1 - j), cx + (float)(2 - j), xm, xM);
                const float oy = overlap_clamped(cy + (float)(1 - j), cy + (float)(2 - j), ym, yM);
                const float oz = overlap_clamped(cz + (float)(1 - j), cz + (float)(2 - j), zm, zM);
                sOX[j * NSRC + i] = ox;
                sOY[j * NSRC + i] = oy;
                sOZ[j * NSRC + i] = oz;
                if (CONS) {
                    wx += (double)ox;
                    wy += (double)oy;
                    wz += (double)oz;
                }
            }
            // With (2s)^3 a power of two the scaling by its inverse is exact (the masses skipped
            // above are the only ones below 1e-10, every nonzero overlap is a multiple of an ulp
            // of s or of a coordinate, so no product is subnormal before the last one): m * (vol *
            // inv) equals (m * inv) * vol bit for bit, and the gather saves a multiply per pair.
            // A scaled mass that overflows is sent to the general path below.
            float sm;
            if (CONS) {
                const double pw = (wx * wy) * wz;
                sm = pw > 0.0 ? (float)((double)m / pw) : 0.0f;
            } else {
                sm = POW2 ? m * inv_tv : m;
            }
            sM[i] = sm;
            // every source is checked by the one tile that owns it
            const int ux = x0 - 1 + sx, uy = y0 - 1 + sy, uz = z0 - 1 + sz;
            if (ux >= x0 && ux < x0 + TX && uy >= y0 && uy < y0 + TY && uz >= z0 && uz < z0 + TZ && ux < W && uy < H &&
                uz < D) {
                // floor(xm) < c - 2  <=>  xm < c - 2 for integer c (likewise ceil); a non-finite
                // value fails the comparisons and is classified on the slow side
                const bool ok = xm >= cx - 2.0f && xM <= cx + 2.0f && ym >= cy - 2.0f && yM <= cy + 2.0f &&
                                zm >= cz - 2.0f && zM <= cz + 2.0f && (CONS || fabsf(sm) <= 3.40282347e38f);
                if (!ok)
                    local_flag |= (isfinite(xm) && isfinite(xM) && isfinite(ym) && isfinite(yM) && isfinite(zm) &&
                                   isfinite(zM) && isfinite(m))
                                      ? 1
                                      : 2;
            }
        }
        // next channel's affinity and masses: into registers now, the affinity into LDS after the
        // barrier (the gather does not read it)
        if (c + 1 < C) {
            const float* Un = U + (long long)(c + 1) * N;
#pragma unroll
            for (int q = 0; q < LPT; ++q) upre[q] = Un[goff[q]];
            const float* An = A + (long long)(c + 1) * N;
#pragma unroll
            for (int q = 0; q < SPT; ++q) msrc[q] = An[soff[q]];
        }
        lds_barrier();
        if (c + 1 < C) {
#pragma unroll
            for (int q = 0; q < LPT; ++q)
                if (tid + q * NT < NLOAD) sU[tid + q * NT] = upre[q];
        }
        // ---- gather: strip of SR destinations ----
        // A zero-volume pair contributes m * 0 = +-0; added to an accumulator that started at +0
        // it changes no bit, so every pair is added without the reference's `volume > 0` test.
        // Waves with no periodic edge in reach take the unrotated path, whose LDS addresses are
        // one base plus constants; waves touching an edge visit in rotated order.
        float acc[SR];
        double accd[CONS ? SR : 1];
#pragma unroll
        for (int r = 0; r < SR; ++r) acc[r] = 0.0f;
#pragma unroll
        for (int r = 0; r < (CONS ? SR : 1); ++r) accd[r] = 0.0;
        if (!live) {
        } else if (__all(rx == 0 && ry == 0 && rz == 0 && inside)) {
#pragma unroll
            for (int a = 0; a < SROWS; ++a) {  // strip source plane a <-> grid plane gdz0 - 1 + a
#pragma unroll
                for (int jy = 0; jy < 4; ++jy) {
#pragma unroll
                    for (int jx = 0; jx < 4; ++jx) {
                        const int si = base + (a * SY + jy) * SX + jx;
                        const float m = sM[si];
                        if (CONS) {
                            const double ad = (double)sOX[jx * NSRC + si] * (double)sOY[jy * NSRC + si];
#pragma unroll
                            for (int r = 0; r < SR; ++r)
                                if (a - r >= 0 && a - r <= 3)
                                    accd[CONS ? r : 0] += (double)m * (ad * (double)sOZ[(a - r) * NSRC + si]);
                            continue;
                        }
                        const float area = sOX[jx * NSRC + si] * sOY[jy * NSRC + si];
#pragma unroll
                        for (int r = 0; r < SR; ++r) {
                            if (a - r >= 0 && a - r <= 3) {
                                const float vol = area * sOZ[(a - r) * NSRC + si];
                                acc[r] += POW2 ? m * vol : m * (vol / tv);
                            }
                        }
                    }
                }
            }
        } else if (inside) {
            // a strip at a periodic edge: the same pairs in rotated order (per lane in x and y,
            // per strip in z), the loop over planes rolled, every pair formed from in-bounds LDS
            // entries and added only where the source plane reaches the destination
            int xo[4], yoS[4], xoN[4], yoN[4];  // slot k -> column jx (and jx * NSRC), row jy * SX (jy * NSRC)
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int jx = (k + rx) & 3, jy = (k + ry) & 3;
                xo[k] = jx;
                xoN[k] = jx * NSRC;
                yoS[k] = jy * SX;
                yoN[k] = jy * NSRC;
            }
#pragma unroll 1
            for (int it = 0; it < SROWS; ++it) {
                int a = it + rz;
                if (a >= SROWS) a -= SROWS;
                const int pb = base + a * SY * SX;
#pragma unroll
                for (int ky = 0; ky < 4; ++ky) {
                    const int rb = pb + yoS[ky];
#pragma unroll
                    for (int kx = 0; kx < 4; ++kx) {
                        const int si = rb + xo[kx];
                        const float m = sM[si];
                        const float ox = sOX[xoN[kx] + si], oy = sOY[yoN[ky] + si];
                        if (CONS) {
                            const double ad = (double)ox * (double)oy;
#pragma unroll
                            for (int r = 0; r < SR; ++r) {
                                const int jz = a - r;
                                const bool ok = (unsigned)jz <= 3u;
                                const double v = (double)m * (ad * (double)sOZ[(ok ? jz : 0) * NSRC + si]);
                                accd[CONS ? r : 0] = ok ? accd[CONS ? r : 0] + v : accd[CONS ? r : 0];
                            }
                            continue;
                        }
                        const float area = ox * oy;
#pragma unroll
                        for (int r = 0; r < SR; ++r) {
                            const int jz = a - r;
                            const bool ok = (unsigned)jz <= 3u;
                            const float vol = area * sOZ[(ok ? jz : 0) * NSRC + si];
                            const float v = POW2 ? m * vol : m * (vol / tv);
                            acc[r] = ok ? acc[r] + v : acc[r];
                        }
                    }
                }
            }
        }
        if (CONS) {
#pragma unroll
            for (int r = 0; r < SR; ++r) acc[r] = (float)accd[CONS ? r : 0];
        }
        if (inside) {
            float* Oc = Aout + (long long)c * N + gdy * W + gdx;
#pragma unroll
            for (int r = 0; r < SR; ++r)
                if (gdz0 + r < D) Oc[(gdz0 + r) * HW] = acc[r];
        }
        if (c + 1 < C) lds_barrier();  // the next channel's source phase overwrites the overlaps
    }
    if (local_flag) atomicOr(flag, local_flag);
}

bool stencil3d_supported(const Geo& g, float s) {
    // distinct cells in every strip's window; int offsets; the 64-source window assumes s <= 1
    const float tv = (2.0f * s) * (2.0f * s) * (2.0f * s);
    return g.D > 1 && g.W >= 4 && g.H >= 4 && g.D >= st3::Tile3::SROWS && g.N <= 0x7fffffffLL && s > 0.0f &&
           s <= 1.0f && tv >= 1e-10f;
}

bool launch_stencil3d(hipStream_t st, const float* A, const float* U, float* Aout, int C, const Geo& g, float dt,
                      float beta_a, float nexp, float s, int* flag, bool exact) {
    using TL = st3::Tile3;
    if (!stencil3d_supported(g, s)) return false;
    const float tv = (2.0f * s) * (2.0f * s) * (2.0f * s);  // reintegration3d.rs:134
    int e;
    const bool pow2 = frexpf(tv, &e) == 0.5f && tv >= 1.17549435e-38f && tv < INFINITY;
    const int tiles_x = (g.W + TL::TX - 1) / TL::TX, tiles_y = (g.H + TL::TY - 1) / TL::TY,
              tiles_z = (g.D + TL::TZ - 1) / TL::TZ;
    const int ntiles = tiles_x * tiles_y * tiles_z;
    const dim3 grid(xcd_row_grid(tiles_x, ntiles / tiles_x));
    const int N = (int)g.N;
    if (exact)
        hipLaunchKernelGGL((k_stencil3d<false, true, TL>), grid, dim3(TL::NT), 0, st, A, U, Aout, C, g.W, g.H, g.D, N,
                           tiles_x, tiles_y, ntiles, dt, beta_a, nexp, s, tv, 1.0f / tv, flag);
    else if (pow2)
        hipLaunchKernelGGL((k_stencil3d<true, false, TL>), grid, dim3(TL::NT), 0, st, A, U, Aout, C, g.W, g.H, g.D, N,
                           tiles_x, tiles_y, ntiles, dt, beta_a, nexp, s, tv, 1.0f / tv, flag);
    else
        hipLaunchKernelGGL((k_stencil3d<false, false, TL>), grid, dim3(TL::NT), 0, st, A, U, Aout, C, g.W, g.H, g.D, N,
                           tiles_x, tiles_y, ntiles, dt, beta_a, nexp, s, tv, 1.0f / tv, flag);
    return true;
}

// ---------------------------------------------------------------------------------------------
// General path.  (1) boxes: per (channel, source) the box [lo, hi] in the source's wrapped
// frame and the mass (inactive sources get an empty box), plus the radius every active source
// needs, via atomicMax.  (2) gather_general: per (channel, destination), every source in the
// (2R+1)^d window in ascending raster order, every periodic image of the destination inside
// the source's clamped index range (iz, iy, ix ascending, as the scatter loops).
// ---------------------------------------------------------------------------------------------
__device__ __forceinline__ void source_box(int dim, float m, int gxs, int gys, int gzs, float fx, float fy, float fz,
                                           float dt, float s, int W, int H, int D, float4& lo, float4& hi,
                                           int& need) {
    need = 0;
    if (fabsf(m) < 1e-10f) {
        lo = make_float4(INFINITY, INFINITY, INFINITY, 0.0f);
        hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.0f);
        return;
    }
    const float dxp = (float)gxs + dt * fx;
    const float dyp = (float)gys + dt * fy;
    const float dzp = dim == 3 ? (float)gzs + dt * fz : 0.0f;
    const float ta = dim == 3 ? (2.0f * s) * (2.0f * s) * (2.0f * s) : (2.0f * s) * (2.0f * s);
    if (ta < 1e-10f) {  // point fallback (reintegration.rs:117-124): all mass to round(dest)
        lo = make_float4(dxp, dyp, dzp, m);
        hi = make_float4(dxp, dyp, dzp, 1.0f);  // w = 1 marks point mode
        int rx = abs(sat_i32(roundf(dxp)) - gxs), ry = abs(sat_i32(roundf(dyp)) - gys);
        need = max(rx, ry);
        if (dim == 3) need = max(need, abs(sat_i32(roundf(dzp)) - gzs));
        return;
    }
    lo = make_float4(dxp - s, dyp - s, dzp - s, m);
    hi = make_float4(dxp + s, dyp + s, dzp + s, 0.0f);
    const int ixl = max(sat_i32(floorf(lo.x)), -W), ixh = min(sat_i32(ceilf(hi.x)), 2 * W);
    const int iyl = max(sat_i32(floorf(lo.y)), -H), iyh = min(sat_i32(ceilf(hi.y)), 2 * H);
    need = max(max(gxs - ixl, ixh - gxs), max(gys - iyl, iyh - gys));
    if (dim == 3) {
        const int izl = max(sat_i32(floorf(lo.z)), -D), izh = min(sat_i32(ceilf(hi.z)), 2 * D);
        need = max(need, max(gzs - izl, izh - gzs));
    }
    if (!(isfinite(lo.x) && isfinite(hi.x) && isfinite(lo.y) && isfinite(hi.y) && isfinite(lo.z) &&
          isfinite(hi.z) && isfinite(m)))
        need = -1;  // non-finite
}

__device__ __forceinline__ float asum_at(const float* A, int C, long long N, long long idx) {
    float s = 0.0f;
    for (int c = 0; c < C; ++c) s += A[(long long)c * N + idx];
    return s;
}

__device__ __forceinline__ void sobel27(const float* g, const float* A, int C, long long N, int W, int H, int D,
                                        int x, int y, int z, bool sum_channels, float& vx, float& vy, float& vz) {
    // gradient3d.rs:75-138 with the same grouping and the final / 32
    const int xp = (x + W - 1) % W, xn = (x + 1) % W, yp = (y + H - 1) % H, yn = (y + 1) % H;
    const int zp = (z + D - 1) % D, zn = (z + 1) % D;
    auto P = [&](int zz, int yy, int xx) -> float {
        const long long i = ((long long)zz * H + yy) * W + xx;
        return sum_channels ? asum_at(A, C, N, i) : g[i];
    };
    const float p_mmm = P(zp, yp, xp), p_mm0 = P(zp, yp, x), p_mmp = P(zp, yp, xn);
    const float p_m0m = P(zp, y, xp), p_m00 = P(zp, y, x), p_m0p = P(zp, y, xn);
    const float p_mpm = P(zp, yn, xp), p_mp0 = P(zp, yn, x), p_mpp = P(zp, yn, xn);
    const float p_0mm = P(z, yp, xp), p_0m0 = P(z, yp, x), p_0mp = P(z, yp, xn);
    const float p_00m = P(z, y, xp), p_00p = P(z, y, xn);
    const float p_0pm = P(z, yn, xp), p_0p0 = P(z, yn, x), p_0pp = P(z, yn, xn);
    const float p_pmm = P(zn, yp, xp), p_pm0 = P(zn, yp, x), p_pmp = P(zn, yp, xn);
    const float p_p0m = P(zn, y, xp), p_p00 = P(zn, y, x), p_p0p = P(zn, y, xn);
    const float p_ppm = P(zn, yn, xp), p_pp0 = P(zn, yn, x), p_ppp = P(zn, yn, xn);
    const float gx_zm = -1.0f * (1.0f * p_mmm + 2.0f * p_m0m + 1.0f * p_mpm) + 1.0f * (1.0f * p_mmp + 2.0f * p_m0p + 1.0f * p_mpp);
    const float gx_z0 = -1.0f * (1.0f * p_0mm + 2.0f * p_00m + 1.0f * p_0pm) + 1.0f * (1.0f * p_0mp + 2.0f * p_00p + 1.0f * p_0pp);
    const float gx_zp = -1.0f * (1.0f * p_pmm + 2.0f * p_p0m + 1.0f * p_ppm) + 1.0f * (1.0f * p_pmp + 2.0f * p_p0p + 1.0f * p_ppp);
    vx = (1.0f * gx_zm + 2.0f * gx_z0 + 1.0f * gx_zp) / 32.0f;
    const float gy_zm = -1.0f * (1.0f * p_mmm + 2.0f * p_mm0 + 1.0f * p_mmp) + 1.0f * (1.0f * p_mpm + 2.0f * p_mp0 + 1.0f * p_mpp);
    const float gy_z0 = -1.0f * (1.0f * p_0mm + 2.0f * p_0m0 + 1.0f * p_0mp) + 1.0f * (1.0f * p_0pm + 2.0f * p_0p0 + 1.0f * p_0pp);
    const float gy_zp = -1.0f * (1.0f * p_pmm + 2.0f * p_pm0 + 1.0f * p_pmp) + 1.0f * (1.0f * p_ppm + 2.0f * p_pp0 + 1.0f * p_ppp);
    vy = (1.0f * gy_zm + 2.0f * gy_z0 + 1.0f * gy_zp) / 32.0f;
    const float gz_xm = -1.0f * (1.0f * p_mmm + 2.0f * p_m0m + 1.0f * p_mpm) + 1.0f * (1.0f * p_pmm + 2.0f * p_p0m + 1.0f * p_ppm);
    const float gz_x0 = -1.0f * (1.0f * p_mm0 + 2.0f * p_m00 + 1.0f * p_mp0) + 1.0f * (1.0f * p_pm0 + 2.0f * p_p00 + 1.0f * p_pp0);
    const float gz_xp = -1.0f * (1.0f * p_mmp + 2.0f * p_m0p + 1.0f * p_mpp) + 1.0f * (1.0f * p_pmp + 2.0f * p_p0p + 1.0f * p_ppp);
    vz = (1.0f * gz_xm + 2.0f * gz_x0 + 1.0f * gz_xp) / 32.0f;
}

// grid-stride over C*N (channel-major)
__global__ void k_boxes(const float* __restrict__ A, const float* __restrict__ U, int C, int W, int H, int D,
                        long long N, float dt, float beta_a, float nexp, float s, float4* __restrict__ box_lo,
                        float4* __restrict__ box_hi, int* __restrict__ need_radius) {
    const long long total = (long long)C * N;
    int local = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t / N);
        const long long i = t - (long long)c * N;
        const int x = (int)(i % W);
        const int y = (int)((i / W) % H);
        const int z = (int)(i / ((long long)W * H));
        const float m = A[t];
        float fx, fy, fz = 0.0f;
        const float al = alpha_of(asum_at(A, C, N, i), beta_a, nexp);
        const float oma = 1.0f - al;
        if (D == 1) {
            const int xp = (x + W - 1) % W, xn = (x + 1) % W, yp = (y + H - 1) % H, yn = (y + 1) % H;
            const long long rp = (long long)yp * W, rc = (long long)y * W, rn = (long long)yn * W;
            float gax, gay, gux, guy;
            sobel9(asum_at(A, C, N, rp + xp), asum_at(A, C, N, rp + x), asum_at(A, C, N, rp + xn),
                   asum_at(A, C, N, rc + xp), asum_at(A, C, N, rc + xn), asum_at(A, C, N, rn + xp),
                   asum_at(A, C, N, rn + x), asum_at(A, C, N, rn + xn), gax, gay);
            const float* Uc = U + (long long)c * N;
            sobel9(Uc[rp + xp], Uc[rp + x], Uc[rp + xn], Uc[rc + xp], Uc[rc + xn], Uc[rn + xp], Uc[rn + x],
                   Uc[rn + xn], gux, guy);
            fx = oma * gux - al * gax;
            fy = oma * guy - al * gay;
        } else {
            float gax, gay, gaz, gux, guy, guz;
            sobel27(nullptr, A, C, N, W, H, D, x, y, z, true, gax, gay, gaz);
            sobel27(U + (long long)c * N, A, C, N, W, H, D, x, y, z, false, gux, guy, guz);
            fx = oma * gux - al * gax;
            fy = oma * guy - al * gay;
            fz = oma * guz - al * gaz;
        }
        float4 lo, hi;
        int need;
        source_box(D == 1 ? 2 : 3, m, x, y, z, fx, fy, fz, dt, s, W, H, D, lo, hi, need);
        box_lo[t] = lo;
        box_hi[t] = hi;
        local = (need < 0 || local < 0) ? -1 : max(local, need);
    }
    if (local < 0) atomicOr(need_radius + 1, 2);
    else atomicMax(need_radius, local);
}

__global__ void k_boxes_from_flow(const float* __restrict__ A, const float* __restrict__ fxv,
                                  const float* __restrict__ fyv, const float* __restrict__ fzv, int W, int H, int D,
                                  long long N, float dt, float s, float4* __restrict__ box_lo,
                                  float4* __restrict__ box_hi, int* __restrict__ need_radius) {
    int local = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % W);
        const int y = (int)((i / W) % H);
        const int z = (int)(i / ((long long)W * H));
        float4 lo, hi;
        int need;
        source_box(D == 1 ? 2 : 3, A[i], x, y, z, fxv[i], fyv[i], fzv ? fzv[i] : 0.0f, dt, s, W, H, D, lo, hi, need);
        box_lo[i] = lo;
        box_hi[i] = hi;
        local = (need < 0 || local < 0) ? -1 : max(local, need);
    }
    if (local < 0) atomicOr(need_radius + 1, 2);
    else atomicMax(need_radius, local);
}

// Ascending wrapped coordinates of the window [d - R, d + R] on an axis of n cells
// (all of [0, n) when the window covers the axis).
__device__ __forceinline__ int win_count(int R, int n) { return 2 * R + 1 >= n ? n : 2 * R + 1; }
__device__ __forceinline__ int win_at(int i, int d, int R, int n) {
    if (2 * R + 1 >= n) return i;
    const int lo = d - R, hi = d + R;
    if (lo >= 0 && hi < n) return lo + i;
    if (lo < 0) return i <= hi ? i : lo + n + (i - hi - 1);  // [0, hi] then [lo + n, n)
    const int a = hi - n + 1;                                 // [0, hi - n] then [lo, n)
    return i < a ? i : lo + (i - a);
}

// Sum of a box's overlap widths over its clamped index range [i0, i1] along one axis, ascending,
// in double (the exact-conservation scale).
__device__ double width_sum(int i0, int i1, float lo, float hi) {
    double acc = 0.0;
    for (int i = i0; i <= i1; ++i) acc += (double)overlap1((float)i, (float)(i + 1), lo, hi);
    return acc;
}

// exact: the exact-conservation mode of the fused kernels (q = m / (Sx Sy [Sz]) per source,
// double contributions and sum, one rounding per destination).
// rdev (may be null): the radius is read on the device from rdev[0] (the boxes' atomicMax);
// nothing is written when rdev[1] holds a flag, or when the radius exceeds rmax (then bit 2 of
// rdev[1] is set for the host to see).
__global__ void k_gather_general(const float4* __restrict__ box_lo, const float4* __restrict__ box_hi,
                                 float* __restrict__ Aout, int C, int W, int H, int D, long long N, float s, int R,
                                 int exact, int* __restrict__ rdev, int rmax) {
    if (rdev) {
        if (rdev[1] != 0) return;
        R = max(rdev[0], 1);
        if (R > rmax) {
            if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(rdev + 1, 4);
            return;
        }
    }
    const long long total = (long long)C * N;
    const bool is3 = D > 1;
    const float ta = is3 ? (2.0f * s) * (2.0f * s) * (2.0f * s) : (2.0f * s) * (2.0f * s);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int c = (int)(t / N);
        const long long i = t - (long long)c * N;
        const int dx = (int)(i % W);
        const int dy = (int)((i / W) % H);
        const int dz = (int)(i / ((long long)W * H));
        const float4* lo_c = box_lo + (long long)c * N;
        const float4* hi_c = box_hi + (long long)c * N;
        const int nzw = is3 ? win_count(R, D) : 1, nyw = win_count(R, H), nxw = win_count(R, W);
        float acc = 0.0f;
        double accd = 0.0;
        for (int a = 0; a < nzw; ++a) {
            const int sz = is3 ? win_at(a, dz, R, D) : 0;
            for (int b = 0; b < nyw; ++b) {
                const int sy = win_at(b, dy, R, H);
                for (int e = 0; e < nxw; ++e) {
                    const int sx = win_at(e, dx, R, W);
                    const long long si = ((long long)sz * H + sy) * W + sx;
                    const float4 lo = lo_c[si];
                    const float4 hi = hi_c[si];
                    if (!(lo.x <= hi.x)) continue;  // inactive source (empty box)
                    const float m = lo.w;
                    if (hi.w != 0.0f) {  // point mode: lo holds dest
                        const int nx = wrapi(sat_i32(roundf(lo.x)), W);
                        const int ny = wrapi(sat_i32(roundf(lo.y)), H);
                        const int nz = is3 ? wrapi(sat_i32(roundf(lo.z)), D) : 0;
                        if (nx == dx && ny == dy && nz == dz) {
                            if (exact) accd += (double)m;
                            else acc += m;
                        }
                        continue;
                    }
                    const int ixl = max(sat_i32(floorf(lo.x)), -W), ixh = min(sat_i32(ceilf(hi.x)), 2 * W);
                    const int iyl = max(sat_i32(floorf(lo.y)), -H), iyh = min(sat_i32(ceilf(hi.y)), 2 * H);
                    int izl = 0, izh = 0;
                    if (is3) {
                        izl = max(sat_i32(floorf(lo.z)), -D);
                        izh = min(sat_i32(ceilf(hi.z)), 2 * D);
                    }
                    double q = 0.0;
                    if (exact) {
                        double pw = width_sum(ixl, ixh, lo.x, hi.x) * width_sum(iyl, iyh, lo.y, hi.y);
                        if (is3) pw = pw * width_sum(izl, izh, lo.z, hi.z);
                        if (!(pw > 0.0)) continue;
                        q = (double)(float)((double)m / pw);
                    }
                    for (int kz = (is3 ? -1 : 0); kz <= (is3 ? 2 : 0); ++kz) {
                        const int iz = dz + kz * D;
                        if (is3 && (iz < izl || iz > izh)) continue;
                        float od = 1.0f;
                        if (is3) od = overlap1((float)iz, (float)(iz + 1), lo.z, hi.z);
                        for (int ky = -1; ky <= 2; ++ky) {
                            const int iy = dy + ky * H;
                            if (iy < iyl || iy > iyh) continue;
                            const float oh = overlap1((float)iy, (float)(iy + 1), lo.y, hi.y);
                            for (int kx = -1; kx <= 2; ++kx) {
                                const int ix = dx + kx * W;
                                if (ix < ixl || ix > ixh) continue;
                                const float ow = overlap1((float)ix, (float)(ix + 1), lo.x, hi.x);
                                if (exact) {
                                    const double ad = (double)ow * (double)oh;
                                    const double vd = is3 ? ad * (double)od : ad;
                                    if (vd > 0.0) accd += q * vd;
                                    continue;
                                }
                                const float vol = is3 ? ow * oh * od : ow * oh;
                                if (vol > 0.0f) {
                                    const float frac = vol / ta;
                                    acc += m * frac;
                                }
                            }
                        }
                    }
                }
            }
        }
        Aout[t] = exact ? (float)accd : acc;
    }
}

void launch_boxes(hipStream_t st, const float* A, const float* U, int C, const Geo& g, float dt, float beta_a,
                  float nexp, float s, float4* box_lo, float4* box_hi, int* need_radius) {
    const long long total = (long long)C * g.N;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 65535);
    hipLaunchKernelGGL(k_boxes, dim3(blocks), dim3(256), 0, st, A, U, C, g.W, g.H, g.D, g.N, dt, beta_a, nexp, s,
                       box_lo, box_hi, need_radius);
}

void launch_boxes_from_flow(hipStream_t st, const float* A, const float* fx, const float* fy, const float* fz,
                            const Geo& g, float dt, float s, float4* box_lo, float4* box_hi, int* need_radius) {
    const int blocks = (int)std::min<long long>((g.N + 255) / 256, 65535);
    hipLaunchKernelGGL(k_boxes_from_flow, dim3(blocks), dim3(256), 0, st, A, fx, fy, fz, g.W, g.H, g.D, g.N, dt, s,
                       box_lo, box_hi, need_radius);
}

void launch_gather_general(hipStream_t st, const float4* box_lo, const float4* box_hi, float* Aout, int C,
                           const Geo& g, float s, int radius, bool exact, int* rdev, int rmax) {
    const long long total = (long long)C * g.N;
    const int blocks = (int)std::min<long long>((total + 255) / 256, 65535);
    hipLaunchKernelGGL(k_gather_general, dim3(blocks), dim3(256), 0, st, box_lo, box_hi, Aout, C, g.W, g.H, g.D,
                       g.N, s, radius, exact ? 1 : 0, rdev, rmax);
}

// ---------------------------------------------------------------------------------------------
// Stage kernels behind the stage entry points.
// ---------------------------------------------------------------------------------------------
__global__ void k_sobel2d(const float* __restrict__ g, float* __restrict__ gx, float* __restrict__ gy, int W, int H) {
    const long long n = (long long)W * H;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % W), y = (int)(i / W);
        const int xp = (x + W - 1) % W, xn = (x + 1) % W, yp = (y + H - 1) % H, yn = (y + 1) % H;
        const long long rp = (long long)yp * W, rc = (long long)y * W, rn = (long long)yn * W;
        float vx, vy;
        sobel9(g[rp + xp], g[rp + x], g[rp + xn], g[rc + xp], g[rc + xn], g[rn + xp], g[rn + x], g[rn + xn], vx, vy);
        gx[i] = vx;
        gy[i] = vy;
    }
}

__global__ void k_flow2d(const float* __restrict__ gux, const float* __restrict__ guy, const float* __restrict__ gax,
                         const float* __restrict__ gay, const float* __restrict__ mass, long long n, float beta_a,
                         float nexp, float* __restrict__ fx, float* __restrict__ fy) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float al = alpha_of(mass[i], beta_a, nexp);
        const float oma = 1.0f - al;
        fx[i] = oma * gux[i] - al * gax[i];
        fy[i] = oma * guy[i] - al * gay[i];
    }
}

__global__ void k_growth_acc(const float* __restrict__ conv, float* __restrict__ target, long long n, float weight,
                             float mu, float inv) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const float d = conv[i] - mu;
        const float g = 2.0f * expf(-d * d * inv) - 1.0f;
        target[i] += weight * g;
    }
}

static int grid_for(long long n) { return (int)std::max<long long>(1, std::min<long long>((n + 255) / 256, 65535)); }

void launch_sobel2d(hipStream_t st, const float* grid, float* gx, float* gy, int W, int H) {
    hipLaunchKernelGGL(k_sobel2d, dim3(grid_for((long long)W * H)), dim3(256), 0, st, grid, gx, gy, W, H);
}

void launch_flow2d(hipStream_t st, const float* gux, const float* guy, const float* gax, const float* gay,
                   const float* mass, long long n, float beta_a, float nexp, float* fx, float* fy) {
    hipLaunchKernelGGL(k_flow2d, dim3(grid_for(n)), dim3(256), 0, st, gux, guy, gax, gay, mass, n, beta_a, nexp, fx, fy);
}

void launch_growth_acc(hipStream_t st, const float* conv, float* target, long long n, float weight, float mu,
                       float sigma) {
    const float sigma_sq_2 = 2.0f * sigma * sigma;  // growth.rs:51-52
    const float inv = 1.0f / sigma_sq_2;
    hipLaunchKernelGGL(k_growth_acc, dim3(grid_for(n)), dim3(256), 0, st, conv, target, n, weight, mu, inv);
}

// ---------------------------------------------------------------------------------------------
// FP64 mass: block b of channel c sums the fixed chunk [b*chunk, (b+1)*chunk) with a fixed
// per-thread stride and a fixed LDS tree; the second kernel adds the partials of each channel
// in a fixed tree and the channels in order.  Same bits on every run.
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mass_partial(const float* __restrict__ A, long long n, int nparts,
                                                      double* __restrict__ partials) {
    __shared__ double red[256];
    const int c = blockIdx.y, b = blockIdx.x;
    const long long chunk = (n + nparts - 1) / nparts;
    const long long beg = (long long)b * chunk;
    const long long end = beg + chunk < n ? beg + chunk : n;
    const float* a = A + (long long)c * n;
    double s = 0.0;
    for (long long i = beg + threadIdx.x; i < end; i += 256) s += (double)a[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[(long long)c * nparts + b] = red[0];
}

__global__ __launch_bounds__(256) void k_mass_final(const double* __restrict__ partials, int C, int nparts,
                                                    double* __restrict__ out) {
    __shared__ double red[256];
    double total = 0.0;
    for (int c = 0; c < C; ++c) {
        double s = 0.0;
        for (int i = threadIdx.x; i < nparts; i += 256) s += partials[(long long)c * nparts + i];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int w = 128; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[c] = red[0];
            total += red[0];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[C] = total;
}

void launch_mass_f64(hipStream_t st, const float* A, int C, long long n, double* partials, int nparts, double* out) {
    hipLaunchKernelGGL(k_mass_partial, dim3(nparts, C), dim3(256), 0, st, A, n, nparts, partials);
    hipLaunchKernelGGL(k_mass_final, dim3(1), dim3(256), 0, st, partials, C, nparts, out);
}

__global__ void k_clear_ints(int* p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0;
}

void launch_clear_ints(hipStream_t st, int* p, int n) { hipLaunchKernelGGL(k_clear_ints, dim3(1), dim3(64), 0, st, p, n); }

// ---------------------------------------------------------------------------------------------
// Population fitness: centre of mass of the channel sum of every member (fitness.rs:150-152,
// 196-218: cells with m > 0 contribute m, x*m, y*m), reduced in FP64 in a fixed order.  One
// workgroup per member; out[3*member + {0,1,2}] = (sum m, sum x*m, sum y*m).
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_batch_moments(const MemberDesc* __restrict__ mem, int cur,
                                                       double* __restrict__ out) {
    __shared__ double red[3][256];
    const MemberDesc& m = mem[blockIdx.x];
    const float* A = m.A[cur];
    const long long n = (long long)m.g.W * m.g.H;  // 2D members
    double s0 = 0.0, sx = 0.0, sy = 0.0;
    for (long long i = threadIdx.x; i < n; i += 256) {
        float v = 0.0f;
        for (int c = 0; c < m.C; ++c) v += A[(long long)c * m.g.N + i];  // channel_sum: 0 + A_0 + ...
        if (v > 0.0f) {
            s0 += (double)v;
            sx += (double)(i % m.g.W) * (double)v;
            sy += (double)(i / m.g.W) * (double)v;
        }
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = sx;
    red[2][threadIdx.x] = sy;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) out[3 * blockIdx.x + k] = red[k][0];
}

void launch_batch_moments(hipStream_t st, const MemberDesc* mem, int nmem, int cur, double* out) {
    hipLaunchKernelGGL(k_batch_moments, dim3(nmem), dim3(256), 0, st, mem, cur, out);
}

}  // namespace hp
