// stencil2d.hpp -- the fused 2D stencil's tile core (channel sum, Sobel, alpha, flow, box,
// overlaps and the gather reintegration of one tile), shared by the stencil kernels and the
// batched persistent step.  Every function here must be compiled without FP contraction: a file
// that includes it with contraction on elsewhere turns it off first (#pragma clang fp
// contract(off)), so that the tile reproduces the reference's f32 arithmetic operation for
// operation.
//
// Reference stages restated (paths relative to the reference repository):
//   channel sum            src/compute/propagator.rs:304-310  (0 + A_0 + A_1 + ...)
//   Sobel 2D               src/compute/gradient.rs:68-108
//   alpha, flow            src/compute/flow.rs:14-32, 68-87
//   reintegration          src/compute/reintegration.rs:60-174
#pragma once

#include <math.h>

#include "stencil_common.hpp"

namespace hp {

// Tile order over the 8 XCDs (workgroup b runs on XCD b % 8, each with its own L2).  Rows of
// tiles (tiles_x consecutive tiles along x) are dealt to the XCDs in turn: an XCD's L2 sees
// runs of x-neighbours, which share halo columns, and every XCD gets rows from the whole grid,
// so work that depends on the data (regions with no mass skip most of theirs) spreads evenly.
// Padding workgroups get a tile index past the end.
__device__ __forceinline__ int xcd_row_tile(int b, int tiles_x, int nrows) {
    const int j = b >> 3, q = j / tiles_x;
    const int row = (b & 7) + 8 * q, col = j - q * tiles_x;
    return row < nrows ? row * tiles_x + col : 0x7fffffff;
}
inline int xcd_row_grid(int tiles_x, int nrows) { return 8 * ((nrows + 7) / 8) * tiles_x; }

// The same with blocks of RB consecutive tile rows dealt to the XCDs in turn, column-major
// inside a block (the RB vertically adjacent tiles of one column run one after another), so
// the halo rows that vertically adjacent tiles share are fetched into one L2.  RB = 1 is
// xcd_row_tile.
template <int RB>
__device__ __forceinline__ int xcd_rows_tile(int b, int tiles_x, int nrows) {
    const int j = b >> 3, per = RB * tiles_x;
    const int q = j / per, r = j - q * per;
    const int col = r / RB, sub = r - col * RB;
    const int row = ((b & 7) + 8 * q) * RB + sub;
    return row < nrows ? row * tiles_x + col : 0x7fffffff;
}
template <int RB>
inline int xcd_rows_grid(int tiles_x, int nrows) { return 8 * RB * ((nrows + 8 * RB - 1) / (8 * RB)) * tiles_x; }

// ---------------------------------------------------------------------------------------------
// Fused 2D stencil.
//
// Window.  A source at x whose box [x + d - s, x + d + s] has |d| = |dt F| < 1 and s <= 1
// overlaps only cells x-2 .. x+1 (monotone f32 rounding keeps that true for the computed box),
// so a destination receives from sources at offsets o = source - dest in {-1, 0, 1, 2} per
// axis: 16 sources.  Any active source whose box reaches further sets flag bit 0 and the host
// reruns the step through the general path.
//
// Work split.  One workgroup owns a TX x TY destination tile.  All global reads of the tile
// (the state over the load region, the masses of the workgroup's sources for every channel,
// the first channel's affinity) are issued together up front; later channels' affinity is
// loaded into registers while the previous channel is gathered.  Per channel, every source of
// the (TX+3) x (TY+3) source region computes its flow, its box and its four x and four y
// overlaps with the destinations it can reach (in its own wrapped frame, exactly as the
// scatter does) into LDS.  Each thread then gathers a vertical strip of SR destinations: it
// walks the strip's source rows and, within a row, the four source columns, both in ascending
// wrapped order, so every destination adds its contributions in the scatter's source raster
// order -- also at the periodic edges, where the visiting order is rotated.
// ---------------------------------------------------------------------------------------------
// Tile shapes (the plan autotuner picks one per configuration): TX x TY destinations per
// workgroup of NT threads, SR = TX*TY/NT destinations per thread strip.
template <int TX_, int TY_, int NT_>
struct StTile {
    static constexpr int TX = TX_, TY = TY_, NT = NT_, SR = TX * TY / NT;
    static constexpr int SW = TX + 3, SH = TY + 3;  // source region: offsets -1 .. +2 around the tile
    static constexpr int LW = SW + 2, LH = SH + 2;  // + Sobel halo
    static constexpr int NSRC = SW * SH;
    static constexpr int NLOAD = LW * LH;
    static constexpr int LPT = (NLOAD + NT - 1) / NT;  // load-region cells per thread
    static constexpr int SPT = (NSRC + NT - 1) / NT;   // sources per thread
    static constexpr int SROWS = SR + 3;               // source rows of one strip
    // LDS: two affinity tiles, and per source its mass and four x and four y overlaps
    static constexpr int LDS = 4 * (2 * NLOAD + 9 * NSRC);
    static constexpr int WG_PER_CU = 163840 / LDS;  // 160 KiB of LDS per CU
    // waves per SIMD that LDS admits, at most 4: registers are capped so that they do not admit
    // fewer, and never below 128 per lane (a tighter cap makes the multi-channel bodies spill;
    // a body that needs fewer registers still runs as many waves as LDS admits)
    static constexpr int WAVES_PER_EU = WG_PER_CU * NT / 256 < 1 ? 1 : (WG_PER_CU * NT / 256 > 4 ? 4 : WG_PER_CU * NT / 256);
    static_assert(TX * TY == NT * SR && NT % TX == 0, "one strip per thread");
};
using StTile0 = StTile<64, 16, 512>;
using StTile1 = StTile<64, 8, 256>;
using StTile2 = StTile<32, 16, 256>;
using StTile3 = StTile<128, 8, 512>;
using StTile4 = StTile<64, 16, 256>;  // 4-destination strips: fewer source loads per destination
using StTile5 = StTile<128, 8, 256>;
constexpr int kStencilTiles = 6;

// CC: channel count known at compile time (1..4), or 0 for any count (masses re-read per channel).
//
// Per-thread work is laid out so that nothing channel-independent is recomputed per channel:
// the alpha phase leaves each owned source's 1 - alpha, alpha * dSum/dx and alpha * dSum/dy in
// registers (the thread that owns source i in the alpha phase owns it in every channel's source
// phase), and the four y overlaps are stored planar so that the gather reads only the ones a
// strip row uses.  The channel-sum tile shares memory with the second affinity buffer, which
// is first written after the alpha phase has read it.
//
// Empty regions: the reference skips every source with |m| < 1e-10, so a channel none of whose
// sources in the tile's region passes that test sends no mass into the tile and its outputs are
// +0, as the reference's.  Such a channel skips its source and gather phases (the test is
// uniform over the workgroup); a tile with no active channel stores zeros after its loads.
//
// CONS (exact-conservation mode, no reference counterpart): each source's mass is scaled once to
// q = m / (Sx * Sy), Sx and Sy the sums of its overlap widths along x and y (ascending cells, in
// double), rounded to f32; every destination adds q * (ow * oh) in double, in the same source
// order, and rounds its total once.  A source's contributions then sum to its mass up to two
// roundings, and the step conserves mass to FP64-oracle level; the reference's own arithmetic
// drifts by about 2e-6 per 1000 steps.
//
// Gather of one strip of SR destinations.  Source rows are visited in the order ROTY, ROTY + 1,
// ... wrapping at SROWS (the row rotation depends only on the strip's row, so it is a compile-time
// instance chosen per wave); source columns in the lane's own order, column i of strip row 0
// being source cb[i] with x overlap ob[i] (flat indices into the mass and x-overlap planes).  A
// destination adds its contributions in that order, which is ascending wrapped raster order of
// the sources.  With cb[i] = base + i every address is one base plus constants.
template <int ROTY, bool POW2, bool CONS, class TL>
__device__ __forceinline__ void gather_strip(const float* __restrict__ sM, const float* __restrict__ sOX,
                                             const float* __restrict__ sOY, const int (&cb)[4], const int (&ob)[4],
                                             float ta, float (&acc)[TL::SR], double (&accd)[CONS ? TL::SR : 1]) {
    constexpr int SW = TL::SW, SR = TL::SR, SROWS = TL::SROWS, NSRC = TL::NSRC;
#pragma unroll
    for (int it = 0; it < SROWS; ++it) {
        const int a = (it + ROTY) % SROWS;  // strip source row a <-> grid row gdy0 - 1 + a
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int si = cb[i] + a * SW;
            const float m = sM[si];
            const float ow = sOX[ob[i] + a * SW];
#pragma unroll
            for (int r = 0; r < SR; ++r) {
                const int jy = a - r;  // o_y + 1
                if (jy >= 0 && jy <= 3) {
                    if (CONS) {
                        accd[CONS ? r : 0] += (double)m * ((double)ow * (double)sOY[jy * NSRC + si]);
                    } else {
                        const float area = ow * sOY[jy * NSRC + si];
                        acc[r] += POW2 ? m * area : m * (area / ta);
                    }
                }
            }
        }
    }
}

// Gather of one strip that touches a periodic edge: source rows in the order roty, roty + 1, ...
// wrapping at SROWS, source columns in the order rotx, rotx + 1, ... mod 4 (both per lane, roty
// in [0, SROWS)), base the lane's source (strip row 0, column o = -1).  The loop over rows stays
// rolled; every (source, destination) pair is formed from in-bounds LDS entries and added only
// where the source row reaches the destination, so each destination adds exactly the pairs of
// the unrotated instance, in ascending wrapped raster order of the sources.
template <bool POW2, bool CONS, class TL>
__device__ __forceinline__ void gather_strip_edge(const float* __restrict__ sM, const float* __restrict__ sOX,
                                                  const float* __restrict__ sOY, int base, int rotx, int roty,
                                                  float ta, float (&acc)[TL::SR],
                                                  double (&accd)[CONS ? TL::SR : 1]) {
    constexpr int SW = TL::SW, SR = TL::SR, SROWS = TL::SROWS, NSRC = TL::NSRC;
    int cb[4], jn[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int jx = (i + rotx) & 3;  // plane jx = o + 1, o = source - dest
        cb[i] = base + jx;
        jn[i] = jx * NSRC;
    }
#pragma unroll 1
    for (int it = 0; it < SROWS; ++it) {
        int a = it + roty;  // strip source row a <-> grid row gdy0 - 1 + a
        if (a >= SROWS) a -= SROWS;
        const int ro = a * SW;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int si = cb[i] + ro;
            const float m = sM[si];
            const float ow = sOX[jn[i] + si];
#pragma unroll
            for (int r = 0; r < SR; ++r) {
                const int jy = a - r;  // o_y + 1
                const bool ok = (unsigned)jy <= 3u;
                const float oy = sOY[(ok ? jy : 0) * NSRC + si];
                if (CONS) {
                    const double v = (double)m * ((double)ow * (double)oy);
                    accd[CONS ? r : 0] = ok ? accd[CONS ? r : 0] + v : accd[CONS ? r : 0];
                } else {
                    const float area = ow * oy;
                    const float v = POW2 ? m * area : m * (area / ta);
                    acc[r] = ok ? acc[r] + v : acc[r];
                }
            }
        }
    }
}

// One source of a 2D tile: its flow from U's Sobel and the alpha terms, its box, its four x and
// four y overlaps with the destination cells it can reach (in its own wrapped frame, as the
// scatter computes them) and its scaled mass, into the LDS planes (plane stride NSRC).  i is the
// source's index in the (TX+3) x (TY+3) source region, (sy, sx) its row and column there, su the
// channel's affinity over the load region, (gx, gy) the source's wrapped grid cell.  A source the
// tile owns whose box leaves the 16-source window sets bit 0 of local_flag (bit 1 when a value is
// not finite).
template <bool POW2, bool CONS, class TL>
__device__ __forceinline__ void source_one(int i, int sy, int sx, float m, const float* __restrict__ su, float aOm,
                                           float aGx, float aGy, int gx, int gy, int x0, int y0, int W, int H,
                                           float dt, float s, float inv_ta, float* __restrict__ sM,
                                           float* __restrict__ sOX, float* __restrict__ sOY, int& local_flag) {
    constexpr int LW = TL::LW, NSRC = TL::NSRC, TX = TL::TX, TY = TL::TY;
    if (fabsf(m) < 1e-10f) {  // reintegration.rs:76: a skipped source adds nothing
        sM[i] = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) sOX[j * NSRC + i] = sOY[j * NSRC + i] = 0.0f;
        return;
    }
    const float* r0 = su + sy * LW + sx;
    const float* r1 = r0 + LW;
    const float* r2 = r1 + LW;
    float gux, guy;
    sobel9(r0[0], r0[1], r0[2], r1[0], r1[2], r2[0], r2[1], r2[2], gux, guy);
    const float fx = aOm * gux - aGx;
    const float fy = aOm * guy - aGy;
    // integer-valued floats below 2^24: (float)(g + k) == (float)g + k exactly
    const float cx = (float)gx, cy = (float)gy;
    const float dxp = cx + dt * fx;  // reintegration.rs:81-82
    const float dyp = cy + dt * fy;
    const float xm = dxp - s, xM = dxp + s, ym = dyp - s, yM = dyp + s;
    // overlaps with destination cells g + 1 - j, j = o + 1 (o = source - dest): cell
    // [c + 1 - j, c + 2 - j)
    float ovx[4], ovy[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ovx[j] = overlap_clamped(cx + (float)(1 - j), cx + (float)(2 - j), xm, xM);
        ovy[j] = overlap_clamped(cy + (float)(1 - j), cy + (float)(2 - j), ym, yM);
        sOX[j * NSRC + i] = ovx[j];
        sOY[j * NSRC + i] = ovy[j];
    }
    // With (2s)^2 a power of two the scaling by its inverse is exact (the masses skipped above are
    // the only ones below 1e-10, every nonzero overlap is a multiple of an ulp of s or of a
    // coordinate, so no product is subnormal before the last one): m * ((ow*oh) * inv) equals
    // (m * inv) * (ow*oh) bit for bit, and the gather saves a multiply per pair.  A scaled mass
    // that overflows is sent to the general path below.
    float sm;
    if (CONS) {  // widths summed over cells c-2 .. c+1, ascending
        const double wx = (((0.0 + (double)ovx[3]) + (double)ovx[2]) + (double)ovx[1]) + (double)ovx[0];
        const double wy = (((0.0 + (double)ovy[3]) + (double)ovy[2]) + (double)ovy[1]) + (double)ovy[0];
        const double pw = wx * wy;
        sm = pw > 0.0 ? (float)((double)m / pw) : 0.0f;
    } else {
        sm = POW2 ? m * inv_ta : m;
    }
    sM[i] = sm;
    // every source is checked by the one tile that owns it
    const int ux = x0 - 1 + sx, uy = y0 - 1 + sy;
    if (ux >= x0 && ux < x0 + TX && uy >= y0 && uy < y0 + TY && ux < W && uy < H) {
        // floor(xm) < c - 2  <=>  xm < c - 2 for integer c (likewise ceil); a non-finite value
        // fails the comparisons and is classified on the slow side
        const bool inside = xm >= cx - 2.0f && xM <= cx + 2.0f && ym >= cy - 2.0f && yM <= cy + 2.0f &&
                            (CONS || fabsf(sm) <= 3.40282347e38f);
        if (!inside)
            local_flag |= (isfinite(xm) && isfinite(xM) && isfinite(ym) && isfinite(yM) && isfinite(m)) ? 1 : 2;
    }
}

// The tile's LDS arrays, passed by pointer so that a kernel running several passes can lay them
// over one buffer (stencil2d_body owns static ones).
template <class TL>
struct St2Lds {
    float sU[2][TL::NLOAD];
    float sM[TL::NSRC];
    float sOX[4][TL::NSRC];
    float sOY[4][TL::NSRC];
    unsigned sAct[TL::NT / 64];
};

// PC / PW: the pointer types of the state and affinity reads and of the output writes -- plain
// pointers for kernel arguments, gptr<const float> / gptr<float> for pointers read from a
// member descriptor (global loads and stores rather than flat ones, whose outstanding count
// the LDS waits of the gather would otherwise include).
template <int CC, bool POW2, bool CONS, class TL, class PC = const float*, class PW = float*>
__device__ __forceinline__ void stencil2d_core(float (*__restrict__ sU)[TL::NLOAD], float* __restrict__ sM,
                                               float (*__restrict__ sOX)[TL::NSRC], float (*__restrict__ sOY)[TL::NSRC],
                                               unsigned* __restrict__ sAct, PC __restrict__ A,
                                               PC __restrict__ U, PW __restrict__ Aout, int Cdyn, int W,
                                               int H, long long N, int tiles_x, int tile, float dt, float beta_a,
                                               float nexp, float s, float ta, float inv_ta, int* __restrict__ flag) {
    constexpr int TX = TL::TX, TY = TL::TY, NT = TL::NT, SR = TL::SR, SW = TL::SW;
    constexpr int LW = TL::LW, NSRC = TL::NSRC, NLOAD = TL::NLOAD, LPT = TL::LPT, SPT = TL::SPT;
    constexpr int SROWS = TL::SROWS;
    const int C = CC > 0 ? CC : Cdyn;
    float* const sSum = sU[1];

    const int x0 = (tile % tiles_x) * TX, y0 = (tile / tiles_x) * TY;
    const int tid = threadIdx.x;
    // load-region coordinates lie in [-2, n + TX + 1): one wrap step suffices unless the grid is
    // narrower than the tile plus its halo
    const bool multi = W < TX + 2 || H < TY + 2;

    // ---- every global read of the tile, issued together ----
    int goff[LPT];
#pragma unroll
    for (int q = 0; q < LPT; ++q) {
        const int i = tid + q * NT;
        const int ly = i / LW, lx = i - ly * LW;
        goff[q] = i < NLOAD ? wrap_tile(y0 - 2 + ly, H, multi) * W + wrap_tile(x0 - 2 + lx, W, multi) : 0;
    }
    int sgx[SPT], sgy[SPT];  // wrapped grid coordinates of the owned sources
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int i = tid + q * NT;
        const int sy = i / SW, sx = i - sy * SW;
        sgx[q] = i < NSRC ? wrap_tile(x0 - 1 + sx, W, multi) : 0;
        sgy[q] = i < NSRC ? wrap_tile(y0 - 1 + sy, H, multi) : 0;
    }
    // Every global read of the load phase is issued before any of them is used: the owned
    // sources' masses, the first affinity and each channel's state over the load region (for a
    // channel count known at compile time; the run-time count reads the state in the sum loop).
    // Used as soon as loaded, they cost a memory latency per group.
    float msrc[CC > 0 ? CC : 1][SPT];
    float aval[CC > 0 ? CC : 1][LPT];
    float upre[LPT];
    if (CC > 0) {
#pragma unroll
        for (int c = 0; c < (CC > 0 ? CC : 1); ++c)
#pragma unroll
            for (int q = 0; q < SPT; ++q) msrc[c][q] = A[(long long)c * N + (long long)sgy[q] * W + sgx[q]];
    }
#pragma unroll
    for (int q = 0; q < LPT; ++q) upre[q] = U[goff[q]];
    if (CC > 0) {
#pragma unroll
        for (int c = 0; c < (CC > 0 ? CC : 1); ++c)
#pragma unroll
            for (int q = 0; q < LPT; ++q) aval[c][q] = A[(long long)c * N + goff[q]];
    }
    // bit c: channel c has a source in the region that the reference does not skip.  A channel
    // without one receives no mass anywhere in the tile.  (With the channel count known only at
    // run time the masses are not at hand and every channel counts as active.)
    unsigned act_t = CC > 0 ? 0u : ~0u;
    if (CC > 0) {
#pragma unroll
        for (int c = 0; c < (CC > 0 ? CC : 1); ++c)
#pragma unroll
            for (int q = 0; q < SPT; ++q)
                if (tid + q * NT < NSRC && active_mass(msrc[c][q])) act_t |= 1u << c;
        act_t = wave_or(act_t);
    }
    if ((tid & 63) == 0) sAct[tid >> 6] = act_t;
#pragma unroll
    for (int q = 0; q < LPT; ++q) {  // channel sum (propagator.rs:304-310): 0 + A_0 + A_1 + ...
        float sum = 0.0f;
        if (CC > 0) {
#pragma unroll
            for (int c = 0; c < (CC > 0 ? CC : 1); ++c) sum += aval[c][q];
        } else {
            for (int c = 0; c < C; ++c) sum += A[(long long)c * N + goff[q]];
        }
        const int i = tid + q * NT;
        if (i < NLOAD) {
            sSum[i] = sum;
            sU[0][i] = upre[q];
        }
    }
    // the next channel's affinity travels in registers ahead of its use: channel c + 1's is
    // loaded before channel c's source phase (its latency hides behind that phase and the
    // gather before it) and goes into the other LDS buffer after channel c's first barrier.  The
    // loads are unconditional (past the last channel they reread it) and are issued before the
    // previous gather's stores, so that nothing waits for the stores.
    {
        const auto Un = U + (long long)(C > 1 ? 1 : 0) * N;
#pragma unroll
        for (int q = 0; q < LPT; ++q) upre[q] = Un[goff[q]];
    }
    lds_barrier();
    unsigned act = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; ++w) act |= sAct[w];
    if (act == 0) {  // no mass reaches the tile: every output is +0, as the reference's
        const int gdx = x0 + tid % TX, gdy0 = y0 + (tid / TX) * SR;
        if (gdx < W) {
#pragma unroll
            for (int c = 0; c < (CC > 0 ? CC : 1); ++c) {
                const auto Oc = Aout + (long long)c * N;
#pragma unroll
                for (int r = 0; r < SR; ++r)
                    if (gdy0 + r < H) Oc[(long long)(gdy0 + r) * W + gdx] = 0.0f;
            }
        }
        return;
    }
    // mass gradient and alpha of the owned sources (flow.rs:68-87: F = (1-a) dU - a dSum).  The
    // loop stays rolled (alpha_of may expand to a powf); each thread parks its own results in
    // the overlap planes, which are free until the first source phase, and reads them back
    // into registers -- the same thread writes and reads every entry, so no barrier is needed.
#pragma unroll 1
    for (int i = tid; i < NSRC; i += NT) {
        const int sy = i / SW, sx = i - sy * SW;
        const float* r0 = sSum + sy * LW + sx;
        const float* r1 = r0 + LW;
        const float* r2 = r1 + LW;
        float gx, gy;
        sobel9(r0[0], r0[1], r0[2], r1[0], r1[2], r2[0], r2[1], r2[2], gx, gy);
        const float al = alpha_of(r1[1], beta_a, nexp);
        sOX[0][i] = 1.0f - al;
        sOX[1][i] = al * gx;
        sOX[2][i] = al * gy;
    }
    float aOm[SPT], aGx[SPT], aGy[SPT];
#pragma unroll
    for (int q = 0; q < SPT; ++q) {
        const int i = tid + q * NT;
        const bool ok = i < NSRC;
        aOm[q] = ok ? sOX[0][i] : 0.0f;
        aGx[q] = ok ? sOX[1][i] : 0.0f;
        aGy[q] = ok ? sOX[2][i] : 0.0f;
    }

    // strip geometry: destination column gdx, rows gdy0 .. gdy0 + SR - 1
    const int tx = tid % TX, ty0 = (tid / TX) * SR;
    const int gdx = x0 + tx, gdy0 = y0 + ty0;
    // ascending wrapped order: start with the first source column / row that does not wrap low
    const int rotx = gdx < 1 ? 1 - gdx : (gdx + 2 >= W ? 4 - (gdx + 2 - (W - 1)) : 0);
    int roty = gdy0 < 1 ? 1 - gdy0 : (gdy0 + SR + 1 >= H ? SROWS - (gdy0 + SR + 1 - (H - 1)) : 0);
    if (roty < 0 || roty >= SROWS) roty = 0;  // strips below the grid's last row store nothing

    int local_flag = 0;
#pragma unroll 1
    for (int c = 0; c < C; ++c) {
        const float* su = sU[c & 1];
        const bool live = (act >> (c & 31)) & 1u;  // uniform over the workgroup; all ones for CC = 0
        // ---- sources: flow, box, overlaps ----
#pragma unroll
        for (int q = 0; q < SPT; ++q) {
            const int i = tid + q * NT;
            if (i >= NSRC || !live) break;
            const int sy = i / SW, sx = i - sy * SW;
            float m;
            if (CC > 0) {
                m = 0.0f;
#pragma unroll
                for (int cc = 0; cc < (CC > 0 ? CC : 1); ++cc)
                    if (cc == c) m = msrc[cc][q];
            } else {
                m = A[(long long)c * N + (long long)sgy[q] * W + sgx[q]];
            }
            source_one<POW2, CONS, TL>(i, sy, sx, m, su, aOm[q], aGx[q], aGy[q], sgx[q], sgy[q], x0, y0, W, H, dt, s,
                                       inv_ta, sM, sOX[0], sOY[0], local_flag);
        }
        lds_barrier();
        if (c + 1 < C) {
#pragma unroll
            for (int q = 0; q < LPT; ++q)
                if (tid + q * NT < NLOAD) sU[(c + 1) & 1][tid + q * NT] = upre[q];
        }
        // ---- gather: strip of SR destinations ----
        // A zero-area pair contributes m * 0 = +-0; added to an accumulator that started at +0 it
        // changes no bit, so every pair is added without the reference's `area > 0` test.
        // Waves with no periodic edge in reach take the unrotated instance, whose LDS addresses
        // are one base plus constants; a wave touching an edge takes the instance of its row
        // rotation with per-lane column bases.  (Lanes past the grid's right edge gather
        // in-bounds LDS entries and store nothing.)
        float acc[SR];
        double accd[CONS ? SR : 1];
#pragma unroll
        for (int r = 0; r < SR; ++r) acc[r] = 0.0f;
#pragma unroll
        for (int r = 0; r < (CONS ? SR : 1); ++r) accd[r] = 0.0;
        if (live) {
            const int base = ty0 * SW + tx;  // source (strip row 0, column o = -1) of the strip
            if (__all(rotx == 0 && roty == 0 && gdx < W)) {
                const int cb[4] = {base, base + 1, base + 2, base + 3};
                const int ob[4] = {base, NSRC + base + 1, 2 * NSRC + base + 2, 3 * NSRC + base + 3};
                gather_strip<0, POW2, CONS, TL>(sM, sOX[0], sOY[0], cb, ob, ta, acc, accd);
            } else if (!CONS && CC >= 1 && CC <= 3 && __all(roty == 0 && gdx < W)) {
                // an x edge only (every wave of an edge tile column; all of them on grids up to
                // two tiles wide): the unrolled gather with the lane's rotated column bases (the
                // exact-conservation and four-or-more-channel bodies, already at 128 registers,
                // would spill and keep the rolled gather)
                int cb[4], ob[4];
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int jx = (i + rotx) & 3;
                    cb[i] = base + jx;
                    ob[i] = jx * NSRC + base + jx;
                }
                gather_strip<0, POW2, CONS, TL>(sM, sOX[0], sOY[0], cb, ob, ta, acc, accd);
            } else {
                gather_strip_edge<POW2, CONS, TL>(sM, sOX[0], sOY[0], base, rotx, roty, ta, acc, accd);
            }
        }
        if (CONS) {
#pragma unroll
            for (int r = 0; r < SR; ++r) acc[r] = (float)accd[CONS ? r : 0];
        }
        {  // channel c + 2's affinity (see above)
            const auto Un = U + (long long)(c + 2 < C ? c + 2 : C - 1) * N;
#pragma unroll
            for (int q = 0; q < LPT; ++q) upre[q] = Un[goff[q]];
        }
        if (gdx < W) {
            const auto Oc = Aout + (long long)c * N;
#pragma unroll
            for (int r = 0; r < SR; ++r)
                if (gdy0 + r < H) Oc[(long long)(gdy0 + r) * W + gdx] = acc[r];
        }
        lds_barrier();  // the next channel's source phase overwrites the overlaps
    }
    if (local_flag) atomicOr(flag, local_flag);
}

}  // namespace hp
