// hp_common.hpp -- shared device/host declarations of the Flow Lenia HIP library.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

namespace hp {

// Work item of the spectral pass: one source spectrum and a run of kernels that
// read it (kernel indices klist[k_first .. k_first + k_count)).
struct SpecWork {
    int src_slot;
    int k_first;
    int k_count;
    int pad;
};

// Work item of the inverse row pass: one target channel and, in kernel-index
// order, the kernels that accumulate into it.
struct TgtWork {
    int target;
    int k_first;
    int k_count;
    int pad;
};

// Growth parameters of one kernel: weight, mu, 1/(2 sigma^2) (growth.rs:51-52), and sigma
// itself for the adjoint's d/dsigma.
struct GrowthParams {
    float weight;
    float mu;
    float inv;
    float sigma;
};

// One kernel of the direct convolution path: source and target channel, actual radius R, its
// (2R+1)^dim taps at taps[tap_off] (raster order, x fastest), its per-column nonzero ranges along
// the last axis at cols[col_off] ((2R+1)^(dim-1) entries), and its growth parameters.
struct DirectKernel {
    int src, tgt, R, tap_off, col_off;
    float weight, mu, inv;
};

// Shapes of one 2D/3D grid and its half spectrum.
struct Geo {
    int W, H, D;      // grid
    int Wh;           // W/2 + 1 complex columns of the half spectrum
    int Wp;           // padded row pitch (complex elements) of a spectrum row
    int Hq, Dq;       // H/2 + 1, D/2 + 1 (folded kernel-spectrum extents)
    long long N;      // W*H*D
    long long spec;   // complex elements of one spectrum: D*H*Wp
};

// Tunable shape of the LDS FFT passes: radix order (0: radix 8 first, 1: radix 4 first) and
// the split of the spectral pass's column block (1 or 2).
struct FftVariant {
    int rp = 0;
    int cs = 1;
};

// One member of a batched population: its buffers and geometry, in device memory.
struct MemberDesc {
    float* A[2];                 // state ping-pong
    float* U;                    // affinity
    float2* X;                   // source half spectra (row-major, pitch Wp)
    float2* Y;                   // per-kernel filtered spectra
    const float* Kq;             // folded kernel spectra
    const int* src_chan;         // source slot -> channel
    const SpecWork* spec_work;   // spectral work items
    const int* klist_spec;
    const TgtWork* tgt_work;     // inverse-row work items
    const int* klist_tgt;
    const GrowthParams* gp;
    const float2* tw_w;          // twiddles of length W
    const float2* tw_h;          // twiddles of length H
    Geo g;
    int C, K, nsrc, ntgt, nspec;
    float dt, beta_a, nexp, s;
    // gradient tape (null unless the member carries one)
    const int2* kmap;            // per kernel (source channel, target channel)
    float* lam[2];               // dL/dA ping-pong, C*N each
    float* conv;                 // convolution outputs, K*N per slot (one slot, or one per taped step)
    float* tape_u;               // U of each taped step, C*N per slot (with stored convolutions)
    float2* tape_x;              // source half spectra of each taped step, nsrc*spec per slot (likewise)
    float* gscr;                 // (2C+3)*N: dL/d(grad U_c) x, y per channel; dL/d(grad A_sum) x, y;
                                 //   dL/dA_sum through alpha
    float* lamU;                 // C*N: dL/dU
    float2* P;                   // K*spec: sum over steps of FFT(g_k) * conj(FFT(A_src)), pitch Wp
    double* pacc;                // K*pairs*3: FP64 sums of dL/d(weight, mu, sigma) per row pair
    float* tape;                 // checkpoint and segment states, C*N each
    double* dkwin;               // per kernel (2R+1)^2 FP64 derivatives of the kernel taps
    const int2* dkw;             // per kernel (offset of its window in dkwin, radius R)
    float* snap;                 // C*N snapshot taken before a multi-step advance
    long long stage_off;         // this member's offset in the contiguous host-transfer staging
};

// Pointers into global memory (HBM).  A pointer read from a descriptor in memory is generic to
// the compiler, which then accesses it with flat loads and stores: those count against the LDS
// counter as well, so every wait for an LDS read also waits for them; they take a full 64-bit
// address per lane; and since a flat access may touch LDS, the compiler keeps each such load
// behind the LDS stores before it (a loop that loads from memory and stores to LDS then waits
// for every load in turn).  Accesses through gptr compile to global loads and stores.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gptr<T> gp(T* p) {
    return (gptr<T>)p;
}
template <class T>
__device__ __forceinline__ gptr<const T> gpc(const T* p) {
    return (gptr<const T>)p;
}

// float2 through a gptr: HIP's float2 is a class whose assignment operators take a generic
// `this`, so its values move through the two-float vector type.  Plain pointers as usual.
using f2v = float __attribute__((ext_vector_type(2)));
__device__ __forceinline__ float2 ld2(const float2* p) { return *p; }
__device__ __forceinline__ float2 ld2(gptr<const float2> p) {
    const f2v v = *(gptr<const f2v>)p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ float2 ld2(gptr<float2> p) { return ld2((gptr<const float2>)p); }
__device__ __forceinline__ void st2(float2* p, float2 a) { *p = a; }
__device__ __forceinline__ void st2(gptr<float2> p, float2 a) {
    f2v v;
    v.x = a.x;
    v.y = a.y;
    *(gptr<f2v>)p = v;
}

// XCD-aware item order for batched passes whose neighbouring items share cache lines (column
// blocks narrower than a 128-byte line, tiles that read their neighbours' rows as halo).
// Workgroup b runs on XCD b % 8.  Items go to the XCDs in chunks of CH consecutive items,
// dealt round-robin: a chunk runs on one XCD at about the same time, so the lines it shares
// are fetched into one L2 once, and the XCDs get near-equal shares of the list.  The grid is
// xcd_chunk_grid<CH>(nitems) workgroups; a workgroup past the last item gets -1.
template <int CH>
__device__ __forceinline__ int xcd_chunk_item(int nitems) {
    const int b = (int)blockIdx.x, x = b % 8, t = b / 8;
    const int lin = ((t / CH) * 8 + x) * CH + t % CH;
    return lin < nitems ? lin : -1;
}
template <int CH>
inline int xcd_chunk_grid(int nitems) {
    return (nitems + 8 * CH - 1) / (8 * CH) * (8 * CH);
}

// ---- launchers implemented in fft_kernels.hip -------------------------------------
// Each returns false if the length is not one of the compiled FFT sizes.
bool fft_supported(int L);
bool launch_rows_fwd(int L, hipStream_t st, const float* A, float2* X, const int* src_chan, int nsrc,
                     const Geo& g, const float2* tw, const FftVariant& v = FftVariant());
bool launch_cols(int L, hipStream_t st, int dir, float2* X, int nspec, const Geo& g, const float2* tw,
                 const FftVariant& v = FftVariant());
bool launch_spectral(int L, int dim, hipStream_t st, const float2* X, float2* Y, const float* Kq,
                     const SpecWork* work, int nwork, const int* klist, const Geo& g, const float2* tw,
                     const FftVariant& v = FftVariant());
bool launch_rows_inv(int L, hipStream_t st, const float2* Y, float* U, const TgtWork* work, int nwork,
                     const int* klist, const GrowthParams* gp, const Geo& g, const float2* tw,
                     const FftVariant& v = FftVariant());
// The same inverse pass writing every kernel's convolution to conv + k*N, with no growth.
bool launch_rows_inv_conv(int L, hipStream_t st, const float2* Y, float* conv, const TgtWork* work, int nwork,
                          const int* klist, const Geo& g, const float2* tw, const FftVariant& v = FftVariant());

// Batched 2D passes over a size class: items[i] = (member, block, slot/work, -).
bool launch_batch_rows_fwd(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur,
                           const FftVariant& v = FftVariant());
bool launch_batch_spectral(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems,
                           const FftVariant& v = FftVariant());
// conv: also store each kernel's convolution into MemberDesc::conv slot cslot (K*N each),
// for the adjoint.
bool launch_batch_rows_inv(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems,
                           bool conv = false, int cslot = 0, const FftVariant& v = FftVariant());
// columns per spectral workgroup for length-L columns
int spectral_block_lines(int L, const FftVariant& v = FftVariant());

// ---- direct convolution for sides the FFT passes do not cover: direct_conv.hip -------------
// Every kernel's convolution into conv + k*N (K*N floats), then, with U non-null, U from them
// (growth, kernel order).  Returns false if the largest radius Rmax does not fit the LDS tile.
bool launch_direct_affinity(hipStream_t st, const float* A, float* U, float* conv, const DirectKernel* kd,
                            const float* taps, const int2* cols, int K, int C, const Geo& g, int Rmax);
int direct_max_radius(int dim);

// ---- launchers implemented in wave_fft.hip (2D, sides 256..2048, column-major spectra) ----
bool wave_fft_supported(int W, int H);
// Wave path: tw is the length-L table exp(-2 pi i m / L) of the pass's line length L (W for the
// row passes, H for the spectral pass); tws holds the per-stage tables of the L/2-point
// transforms, entry i = tw[2 * wave_stage_twiddle_index(L / 2)[i]].
std::vector<int> wave_stage_twiddle_index(int M);
bool launch_rows_fwd_t(hipStream_t st, const float* A, float2* XT, const int* src_chan, int nsrc, const Geo& g,
                       const float2* tw, const float2* tws);
bool launch_spectral_t(hipStream_t st, const float2* XT, float2* YT, const float* KqT, const SpecWork* work,
                       int nwork, const int* klist, const Geo& g, const float2* tw, const float2* tws);
bool launch_rows_inv_t(hipStream_t st, const float2* YT, float* U, const TgtWork* work, int nwork, const int* klist,
                       const GrowthParams* gp, const Geo& g, const float2* tw, const float2* tws);

// ---- launchers implemented in stencil_kernels.hip ----------------------------------
// Fused channel sum + Sobel + flow + gather reintegration with a fixed gather
// radius; sets *flag when any active source's box leaves that radius (bit 0) or
// is not finite (bit 1).  Returns false if the configuration is not covered.
bool stencil2d_supported(const Geo& g, float s, int tile = 0);
// Batched fused stencil: items[i] = (member, tile); members of one launch share C and the
// power-of-two-area property.
bool launch_batch_stencil2d(hipStream_t st, int C, bool pow2, const MemberDesc* mem, const int4* items, int nitems,
                            int cur, int* flag, int tile = 0);
// Tile shapes of the fused stencil: variant t in [0, stencil2d_tile_count()).
int stencil2d_tile_count();
int stencil2d_tile_w(int tile = 0);
// ---- the persistent population step (batch_step.hip) --------------------------------
// One launch per step over four phases' item lists (rows_fwd (member, pair, source), spectral
// (member, column block, work), rows_inv (member, pair, work), stencil (member, tile)), each
// phase's items sorted into segments by class (batch_step_persistent_segment), starts[0 ..
// segments] bounding the segments, with a bounded grid barrier between phases (a barrier that
// waits past spin_limit polls sets bit 2 of flags[0]).  Members' sides must be among the
// compiled lengths; the FFT passes run with radix order 0.
bool batch_step_persistent_size(int L);
int batch_step_persistent_segment(int phase, int L, int C, bool pow2);
int batch_step_persistent_segments();
int batch_step_persistent_tile_w();
int batch_step_persistent_tile_h();
int batch_step_persistent_block_lines(int L);  // spectral column block of length L
int batch_step_persistent_grid();              // workgroups the device holds at once
void launch_batch_step_persistent(hipStream_t st, const MemberDesc* mem, const int4* items, const int* starts,
                                  int cur, int* flags, unsigned* bar, int grid, int spin_limit);
// Per member (sum m, sum x m, sum y m) of the channel sum over cells with m > 0, FP64.
void launch_batch_moments(hipStream_t st, const MemberDesc* mem, int nmem, int cur, double* out);
int stencil2d_tile_h(int tile = 0);
// exact: the exact-conservation mode (FP64 contributions and sums; not the reference's arithmetic)
bool launch_stencil2d(hipStream_t st, const float* A, const float* U, float* Aout, int C, const Geo& g,
                      float dt, float beta_a, float nexp, float s, int* flag, int tile = 0, bool exact = false);
bool stencil3d_supported(const Geo& g, float s);
bool launch_stencil3d(hipStream_t st, const float* A, const float* U, float* Aout, int C, const Geo& g,
                      float dt, float beta_a, float nexp, float s, int* flag, bool exact = false);
// General path: per-source destination boxes (+ needed radius via atomicMax)...
void launch_boxes(hipStream_t st, const float* A, const float* U, int C, const Geo& g, float dt, float beta_a,
                  float nexp, float s, float4* box_lo, float4* box_hi, int* need_radius);
// ...then a gather over any radius with full periodic-image handling.  With rdev the radius is
// read on the device from rdev[0]; a flag in rdev[1] or a radius above rmax (which sets bit 2
// of rdev[1]) leaves Aout unwritten.
void launch_gather_general(hipStream_t st, const float4* box_lo, const float4* box_hi, float* Aout, int C,
                           const Geo& g, float s, int radius, bool exact = false, int* rdev = nullptr,
                           int rmax = 0);
// Stage kernels exposed for the stage entry points.
void launch_sobel2d(hipStream_t st, const float* grid, float* gx, float* gy, int W, int H);
void launch_flow2d(hipStream_t st, const float* gux, const float* guy, const float* gax, const float* gay,
                   const float* mass, long long n, float beta_a, float nexp, float* fx, float* fy);
void launch_growth_acc(hipStream_t st, const float* conv, float* target, long long n, float weight, float mu,
                       float sigma);
// Boxes from given flow fields (for the advect stage entry point).
void launch_boxes_from_flow(hipStream_t st, const float* A, const float* fx, const float* fy, const float* fz,
                            const Geo& g, float dt, float s, float4* box_lo, float4* box_hi, int* need_radius);
// FP64 mass: per-channel partial sums in a fixed order, then a fixed-order final sum.
void launch_mass_f64(hipStream_t st, const float* A, int C, long long n, double* partials, int nparts,
                     double* out);
// Clear a small int buffer with a kernel (graph-capture safe, no memset node).
void launch_clear_ints(hipStream_t st, int* p, int n);

// ---- embedded-parameter step: embedded.hip ------------------------------------------------
// Parameter planes P[c][f][N], f = mu, sigma, weight, beta_a, n; AoS = 5 floats per cell.
void launch_emb_split(hipStream_t st, const float* aos, float* planes, int C, long long N);
void launch_emb_join(hipStream_t st, const float* planes, float* aos, int C, long long N);
// U[t] from the convolutions conv[k] (K x N) and the targets' per-cell growth parameters.
void launch_emb_growth(hipStream_t st, const float* conv, const float* P, const int* ktarget, float* U, int C, int K,
                       long long N);
// F[c][2][N] and the per-channel largest |F| as float bits (maxbits cleared to 0 first).
void launch_emb_flow(hipStream_t st, const float* A, const float* U, const float* P, float* F, unsigned* maxbits,
                     int C, int W, int H);
// The largest |F| per channel of given flows F[c][2][N] (maxbits cleared to 0 first).
void launch_emb_maxmag(hipStream_t st, const float* F, unsigned* maxbits, int C, long long N);
// Gather with parameter mixing; a radius of rmax or more sets err[0] bit 0 and writes nothing.
void launch_emb_advect(hipStream_t st, const float* A, const float* P, const float* F, const unsigned* maxbits,
                       float* Aout, float* Pout, int C, int W, int H, float dt, float s, bool linear,
                       float temperature, int rmax, int* err);

// ---- population fitness statistics: fitness.hip ---------------------------------------------
// Per (sample, member): sum of the channel sum over all cells; over cells with m > 0 the sums
// of m, x*m, y*m, (x^2 + y^2)*m; the largest value (fold from 0); cells above 1e-6; cells below
// 0; the sum of squares; the sum of left and upper neighbour differences (final state only).
constexpr int kFitStats = 10;
// items = (member, part, nparts, -); part receives nitems * kFitStats sums.  snap_base (may be
// null): per member the base of its snapshot slots (W*H floats each), written at `slot`.
void launch_fit_stats(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur, bool with_grad,
                      float* const* snap_base, int slot, double* part);
// out[m * kFitStats + k]: member m's items first[m] .. first[m + 1] - 1 reduced in order.
void launch_fit_reduce(hipStream_t st, const int* first, int nmem, const double* part, double* out);
// pairs = (member, slot a, slot b, -): out[i] = sum over the member's cells of a * b (FP64).
void launch_fit_dots(hipStream_t st, const MemberDesc* mem, const int4* pairs, int npairs, float* const* snap_base,
                     double* out);

// ---- gradient tape: adjoint_stencil.hip ----------------------------------------------------
// items = (member, tile) over adj_tile_w() x adj_tile_h() tiles.
void launch_adj_flow(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur, int li);
void launch_adj_div(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int li);
int adj_tile_w();
int adj_tile_h();
// items = (member, part, nparts).  to_tape: tape[slot] <- A[cur], else A[cur] <- tape[slot];
// slot < 0 addresses MemberDesc::snap instead of the tape.
void launch_tape_copy(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur, int slot,
                      bool to_tape);
void launch_tape_clear(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems);
// items as launch_tape_copy.  to_stage: stage[stage_off + j] <- lam[li][j], else the reverse.
void launch_stage_lam(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, float* stage, int li,
                      bool to_stage);
// items as launch_tape_copy.  to_tape: tape_u[slot] <- U, else U <- tape_u[slot].
void launch_u_copy(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int slot, bool to_tape);
// Every member's A[0] -> tape[slot], U -> tape_u[u_slot] and X -> tape_x[u_slot] (descriptor
// pointers only; a negative slot leaves those pointers).
void launch_point_slots(hipStream_t st, MemberDesc* mem, int n, int slot, int u_slot);
// ---- gradient tape: adjoint_fft.hip ---------------------------------------------------------
bool adj_fft_supported(int L);
// items = (member, row pair, kernel); reads conv slot cslot
bool launch_adj_rows_fwd(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cslot);
// items = (member, column block of adj_spectral_block_lines(L) columns, source work item)
bool launch_adj_spectral(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems);
int adj_spectral_block_lines(int L);
// items as the forward row pass (member, row pair, source slot)
bool launch_adj_rows_inv(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int li);
// items = (member, kernel)
// The kernel-tap windows: row_items (member, kernel, row block, poff) write each block's window to
// part + poff + block * (2R+1)^2; sum_items (member, kernel, blocks, poff) add them into dkwin.
void launch_window_dft(hipStream_t st, const MemberDesc* mem, const int4* row_items, int nrow_items,
                       const int4* sum_items, int nsum_items, double* part);
int window_dft_rows();  // spectrum rows per block
int window_dft_max_side();
int window_dft_max_taps();

}  // namespace hp
