// fft_passes.hpp -- the LDS Stockham FFT passes of the Flow Lenia convolution (templates).
//
// The reference convolves every kernel with a full complex 2D/3D FFT of the
// source channel (CachedConvolver::convolve_with_kernel_scratch, fft.rs:288-324;
// fft3d.rs:308-332) and then applies growth_accumulate per kernel in kernel order
// (propagator.rs:270-279, growth.rs:44-59).  Here the same mathematics runs as
// three memory passes over a half spectrum (the input is real):
//   rows_fwd  : two real rows packed as one complex line, FFT along x, untangled
//               into two Hermitian half rows (R2C);
//   spectral  : FFT along the last axis (y in 2D, z in 3D), product with the
//               kernel spectrum -- real and even in every axis because the ring
//               kernel is point-symmetric (kernel.rs:41-64), so stored folded as
//               (D/2+1)(H/2+1)(W/2+1) floats -- and the inverse FFT along the same
//               axis, written per kernel;
//   rows_inv  : Hermitian completion of two half rows, inverse FFT along x,
//               Re * 1/(W*H*D) (fft.rs:389-392), growth G(u) and the weighted
//               accumulate of every kernel that targets the channel, in kernel
//               index order, all in registers; one write of the affinity U.
// 3D adds a forward y pass after rows_fwd and an inverse y pass before rows_inv.
//
// Every transform is an in-place Stockham autosort FFT in LDS: each stage loads
// its radix-R butterfly inputs into registers, synchronises, and stores the
// outputs; radices 8/4/2/3 cover lengths 2^a * 3^b.  Twiddles come from a table
// exp(-2 pi i m / L) (host FP64, rounded to f32) staged in LDS.
//
// The kernel templates are instantiated by the fft_g*.hip files, one group of lengths each
// (HP_DEFINE_FFT_LAUNCHERS), so that the many (length, radix order, column split) variants
// compile in parallel; fft_kernels.hip dispatches a length to its group.
#pragma once

#include "fft_lds.hpp"

namespace hp {


// ---- pass 1: R2C along x on pairs of real rows ----------------------------------------
// grid (ceil(rows/2), nsrc); rows = H*D; src_chan[s] = channel of source slot s.
// The pass bodies take their LDS arrays as pointers (the *_core functions), so that a kernel that
// runs several passes can lay them over one buffer; the *_body wrappers own static arrays.
template <int L, int NT, int RP>
__device__ __forceinline__ void rows_fwd_core(float2* __restrict__ buf, float2* __restrict__ tw,
                                              const float* __restrict__ A, float2* __restrict__ X,
                                              const int* __restrict__ src_chan, int rows, int Wp, long long N,
                                              long long spec, const float2* __restrict__ twg, int pair, int s) {
    const int y0 = 2 * pair;
    const bool has1 = (y0 + 1) < rows;
    const float* a = A + (long long)src_chan[s] * N + (long long)y0 * L;
    stage_twiddles<NT, L>(tw, twg);
    // the row pair's loads first, then the LDS stores: in one loop a flat load (the batched
    // kernels' pointers come from a descriptor) waits behind the previous LDS store
    constexpr int PER = (L + NT - 1) / NT;
    float va[PER], vb[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        if (L % NT == 0 || i < L) {
            va[q] = a[i];
            vb[q] = has1 ? a[L + i] : 0.0f;
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        if (L % NT == 0 || i < L) buf[i] = make_float2(va[q], vb[q]);
    }
    __syncthreads();
    lds_fft<L, -1, 1, NT, L, RP>(buf, tw);
    float2* x0 = X + (long long)s * spec + (long long)y0 * Wp;
    float2* x1 = x0 + Wp;
    for (int k = threadIdx.x; k <= L / 2; k += NT) {
        const float2 zk = buf[k];
        const float2 zm = buf[k == 0 ? 0 : L - k];
        x0[k] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        if (has1) x1[k] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
    }
}

template <int L, int NT, int RP>
__device__ __forceinline__ void rows_fwd_body(const float* __restrict__ A, float2* __restrict__ X,
                                              const int* __restrict__ src_chan, int rows, int Wp, long long N,
                                              long long spec, const float2* __restrict__ twg, int pair, int s) {
    __shared__ float2 buf[L];
    __shared__ float2 tw[L];
    rows_fwd_core<L, NT, RP>(buf, tw, A, X, src_chan, rows, Wp, N, spec, twg, pair, s);
}

template <int L, int NT, int RP>
__global__ __launch_bounds__(NT) void k_rows_fwd(const float* __restrict__ A, float2* __restrict__ X,
                                                 const int* __restrict__ src_chan, int rows, int Wp,
                                                 long long N, long long spec, const float2* __restrict__ twg) {
    rows_fwd_body<L, NT, RP>(A, X, src_chan, rows, Wp, N, spec, twg, blockIdx.x, blockIdx.y);
}

// batched: item = (member, row pair, source slot)
template <int L, int NT, int RP>
__global__ __launch_bounds__(NT) void kb_rows_fwd(const MemberDesc* __restrict__ mem, const int4* __restrict__ items,
                                                  int cur) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    rows_fwd_body<L, NT, RP>(m.A[cur], m.X, m.src_chan, m.g.H * m.g.D, m.g.Wp, m.g.N, m.g.spec, m.tw_w, it.y, it.z);
}

// ---- 3D only: FFT along y, in place, forward or inverse ------------------------------------
// grid (ceil(Wh/CB), D, nspec).  Lines are the kx columns of one z plane.
template <int L, int CB, int NT, int DIR, int RP>
__global__ __launch_bounds__(NT) void k_cols(float2* __restrict__ X, int Wh, int Wp, long long spec,
                                             const float2* __restrict__ twg) {
    constexpr int LP = L + 1;
    __shared__ float2 buf[CB * LP];
    __shared__ float2 tw[L];
    const int kx0 = blockIdx.x * CB;
    float2* base = X + (long long)blockIdx.z * spec + (long long)blockIdx.y * L * Wp;
    stage_twiddles<NT, L>(tw, twg);
    constexpr int PER = (CB * L + NT - 1) / NT;
    float2 v[PER];  // every load first, then the LDS stores (see rows_fwd_core)
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        const int c = i % CB, e = i / CB;
        const int kx = kx0 + c;
        if ((CB * L) % NT == 0 || i < CB * L)
            v[q] = kx < Wh ? base[(long long)e * Wp + kx] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        if ((CB * L) % NT == 0 || i < CB * L) buf[(i % CB) * LP + i / CB] = v[q];
    }
    __syncthreads();
    lds_fft<L, DIR, CB, NT, LP, RP>(buf, tw);
    for (int i = threadIdx.x; i < CB * L; i += NT) {
        const int c = i % CB, e = i / CB;
        const int kx = kx0 + c;
        if (kx < Wh) base[(long long)e * Wp + kx] = buf[c * LP + e];
    }
}

// ---- pass 2: FFT along the last axis, x kernel spectrum, inverse FFT ----------------------
// DIM 2: lines are kx (element stride Wp, L = H).
// DIM 3: lines are ky*Wp + kx (element stride H*Wp, L = D).
// grid (ceil(lines/CB), nwork); work[w] names the source slot and its kernels.
template <int L, int CB, int NT, int DIM, int RP>
__device__ __forceinline__ void spectral_core(float2* __restrict__ buf, float2* __restrict__ tw,
                                              const float2* __restrict__ X, float2* __restrict__ Y,
                                              const float* __restrict__ Kq, const SpecWork* __restrict__ work,
                                              const int* __restrict__ klist, int Wh, int Wp, int H, int Hq,
                                              long long spec, long long kq_stride, const float2* __restrict__ twg,
                                              int blk, int wi) {
    constexpr int LP = L + 1;  // buf holds CB * LP elements
    constexpr int TOT = CB * L;
    constexpr int PER = (TOT + NT - 1) / NT;  // elements per thread; EVEN: no thread runs past TOT
    constexpr bool EVEN = TOT % NT == 0;
    const SpecWork wk = work[wi];
    const long long nlines = DIM == 2 ? (long long)Wh : (long long)H * Wp;
    const long long es = DIM == 2 ? (long long)Wp : (long long)H * Wp;
    const long long line0 = (long long)blk * CB;
    const float2* xs = X + (long long)wk.src_slot * spec;
    stage_twiddles<NT, L>(tw, twg);
    float2 reg[PER];
    int koff[PER];  // offset of this element's folded kernel-spectrum value, or -1 if the line is padding
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        koff[q] = -1;
        if (!EVEN && i >= TOT) continue;
        const int c = i % CB, e = i / CB;
        const long long line = line0 + c;
        int kx, ky;
        if (DIM == 2) {
            kx = (int)line;
            ky = e;
        } else {
            ky = (int)(line / Wp);
            kx = (int)(line - (long long)ky * Wp);
        }
        const bool valid = line < nlines && kx < Wh;
        reg[q] = valid ? xs[(long long)e * es + line] : make_float2(0.f, 0.f);
        if (DIM == 2) {
            const int kyf = ky <= L / 2 ? ky : L - ky;
            koff[q] = valid ? kyf * Wh + kx : -1;
        } else {
            const int kyf = ky <= H / 2 ? ky : H - ky;
            const int kzf = e <= L / 2 ? e : L - e;
            koff[q] = valid ? (kzf * Hq + kyf) * Wh + kx : -1;
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {  // the LDS stores after every load (see rows_fwd_core)
        const int i = threadIdx.x + q * NT;
        if (EVEN || i < TOT) buf[(i % CB) * LP + i / CB] = reg[q];
    }
    __syncthreads();
    lds_fft<L, -1, CB, NT, LP, RP>(buf, tw);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        reg[q] = EVEN || i < TOT ? buf[(i % CB) * LP + i / CB] : make_float2(0.f, 0.f);
    }
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = klist[wk.k_first + j];
        const float* kq = Kq + (long long)k * kq_stride;
        __syncthreads();  // previous kernel's outputs have been read
        float kv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) kv[q] = koff[q] >= 0 ? kq[koff[q]] : 0.0f;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            if (!EVEN && i >= TOT) continue;
            buf[(i % CB) * LP + i / CB] = make_float2(reg[q].x * kv[q], reg[q].y * kv[q]);
        }
        __syncthreads();
        lds_fft<L, +1, CB, NT, LP, RP>(buf, tw);
        float2* ys = Y + (long long)k * spec;
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            const int c = i % CB, e = i / CB;
            if (koff[q] >= 0) ys[(long long)e * es + line0 + c] = buf[c * LP + e];
        }
    }
}

template <int L, int CB, int NT, int DIM, int RP>
__device__ __forceinline__ void spectral_body(const float2* __restrict__ X, float2* __restrict__ Y,
                                              const float* __restrict__ Kq, const SpecWork* __restrict__ work,
                                              const int* __restrict__ klist, int Wh, int Wp, int H, int Hq,
                                              long long spec, long long kq_stride, const float2* __restrict__ twg,
                                              int blk, int wi) {
    __shared__ float2 buf[CB * (L + 1)];
    __shared__ float2 tw[L];
    spectral_core<L, CB, NT, DIM, RP>(buf, tw, X, Y, Kq, work, klist, Wh, Wp, H, Hq, spec, kq_stride, twg, blk, wi);
}

template <int L, int CB, int NT, int DIM, int RP>
__global__ __launch_bounds__(NT) void k_spectral(const float2* __restrict__ X, float2* __restrict__ Y,
                                                 const float* __restrict__ Kq, const SpecWork* __restrict__ work,
                                                 const int* __restrict__ klist, int Wh, int Wp, int H, int Hq,
                                                 long long spec, long long kq_stride,
                                                 const float2* __restrict__ twg) {
    spectral_body<L, CB, NT, DIM, RP>(X, Y, Kq, work, klist, Wh, Wp, H, Hq, spec, kq_stride, twg, blockIdx.x,
                                      blockIdx.y);
}

// batched (2D): item = (member, column block, work item).  Blocks of 16 or more columns
// share few lines with their neighbours: launch order (the XCD-aware order of the adjoint's
// narrower blocks measured slower here at sides 128 and 256).
template <int L, int CB, int NT, int RP>
__global__ __launch_bounds__(NT) void kb_spectral(const MemberDesc* __restrict__ mem, const int4* __restrict__ items) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    spectral_body<L, CB, NT, 2, RP>(m.X, m.Y, m.Kq, m.spec_work, m.klist_spec, m.g.Wh, m.g.Wp, m.g.H, m.g.Hq, m.g.spec,
                                (long long)m.g.Dq * m.g.Hq * m.g.Wh, m.tw_h, it.y, it.z);
}

// ---- pass 3: C2R along x on pairs of rows + growth + weighted accumulate -----------------
// grid (ceil(rows/2), nwork); work[w] is a target channel and its kernels in index order.
// One kernel's growth term w * G(c) (growth_accumulate, growth.rs:51-57: g = 2 exp(-d*d*inv) - 1,
// t += w*g).  A taped step's U is the one this pass wrote, kept in the tape for the adjoint.
__device__ __forceinline__ float growth_term(float c, const GrowthParams& p) {
#pragma clang fp contract(off)
    const float d = c - p.mu;
    const float g = 2.0f * expf(-d * d * p.inv) - 1.0f;
    return p.weight * g;
}

// CONV: also store every kernel's convolution rows (conv + k*N), for the adjoint's recompute.
// GROW false: the convolutions only (no growth, no U), for growth that reads per-cell values.
template <int L, int NT, bool CONV, int RP, bool GROW = true>
__device__ __forceinline__ void rows_inv_core(float2* __restrict__ buf, float2* __restrict__ tw,
                                              const float2* __restrict__ Y, float* __restrict__ U,
                                              const TgtWork* __restrict__ work, const int* __restrict__ klist,
                                              const GrowthParams* __restrict__ gp, int rows, int Wp, long long N,
                                              long long spec, float scale, const float2* __restrict__ twg, int pair,
                                              int wi, float* __restrict__ conv = nullptr) {
    constexpr int PER = (L + NT - 1) / NT;
    const TgtWork wk = work[wi];
    const int y0 = 2 * pair;
    const bool has1 = (y0 + 1) < rows;
    stage_twiddles<NT, L>(tw, twg);
    float acc0[PER], acc1[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) acc0[q] = acc1[q] = 0.0f;
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = klist[wk.k_first + j];
        const float2* r0 = Y + (long long)k * spec + (long long)y0 * Wp;
        const float2* r1 = r0 + Wp;
        // Hermitian completion: X0[n] = conj(X0[L-n]) for n > L/2, likewise X1; the packed
        // line is Z = X0 + i X1, whose inverse is x0 + i x1.  Every load first, then the LDS
        // stores (see rows_fwd_core).
        float2 ra[PER], rb[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int n = threadIdx.x + q * NT;
            if (L % NT == 0 || n < L) {
                const int m = n <= L / 2 ? n : L - n;
                ra[q] = r0[m];
                rb[q] = has1 ? r1[m] : make_float2(0.f, 0.f);
            }
        }
        __syncthreads();  // previous kernel's results have been read
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int n = threadIdx.x + q * NT;
            if (L % NT == 0 || n < L) {
                const float2 a = ra[q], b = rb[q];
                buf[n] = n <= L / 2 ? make_float2(a.x - b.y, a.y + b.x) : make_float2(a.x + b.y, b.x - a.y);
            }
        }
        __syncthreads();
        lds_fft<L, +1, 1, NT, L, RP>(buf, tw);
        GrowthParams p{};
        if (GROW) p = gp[k];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int m = threadIdx.x + q * NT;
            if (L % NT == 0 || m < L) {
                const float2 z = buf[m];
                {
#pragma clang fp contract(off)
                    // growth_accumulate (growth.rs:51-57): g = 2 exp(-d*d*inv) - 1; t += w*g
                    const float c0 = z.x * scale, c1 = z.y * scale;
                    if (CONV) {
                        float* cv = conv + (long long)k * N + (long long)y0 * L + m;
                        cv[0] = c0;
                        if (has1) cv[L] = c1;
                    }
                    if (GROW) {
                        acc0[q] += growth_term(c0, p);
                        acc1[q] += growth_term(c1, p);
                    }
                }
            }
        }
    }
    if (!GROW) return;
    float* u0 = U + (long long)wk.target * N + (long long)y0 * L;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int m = threadIdx.x + q * NT;
        if (L % NT == 0 || m < L) {
            u0[m] = acc0[q];
            if (has1) u0[L + m] = acc1[q];
        }
    }
}

template <int L, int NT, bool CONV, int RP, bool GROW = true>
__device__ __forceinline__ void rows_inv_body(const float2* __restrict__ Y, float* __restrict__ U,
                                              const TgtWork* __restrict__ work, const int* __restrict__ klist,
                                              const GrowthParams* __restrict__ gp, int rows, int Wp, long long N,
                                              long long spec, float scale, const float2* __restrict__ twg, int pair,
                                              int wi, float* __restrict__ conv = nullptr) {
    __shared__ float2 buf[L];
    __shared__ float2 tw[L];
    rows_inv_core<L, NT, CONV, RP, GROW>(buf, tw, Y, U, work, klist, gp, rows, Wp, N, spec, scale, twg, pair, wi,
                                         conv);
}

template <int L, int NT, int RP>
__global__ __launch_bounds__(NT) void k_rows_inv(const float2* __restrict__ Y, float* __restrict__ U,
                                                 const TgtWork* __restrict__ work, const int* __restrict__ klist,
                                                 const GrowthParams* __restrict__ gp, int rows, int Wp,
                                                 long long N, long long spec, float scale,
                                                 const float2* __restrict__ twg) {
    rows_inv_body<L, NT, false, RP>(Y, U, work, klist, gp, rows, Wp, N, spec, scale, twg, blockIdx.x, blockIdx.y);
}

// every kernel's convolution into conv + k*N, no growth (grid and arguments as k_rows_inv)
template <int L, int NT, int RP>
__global__ __launch_bounds__(NT) void k_rows_inv_conv(const float2* __restrict__ Y, float* __restrict__ conv,
                                                      const TgtWork* __restrict__ work, const int* __restrict__ klist,
                                                      int rows, int Wp, long long N, long long spec, float scale,
                                                      const float2* __restrict__ twg) {
    rows_inv_body<L, NT, true, RP, false>(Y, nullptr, work, klist, nullptr, rows, Wp, N, spec, scale, twg, blockIdx.x,
                                          blockIdx.y, conv);
}

// batched: item = (member, row pair, target work item); CONV: convolutions into conv slot cslot
template <int L, int NT, bool CONV, int RP>
__global__ __launch_bounds__(NT) void kb_rows_inv(const MemberDesc* __restrict__ mem, const int4* __restrict__ items,
                                                  int cslot) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    rows_inv_body<L, NT, CONV, RP>(m.Y, m.U, m.tgt_work, m.klist_tgt, m.gp, m.g.H * m.g.D, m.g.Wp, m.g.N, m.g.spec,
                               1.0f / (float)m.g.N, m.tw_w, it.y, it.z,
                               CONV ? m.conv + (long long)cslot * m.K * m.g.N : nullptr);
}

// ---- host-side helpers ------------------------------------------------------------------
// HP_RP / HP_CS bind the run-time variant to the compile-time RP / CS of the call inside.
#define HP_RP(v, ...)                    \
    do {                                 \
        if ((v).rp == 1) {               \
            constexpr int RP = 1;        \
            __VA_ARGS__;                 \
        } else {                         \
            constexpr int RP = 0;        \
            __VA_ARGS__;                 \
        }                                \
    } while (0)
#define HP_CS(v, ...)                    \
    do {                                 \
        if ((v).cs == 2) {               \
            constexpr int CS = 2;        \
            __VA_ARGS__;                 \
        } else {                         \
            constexpr int CS = 1;        \
            __VA_ARGS__;                 \
        }                                \
    } while (0)

namespace detail {
template <int L, int RP>
inline void rows_fwd(hipStream_t st, const float* A, float2* X, const int* src_chan, int nsrc, const Geo& g,
              const float2* tw) {
    const int rows = g.H * g.D;
    dim3 grid((rows + 1) / 2, nsrc);
    hipLaunchKernelGGL((k_rows_fwd<L, Cfg<L>::ROW_NT, RP>), grid, dim3(Cfg<L>::ROW_NT), 0, st, A, X, src_chan, rows,
                       g.Wp, g.N, g.spec, tw);
}

template <int L, int RP, int CS>
inline void cols(hipStream_t st, int dir, float2* X, int nspec, const Geo& g, const float2* tw) {
    constexpr int CB = Cfg<L, CS>::CBC3, NT = Cfg<L, CS>::COLC3_NT;
    dim3 grid((g.Wh + CB - 1) / CB, g.D, nspec);
    if (dir < 0)
        hipLaunchKernelGGL((k_cols<L, CB, NT, -1, RP>), grid, dim3(NT), 0, st, X, g.Wh, g.Wp, g.spec, tw);
    else
        hipLaunchKernelGGL((k_cols<L, CB, NT, 1, RP>), grid, dim3(NT), 0, st, X, g.Wh, g.Wp, g.spec, tw);
}

template <int L, int RP, int CS>
inline void spectral(int dim, hipStream_t st, const float2* X, float2* Y, const float* Kq, const SpecWork* work,
              int nwork, const int* klist, const Geo& g, const float2* tw) {
    const long long kq_stride = (long long)g.Dq * g.Hq * g.Wh;
    if (dim == 2) {
        constexpr int CB = Cfg<L, CS>::CB, NT = Cfg<L, CS>::COL_NT;
        dim3 grid((g.Wh + CB - 1) / CB, nwork);
        hipLaunchKernelGGL((k_spectral<L, CB, NT, 2, RP>), grid, dim3(NT), 0, st, X, Y, Kq, work, klist, g.Wh, g.Wp,
                           g.H, g.Hq, g.spec, kq_stride, tw);
    } else {
        constexpr int CB = Cfg<L, CS>::CBS3, NT = Cfg<L, CS>::SPEC3_NT;
        const long long lines = (long long)g.H * g.Wp;
        dim3 grid((unsigned)((lines + CB - 1) / CB), nwork);
        hipLaunchKernelGGL((k_spectral<L, CB, NT, 3, RP>), grid, dim3(NT), 0, st, X, Y, Kq, work, klist, g.Wh, g.Wp,
                           g.H, g.Hq, g.spec, kq_stride, tw);
    }
}

template <int L, int RP>
inline void rows_inv(hipStream_t st, const float2* Y, float* U, const TgtWork* work, int nwork, const int* klist,
              const GrowthParams* gp, const Geo& g, const float2* tw) {
    const int rows = g.H * g.D;
    dim3 grid((rows + 1) / 2, nwork);
    const float scale = 1.0f / (float)(g.N);  // fft.rs:389, fft3d.rs:449: 1 / (W*H*D) as f32
    hipLaunchKernelGGL((k_rows_inv<L, Cfg<L>::ROW_NT, RP>), grid, dim3(Cfg<L>::ROW_NT), 0, st, Y, U, work, klist, gp,
                       rows, g.Wp, g.N, g.spec, scale, tw);
}

template <int L, int RP>
inline void rows_inv_conv(hipStream_t st, const float2* Y, float* conv, const TgtWork* work, int nwork,
                          const int* klist, const Geo& g, const float2* tw) {
    const int rows = g.H * g.D;
    dim3 grid((rows + 1) / 2, nwork);
    const float scale = 1.0f / (float)(g.N);
    hipLaunchKernelGGL((k_rows_inv_conv<L, Cfg<L>::ROW_NT, RP>), grid, dim3(Cfg<L>::ROW_NT), 0, st, Y, conv, work,
                       klist, rows, g.Wp, g.N, g.spec, scale, tw);
}
}  // namespace detail

// Launchers for the lengths of one group G (SIZES(X) lists them); each returns false for a
// length outside the group.
#define HP_DEFINE_FFT_LAUNCHERS(G, SIZES)                                                                       \
    bool launch_rows_fwd_##G(int L, hipStream_t st, const float* A, float2* X, const int* src_chan, int nsrc,   \
                             const Geo& g, const float2* tw, const FftVariant& v) {                             \
        SIZES(HP_X_ROWS_FWD)                                                                                    \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_cols_##G(int L, hipStream_t st, int dir, float2* X, int nspec, const Geo& g, const float2* tw,  \
                         const FftVariant& v) {                                                                 \
        SIZES(HP_X_COLS)                                                                                        \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_spectral_##G(int L, int dim, hipStream_t st, const float2* X, float2* Y, const float* Kq,       \
                             const SpecWork* work, int nwork, const int* klist, const Geo& g, const float2* tw, \
                             const FftVariant& v) {                                                             \
        SIZES(HP_X_SPECTRAL)                                                                                    \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_rows_inv_##G(int L, hipStream_t st, const float2* Y, float* U, const TgtWork* work, int nwork, \
                             const int* klist, const GrowthParams* gp, const Geo& g, const float2* tw,          \
                             const FftVariant& v) {                                                             \
        SIZES(HP_X_ROWS_INV)                                                                                    \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_rows_inv_conv_##G(int L, hipStream_t st, const float2* Y, float* conv, const TgtWork* work,    \
                                  int nwork, const int* klist, const Geo& g, const float2* tw,                 \
                                  const FftVariant& v) {                                                       \
        SIZES(HP_X_ROWS_INV_CONV)                                                                               \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_batch_rows_fwd_##G(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, \
                                   int cur, const FftVariant& v) {                                              \
        SIZES(HP_X_B_ROWS_FWD)                                                                                  \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_batch_spectral_##G(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, \
                                   const FftVariant& v) {                                                       \
        SIZES(HP_X_B_SPECTRAL)                                                                                  \
        return false;                                                                                           \
    }                                                                                                           \
    bool launch_batch_rows_inv_##G(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, \
                                   bool conv, int cslot, const FftVariant& v) {                                 \
        SIZES(HP_X_B_ROWS_INV)                                                                                  \
        return false;                                                                                           \
    }

#define HP_X_ROWS_FWD(n) if (L == n) { HP_RP(v, detail::rows_fwd<n, RP>(st, A, X, src_chan, nsrc, g, tw)); return true; }
#define HP_X_COLS(n) if (L == n) { HP_RP(v, HP_CS(v, detail::cols<n, RP, CS>(st, dir, X, nspec, g, tw))); return true; }
#define HP_X_SPECTRAL(n) if (L == n) { HP_RP(v, HP_CS(v, detail::spectral<n, RP, CS>(dim, st, X, Y, Kq, work, nwork, klist, g, tw))); return true; }
#define HP_X_ROWS_INV(n) if (L == n) { HP_RP(v, detail::rows_inv<n, RP>(st, Y, U, work, nwork, klist, gp, g, tw)); return true; }
#define HP_X_ROWS_INV_CONV(n) if (L == n) { HP_RP(v, detail::rows_inv_conv<n, RP>(st, Y, conv, work, nwork, klist, g, tw)); return true; }
#define HP_X_B_ROWS_FWD(n) if (L == n) { HP_RP(v, hipLaunchKernelGGL((kb_rows_fwd<n, Cfg<n>::ROW_NT, RP>), dim3(nitems), dim3(Cfg<n>::ROW_NT), 0, st, mem, items, cur)); return true; }
#define HP_X_B_SPECTRAL(n) if (L == n) { HP_RP(v, HP_CS(v, hipLaunchKernelGGL((kb_spectral<n, Cfg<n, CS>::CB, Cfg<n, CS>::COL_NT, RP>), dim3(nitems), dim3(Cfg<n, CS>::COL_NT), 0, st, mem, items))); return true; }
#define HP_X_B_ROWS_INV(n) if (L == n) { HP_RP(v, if (conv) hipLaunchKernelGGL((kb_rows_inv<n, Cfg<n>::ROW_NT, true, RP>), dim3(nitems), dim3(Cfg<n>::ROW_NT), 0, st, mem, items, cslot); \
                                         else hipLaunchKernelGGL((kb_rows_inv<n, Cfg<n>::ROW_NT, false, RP>), dim3(nitems), dim3(Cfg<n>::ROW_NT), 0, st, mem, items, 0)); return true; }

// Length groups (one translation unit each).
#define HP_FFT_G0(X) X(8) X(16) X(32) X(64) X(24) X(48)
#define HP_FFT_G1(X) X(128) X(256) X(96) X(192)
#define HP_FFT_G2(X) X(512) X(1024) X(384) X(768)
#define HP_FFT_G3(X) X(2048) X(1536)
#define HP_FFT_G4(X) X(4096) X(3072)

}  // namespace hp
