// adjoint_fft.hip -- reverse-mode stage 1 (convolution + growth) of the 2D step on gfx950.
//
// Forward (propagator.rs:244-280): conv_k = K_k (*) A_src(k) (circular, via FFT, scaled 1/N),
// U_t = sum_{k -> t} w_k G(conv_k; mu_k, sigma_k) with G = 2 exp(-(u - mu)^2 / (2 sigma^2)) - 1.
// Transpose, given dL/dU (from the divergence pass):
//   g_k = dL/dU_tgt(k) * w_k * G'(conv_k),  G'(u) = -4 (u - mu) inv e,  inv = 1/(2 sigma^2)
//   dL/dw_k  = sum dL/dU * (2e - 1)
//   dL/dmu_k = sum dL/dU * w_k * 4 (u - mu) inv e
//   dL/dsigma_k = sum dL/dU * w_k * 2 e (u - mu)^2 / sigma^3
//   dL/dA_src += K_k^T g_k = K_k (*) g_k (the ring kernel is point-symmetric, so its transpose
//                is itself and its spectrum is real: the same folded spectrum multiplies)
//   dL/dK_k[j] = sum_i g_k[i] A_src[i - j], accumulated over steps in the spectral domain as
//                P_k += FFT(g_k) * conj(FFT(A_src)) and brought back once per backward for
//                the (2R+1)^2 taps only (kb_window_dft).
// Three passes as in the forward: R2C rows of g_k (growth derivative fused in, parameter sums
// reduced per row pair in FP64), a column pass per source that sums K_k * FFT(g_k) over the
// source's kernels, accumulates P_k and inverts, and C2R rows that add into dL/dA.
#include "fft_lds.hpp"

namespace hp {

// item = (member, row pair, kernel).  Reads conv_k and dL/dU; writes FFT_x(g_k) into Y.
template <int L, int NT>
__global__ __launch_bounds__(NT) void kb_adj_rows_fwd(const MemberDesc* __restrict__ mem,
                                                      const int4* __restrict__ items, int cslot) {
    __shared__ float2 buf[L];
    __shared__ float2 tw[L];
    __shared__ double red[3][NT];
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const int pair = it.y, k = it.z;
    const int rows = m.g.H, pairs = (rows + 1) / 2;
    const long long N = m.g.N;
    const int y0 = 2 * pair;
    const bool has1 = (y0 + 1) < rows;
    const GrowthParams p = m.gp[k];
    const int2 km = m.kmap[k];
    const auto cv = gpc(m.conv) + ((long long)cslot * m.K + k) * N + (long long)y0 * L;
    const auto lu = gpc(m.lamU) + (long long)km.y * N + (long long)y0 * L;
    const float inv_s3 = 1.0f / (p.sigma * p.sigma * p.sigma);
    stage_twiddles<NT, L>(tw, gpc(m.tw_w));
    // every load of the row pair first (global loads, issued together), then the arithmetic in
    // the order of the element loop below
    constexpr int PER = (L + NT - 1) / NT;
    float uv[PER][2], lv[PER][2];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        const bool ok = L % NT == 0 || i < L;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const bool in = ok && (r == 0 || has1);
            uv[q][r] = in ? cv[r * L + (in ? i : 0)] : 0.0f;
            lv[q][r] = in ? lu[r * L + (in ? i : 0)] : 0.0f;
        }
    }
    double sw = 0.0, smu = 0.0, ssig = 0.0;
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        if (!(L % NT == 0 || i < L)) break;
        float g[2] = {0.0f, 0.0f};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            if (r == 1 && !has1) break;
            const float u = uv[q][r], l = lv[q][r];
            const float d = u - p.mu;
            const float e = expf(-d * d * p.inv);
            g[r] = l * p.weight * (-4.0f * d * p.inv * e);
            sw += (double)(l * (2.0f * e - 1.0f));
            smu += (double)(l * p.weight * 4.0f * d * p.inv * e);
            ssig += (double)(l * p.weight * 2.0f * e * d * d * inv_s3);
        }
        buf[i] = make_float2(g[0], g[1]);
    }
    red[0][threadIdx.x] = sw;
    red[1][threadIdx.x] = smu;
    red[2][threadIdx.x] = ssig;
    __syncthreads();
    lds_fft<L, -1, 1, NT, L>(buf, tw);
    const auto x0 = gp(m.Y) + (long long)k * m.g.spec + (long long)y0 * m.g.Wp;
    const auto x1 = x0 + m.g.Wp;
    for (int kx = threadIdx.x; kx <= L / 2; kx += NT) {
        const float2 zk = buf[kx];
        const float2 zm = buf[kx == 0 ? 0 : L - kx];
        st2(x0 + kx, make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y)));
        if (has1) st2(x1 + kx, make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x)));
    }
    if (threadIdx.x < 3) {  // fixed-order block sum, then into this row pair's own slot
        double t = 0.0;
        for (int j = 0; j < NT; ++j) t += red[threadIdx.x][j];
        m.pacc[((long long)k * pairs + pair) * 3 + threadIdx.x] += t;
    }
}

// Launch shape of the adjoint column pass: a quarter of the forward's columns per
// workgroup, so that the source column, the accumulator and the offsets (4 values each per
// thread) stay in registers.
template <int L>
struct AdjCfg {
    static constexpr int CB = Cfg<L>::CBP / 4 < 1 ? 1 : Cfg<L>::CBP / 4;
    static constexpr int NT = (CB * L) / 4 < 64 ? 64 : ((CB * L) / 4 > 512 ? 512 : (CB * L) / 4);
    static_assert((CB * L) % NT == 0, "adjoint column tile must split evenly");
};

// item = (member, column block, source work item).  X holds FFT_x(A_src) of the recomputed
// step and is overwritten with FFT_x of sum_k K_k (*) g_k; P_k accumulates.
//
// A block of CB < 16 columns covers part of each 128-byte line of the spectrum rows; the
// neighbouring blocks (the next items) cover the rest.  Items run in the XCD-aware order of
// xcd_chunk_item, so each line is fetched into one L2 once, not once per block (the pass read
// 4.3 times its bytes from memory with one item per workgroup in launch order).
template <int L, int CB, int NT>
__global__ __launch_bounds__(NT, 4) void kb_adj_spectral(const MemberDesc* __restrict__ mem,
                                                      const int4* __restrict__ items, int nitems) {
    constexpr int LP = L + 1;
    constexpr int PER = CB * L / NT;
    static_assert((CB * L) % NT == 0, "tile must split evenly over the threads");
    __shared__ float2 buf[CB * LP];
    __shared__ float2 tw[L];
    const int lin = xcd_chunk_item<16>(nitems);
    if (lin < 0) return;  // before any barrier: the whole workgroup leaves
    const int4 it = items[lin];
    const MemberDesc& m = mem[it.x];
    const SpecWork wk = m.spec_work[it.z];
    const int Wh = m.g.Wh, Wp = m.g.Wp;
    const long long spec = m.g.spec, kq_stride = (long long)m.g.Hq * Wh;
    const long long line0 = (long long)it.y * CB;
    const auto xs = gp(m.X) + (long long)wk.src_slot * spec;
    stage_twiddles<NT, L>(tw, gpc(m.tw_h));
    float2 xr[PER], acc[PER];
    int koff[PER];
    // every memory read is issued before the LDS stores and transforms that follow it
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        const int c = i % CB, e = i / CB;
        const long long kx = line0 + c;
        const bool valid = kx < Wh;
        xr[q] = valid ? ld2(xs + (long long)e * Wp + kx) : make_float2(0.f, 0.f);
        const int kyf = e <= L / 2 ? e : L - e;
        koff[q] = valid ? kyf * Wh + (int)kx : -1;
        acc[q] = make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        buf[(i % CB) * LP + i / CB] = xr[q];
    }
    __syncthreads();
    lds_fft<L, -1, CB, NT, LP>(buf, tw);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        xr[q] = buf[(i % CB) * LP + i / CB];
    }
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = m.klist_spec[wk.k_first + j];
        const auto kq = gpc(m.Kq) + (long long)k * kq_stride;
        const auto gs = gpc(m.Y) + (long long)k * spec;
        const auto pk = gp(m.P) + (long long)k * spec;
        float2 gv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            const long long o = (long long)(i / CB) * Wp + line0 + i % CB;
            gv[q] = koff[q] >= 0 ? ld2(gs + o) : make_float2(0.f, 0.f);
        }
        __syncthreads();  // previous reads of buf are done
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            buf[(i % CB) * LP + i / CB] = gv[q];
        }
        __syncthreads();
        lds_fft<L, -1, CB, NT, LP>(buf, tw);
        // P and the kernel spectrum after the transform (issued while it was still in flight
        // they held the registers of a fourth wave per SIMD), all before the arithmetic
        float2 pv[PER];
        float kv[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            const long long o = (long long)(i / CB) * Wp + line0 + i % CB;
            const bool ok = koff[q] >= 0;
            pv[q] = ok ? ld2(pk + o) : make_float2(0.f, 0.f);
            kv[q] = ok ? kq[koff[q]] : 0.0f;
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int i = threadIdx.x + q * NT;
            const int c = i % CB, e = i / CB;
            if (koff[q] < 0) continue;
            const float2 gq = buf[c * LP + e];
            // g * conj(x)
            st2(pk + (long long)e * Wp + line0 + c,
                make_float2(pv[q].x + (gq.x * xr[q].x + gq.y * xr[q].y), pv[q].y + (gq.y * xr[q].x - gq.x * xr[q].y)));
            acc[q].x += kv[q] * gq.x;
            acc[q].y += kv[q] * gq.y;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        buf[(i % CB) * LP + i / CB] = acc[q];
    }
    __syncthreads();
    lds_fft<L, +1, CB, NT, LP>(buf, tw);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        const int c = i % CB, e = i / CB;
        if (koff[q] >= 0) st2(xs + (long long)e * Wp + line0 + c, buf[c * LP + e]);
    }
}

// item = (member, row pair, source slot).  C2R of the slot's summed spectrum, added into
// lam[1 - li] of the slot's channel with the forward's 1/N.
template <int L, int NT>
__global__ __launch_bounds__(NT) void kb_adj_rows_inv(const MemberDesc* __restrict__ mem,
                                                      const int4* __restrict__ items, int li) {
    constexpr int PER = (L + NT - 1) / NT;
    __shared__ float2 buf[L];
    __shared__ float2 tw[L];
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const int rows = m.g.H;
    const int y0 = 2 * it.y;
    const bool has1 = (y0 + 1) < rows;
    const long long N = m.g.N;
    const float scale = 1.0f / (float)N;
    stage_twiddles<NT, L>(tw, m.tw_w);
    const auto r0 = gpc(m.X) + (long long)it.z * m.g.spec + (long long)y0 * m.g.Wp;
    const auto r1 = r0 + m.g.Wp;
    const auto out = gp(m.lam[1 - li]) + (long long)m.src_chan[it.z] * N + (long long)y0 * L;
    // every load first (the two spectrum rows and the accumulated rows), then the LDS stores
    float2 ra[PER], rb[PER];
    float o0[PER], o1[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int n = threadIdx.x + q * NT;
        if (L % NT == 0 || n < L) {
            const int mm = n <= L / 2 ? n : L - n;
            ra[q] = ld2(r0 + mm);
            rb[q] = has1 ? ld2(r1 + mm) : make_float2(0.f, 0.f);
            o0[q] = out[n];
            o1[q] = has1 ? out[L + n] : 0.0f;
        }
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int n = threadIdx.x + q * NT;
        if (L % NT == 0 || n < L) {
            const float2 a = ra[q], b = rb[q];
            buf[n] = n <= L / 2 ? make_float2(a.x - b.y, a.y + b.x) : make_float2(a.x + b.y, b.x - a.y);
        }
    }
    __syncthreads();
    lds_fft<L, +1, 1, NT, L>(buf, tw);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int n = threadIdx.x + q * NT;
        if (L % NT == 0 || n < L) {
            const float2 z = buf[n];
            out[n] = o0[q] + z.x * scale;
            if (has1) out[L + n] = o1[q] + z.y * scale;
        }
    }
}

// dkwin[off + (jy+R)(2R+1) + (jx+R)] = sum over the full spectrum of P_k e^{+2 pi i (kx jx / W +
// ky jy / H)} (unscaled; times 1/N on the host), evaluated from the half spectrum as sum_ky
// sum_{kx <= W/2} m_kx Re(...), m_kx = 1 at kx = 0, W/2 else 2; separable: each block of 8
// spectrum rows is reduced over kx into LDS first, then over its rows into the window.  FP64
// throughout.  Two kernels: kb_window_dft_rows, one workgroup per (member, kernel, row block),
// writes each block's window to part[poff + rb (2R+1)^2 + o]; kb_window_dft_sum adds the blocks
// into dkwin in ascending block order (the order a single pass over the blocks would use), so
// the result does not depend on how the blocks were spread over workgroups.
constexpr int WDFT_NT = 256, WDFT_ROWS = 8, WDFT_MAXS = 129, WDFT_MAXL = 1024;

// item = (member, kernel, row block rb, poff)
__global__ __launch_bounds__(WDFT_NT) void kb_window_dft_rows(const MemberDesc* __restrict__ mem,
                                                              const int4* __restrict__ items,
                                                              double* __restrict__ part) {
    __shared__ double2 twW[WDFT_MAXL];
    __shared__ double2 twH[WDFT_MAXL];
    __shared__ double2 Q[WDFT_ROWS][WDFT_MAXS];
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const int k = it.y, ky0 = it.z * WDFT_ROWS;
    const int W = m.g.W, H = m.g.H, Wh = m.g.Wh, Wp = m.g.Wp;
    const int R = m.dkw[k].y, S = 2 * R + 1;
    double* out = part + it.w + (long long)it.z * S * S;
    const float2* P = m.P + (long long)k * m.g.spec;
    for (int i = threadIdx.x; i < W; i += WDFT_NT) {
        double sn, cs;
        sincospi(2.0 * (double)i / (double)W, &sn, &cs);
        twW[i] = make_double2(cs, sn);
    }
    for (int i = threadIdx.x; i < H; i += WDFT_NT) {
        double sn, cs;
        sincospi(2.0 * (double)i / (double)H, &sn, &cs);
        twH[i] = make_double2(cs, sn);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < WDFT_ROWS * S; t += WDFT_NT) {
        const int r = t / S, jxi = t - r * S;
        const int ky = ky0 + r;
        double re = 0.0, im = 0.0;
        if (ky < H) {
            const int jx = jxi - R;
            const int step = ((jx % W) + W) % W;
            int ph = 0;
            const float2* row = P + (long long)ky * Wp;
            for (int kx = 0; kx < Wh; ++kx) {
                const double2 w = twW[ph];
                const float2 p = row[kx];
                const double mk = (kx == 0 || 2 * kx == W) ? 1.0 : 2.0;
                re += mk * ((double)p.x * w.x - (double)p.y * w.y);
                im += mk * ((double)p.x * w.y + (double)p.y * w.x);
                ph += step;
                if (ph >= W) ph -= W;
            }
        }
        Q[r][jxi] = make_double2(re, im);
    }
    __syncthreads();
    for (int o = threadIdx.x; o < S * S; o += WDFT_NT) {
        const int jyi = o / S, jxi = o - jyi * S;
        const int jy = jyi - R;
        const int step = ((jy % H) + H) % H;
        int ph = (int)(((long long)ky0 * step) % H);
        double a = 0.0;
        for (int r = 0; r < WDFT_ROWS && ky0 + r < H; ++r) {
            const double2 w = twH[ph];
            a += Q[r][jxi].x * w.x - Q[r][jxi].y * w.y;
            ph += step;
            if (ph >= H) ph -= H;
        }
        out[o] = a;
    }
}

// item = (member, kernel, row blocks, poff), blockIdx.y = a chunk of WDFT_NT window entries:
// dkwin += the blocks' windows, block by block (the loads of eight blocks in flight at once)
__global__ __launch_bounds__(WDFT_NT) void kb_window_dft_sum(const MemberDesc* __restrict__ mem,
                                                             const int4* __restrict__ items,
                                                             const double* __restrict__ part) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const int k = it.y, nrb = it.z;
    const int S = 2 * m.dkw[k].y + 1, SS = S * S;
    const int o = (int)blockIdx.y * WDFT_NT + (int)threadIdx.x;
    if (o >= SS) return;
    double* out = m.dkwin + m.dkw[k].x;
    const double* pk = part + it.w + o;
    double acc = out[o];
    int rb = 0;
    for (; rb + 8 <= nrb; rb += 8) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = pk[(long long)(rb + j) * SS];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += v[j];
    }
    for (; rb < nrb; ++rb) acc += pk[(long long)rb * SS];
    out[o] = acc;
}

// ---- launchers -------------------------------------------------------------------------------
#define HP_ADJ_SIZES(X_) X_(8) X_(16) X_(32) X_(64) X_(128) X_(256) X_(512) X_(1024) \
    X_(24) X_(48) X_(96) X_(192) X_(384) X_(768)

bool launch_adj_rows_fwd(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cslot) {
#define X_(n) if (L == n) { hipLaunchKernelGGL((kb_adj_rows_fwd<n, Cfg<n>::ROW_NT>), dim3(nitems), dim3(Cfg<n>::ROW_NT), 0, st, mem, items, cslot); return true; }
    HP_ADJ_SIZES(X_)
#undef X_
    return false;
}

bool launch_adj_spectral(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems) {
#define X_(n) if (L == n) { hipLaunchKernelGGL((kb_adj_spectral<n, AdjCfg<n>::CB, AdjCfg<n>::NT>), dim3(xcd_chunk_grid<16>(nitems)), dim3(AdjCfg<n>::NT), 0, st, mem, items, nitems); return true; }
    HP_ADJ_SIZES(X_)
#undef X_
    return false;
}

bool launch_adj_rows_inv(int L, hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int li) {
#define X_(n) if (L == n) { hipLaunchKernelGGL((kb_adj_rows_inv<n, Cfg<n>::ROW_NT>), dim3(nitems), dim3(Cfg<n>::ROW_NT), 0, st, mem, items, li); return true; }
    HP_ADJ_SIZES(X_)
#undef X_
    return false;
}

int adj_spectral_block_lines(int L) {
#define X_(n) if (L == n) return AdjCfg<n>::CB;
    HP_ADJ_SIZES(X_)
#undef X_
    return 0;
}

bool adj_fft_supported(int L) {
#define X_(n) if (L == n) return true;
    HP_ADJ_SIZES(X_)
#undef X_
    return false;
}

void launch_window_dft(hipStream_t st, const MemberDesc* mem, const int4* row_items, int nrow_items,
                       const int4* sum_items, int nsum_items, double* part) {
    if (nrow_items > 0)
        hipLaunchKernelGGL(kb_window_dft_rows, dim3(nrow_items), dim3(WDFT_NT), 0, st, mem, row_items, part);
    if (nsum_items > 0)
        hipLaunchKernelGGL(kb_window_dft_sum, dim3(nsum_items, (WDFT_MAXS * WDFT_MAXS + WDFT_NT - 1) / WDFT_NT),
                           dim3(WDFT_NT), 0, st, mem, sum_items, (const double*)part);
}

int window_dft_rows() { return WDFT_ROWS; }
int window_dft_max_side() { return WDFT_MAXL; }
int window_dft_max_taps() { return WDFT_MAXS; }

}  // namespace hp
