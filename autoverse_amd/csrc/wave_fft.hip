// wave_fft.hip -- 2D convolution passes for grids whose sides are 256, 512, 1024 or 2048.
//
// Same mathematics as fft_kernels.hip (R2C rows, spectral product along y, C2R rows with
// the growth epilogue; CachedConvolver::convolve_with_kernel_scratch, fft.rs:288-324,
// and growth_accumulate, growth.rs:44-59), laid out for the GPU differently:
//
//   * spectra are stored blocked column-major, XT[slot][y / 8][kx][y % 8]: the 8-row tile of
//     a row pass is one contiguous run, and a y-pass workgroup (4 adjacent kx columns) reads
//     and writes 256-byte runs per 8 rows;
//   * every line transform is done by ONE wavefront: lane l holds the E = L/64 elements
//     l, l+64, l+128, ... in registers, the radix-16/8/4 Stockham stages run in registers
//     and exchange data through a per-wave LDS line (no workgroup barriers), and the first
//     stage reads / last stage writes that same register layout, so a forward transform's
//     output feeds the kernel-spectrum product and the inverse transform directly;
//   * the row passes work on 8 rows per workgroup (4 packed row pairs, one per wave) and
//     move the transposition through an LDS tile, so their global accesses are 64-byte
//     runs per kx column and whole rows in the spatial domain.
#include "hp_common.hpp"

namespace hp {
namespace wv {

constexpr int WAVE = 64;

__host__ __device__ constexpr int stages(int L) { return L == 256 ? 4 : 3; }
__host__ __device__ constexpr int radix(int L, int s) {
    return L == 256 ? 4 : L == 512 ? 8 : L == 1024 ? (s < 2 ? 16 : 4) : (s < 2 ? 16 : 8);
}
__host__ __device__ constexpr int ns_before(int L, int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= radix(L, i);
    return ns;
}
__host__ __device__ constexpr int padded(int L) { return L + L / 16; }
__device__ __forceinline__ int pad(int i) { return i + (i >> 4); }

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <int DIR>
__device__ __forceinline__ float2 mul_i(float2 a) {  // multiply by DIR*i: forward -i, inverse +i
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2& a, float2& b, float2& c, float2& d) {
    const float2 s0 = cadd(a, c), s1 = csub(a, c);
    const float2 t0 = cadd(b, d), t1 = mul_i<DIR>(csub(b, d));
    a = cadd(s0, t0);
    c = csub(s0, t0);
    b = cadd(s1, t1);
    d = csub(s1, t1);
}

template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 4) {
        dft4<DIR>(v[0], v[1], v[2], v[3]);
    } else if constexpr (R == 8) {
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<DIR>(e0, e1, e2, e3);
        dft4<DIR>(o0, o1, o2, o3);
        const float h = 0.70710678118654752440f;
        o1 = cmul(o1, make_float2(h, DIR * h));
        o2 = mul_i<DIR>(o2);
        o3 = cmul(o3, make_float2(-h, DIR * h));
        v[0] = cadd(e0, o0);
        v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1);
        v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2);
        v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3);
        v[7] = csub(e3, o3);
    } else if constexpr (R == 16) {
        // n = 4 n1 + n2: DFT4 over n1, twiddle W16^(n2 k1), DFT4 over n2; out[k1 + 4 k2]
        float2 a[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            a[n2][0] = v[n2];
            a[n2][1] = v[4 + n2];
            a[n2][2] = v[8 + n2];
            a[n2][3] = v[12 + n2];
            dft4<DIR>(a[n2][0], a[n2][1], a[n2][2], a[n2][3]);
        }
        const float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f, h = 0.70710678118654752440f;
        // W16^m for m = n2*k1 in {1,2,3,4,6,9}
        a[1][1] = cmul(a[1][1], make_float2(c1, DIR * s1));
        a[1][2] = cmul(a[1][2], make_float2(h, DIR * h));
        a[1][3] = cmul(a[1][3], make_float2(s1, DIR * c1));
        a[2][1] = cmul(a[2][1], make_float2(h, DIR * h));
        a[2][2] = mul_i<DIR>(a[2][2]);
        a[2][3] = cmul(a[2][3], make_float2(-h, DIR * h));
        a[3][1] = cmul(a[3][1], make_float2(s1, DIR * c1));
        a[3][2] = cmul(a[3][2], make_float2(-h, DIR * h));
        a[3][3] = cmul(a[3][3], make_float2(-c1, -DIR * s1));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 p = a[0][k1], q = a[1][k1], r = a[2][k1], t = a[3][k1];
            dft4<DIR>(p, q, r, t);
            v[k1] = p;
            v[k1 + 4] = q;
            v[k1 + 8] = r;
            v[k1 + 12] = t;
        }
    }
}

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One wavefront transforms one line of L elements held as e[i] = x[lane + 64 i].
// `line` is this wave's private LDS line (padded(L) float2), `tw` the shared table
// exp(-2 pi i m / L).  Output in the same layout.  Unnormalised; DIR -1 forward.
template <int L, int S, int DIR>
__device__ __forceinline__ void wave_stage(float2 (&e)[L / WAVE], float2* line, const float2* tw) {
    constexpr int E = L / WAVE;
    constexpr int R = radix(L, S);
    constexpr int Q = E / R;
    constexpr int NS = ns_before(L, S);
    constexpr int NB = L / R;
    constexpr int TSTEP = L / (NS * R);
    constexpr bool FIRST = S == 0, LAST = S == stages(L) - 1;
    const int lane = threadIdx.x & (WAVE - 1);
    float2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (FIRST) v[q][r] = e[q + r * Q];
            else v[q][r] = line[pad(lane + WAVE * q + r * NB)];
        }
    if (!FIRST) wave_sync();  // every lane has read its inputs before any output lands
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int j = lane + WAVE * q;
        const int k = j % NS;
        if (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[r * k * TSTEP];
                if (DIR > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
        }
        dft<R, DIR>(v[q]);
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (LAST) e[q + r * Q] = v[q][r];
            else line[pad((j / NS) * NS * R + k + r * NS)] = v[q][r];
        }
    }
    if (!LAST) wave_sync();
}

template <int L, int DIR>
__device__ __forceinline__ void wave_fft(float2 (&e)[L / WAVE], float2* line, const float2* tw) {
    wave_stage<L, 0, DIR>(e, line, tw);
    wave_stage<L, 1, DIR>(e, line, tw);
    wave_stage<L, 2, DIR>(e, line, tw);
    if constexpr (stages(L) > 3) wave_stage<L, 3, DIR>(e, line, tw);
}

// XCD-aware block order for the row passes: workgroup b runs on XCD b % 8, so give every XCD a
// contiguous run of row blocks (neighbouring tiles of the spectrum and of the state meet in the
// same L2).
__device__ __forceinline__ bool xcd_block(int nbx, int total, int& bx, int& by) {
    const int b = blockIdx.x;
    const int per = (total + 7) / 8;
    const int lin = (b % 8) * per + b / 8;
    if (lin >= total) return false;
    bx = lin % nbx;
    by = lin / nbx;
    return true;
}

constexpr int WAVES = 4;  // waves per workgroup
constexpr int NT = WAVES * WAVE;
constexpr int ROWS = 2 * WAVES;  // rows per workgroup in the row passes (one packed pair per wave)

// ---- pass 1: R2C of 8 rows, written as one blocked column-major tile ------------------------
// grid (H / 8, nsrc).  XT[slot][y / 8][kx][y % 8], kx < W/2+1.
template <int L>
__global__ __launch_bounds__(NT) void k_rows_fwd_t(const float* __restrict__ A, float2* __restrict__ XT,
                                                   const int* __restrict__ src_chan, int H, long long N,
                                                   long long spec, int nblocks, const float2* __restrict__ twg) {
    constexpr int E = L / WAVE, LH = L / 2 + 1;
    constexpr int LINE = padded(L);
    __shared__ float2 tw[L];
    __shared__ float2 lines[WAVES * LINE > LH * ROWS ? WAVES * LINE : LH * ROWS];  // FFT lines, then the tile
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int bx, s;
    if (!xcd_block(H / ROWS, nblocks, bx, s)) return;
    const int ybase = bx * ROWS;
    for (int i = threadIdx.x; i < L; i += NT) tw[i] = twg[i];
    const float* a0 = A + (long long)src_chan[s] * N + (long long)(ybase + 2 * wave) * L;
    const float* a1 = a0 + L;
    float2 e[E];
#pragma unroll
    for (int i = 0; i < E; ++i) e[i] = make_float2(a0[lane + WAVE * i], a1[lane + WAVE * i]);
    __syncthreads();  // twiddles staged
    float2* line = lines + wave * LINE;
    wave_fft<L, -1>(e, line, tw);
    // untangle Z = X0 + i X1 (Hermitian halves): needs Z[L - k]
#pragma unroll
    for (int i = 0; i < E; ++i) line[pad(lane + WAVE * i)] = e[i];
    wave_sync();
    float2 x0[E / 2 + 1], x1[E / 2 + 1];
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) {
        const int k = lane + WAVE * i;
        if (k <= L / 2) {
            const float2 zk = line[pad(k)];
            const float2 zm = line[pad(k == 0 ? 0 : L - k)];
            x0[i] = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
            x1[i] = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        }
    }
    __syncthreads();  // all waves are done with their lines before the tile overwrites them
    float2* tile = lines;  // [kx][ROWS]
#pragma unroll
    for (int i = 0; i <= E / 2; ++i) {
        const int k = lane + WAVE * i;
        if (k <= L / 2) {
            tile[k * ROWS + 2 * wave] = x0[i];
            tile[k * ROWS + 2 * wave + 1] = x1[i];
        }
    }
    __syncthreads();
    // blocked column-major: the tile [kx][8 rows] is one contiguous run of the spectrum
    float2* xs = XT + (long long)s * spec + (long long)bx * LH * ROWS;
    for (int t = threadIdx.x; t < LH * ROWS; t += NT) xs[t] = tile[t];
}

// ---- pass 2: y FFT of one kx column per wave, x folded kernel spectrum, inverse y FFT ------
// grid (ceil(Wh / 4), nwork).  KqT[k][kx][kyf], kyf < H/2+1.
template <int L>
__global__ __launch_bounds__(NT) void k_spectral_t(const float2* __restrict__ XT, float2* __restrict__ YT,
                                                   const float* __restrict__ KqT, const SpecWork* __restrict__ work,
                                                   const int* __restrict__ klist, int Wh, long long spec,
                                                   long long kq_stride, const float2* __restrict__ twg) {
    constexpr int E = L / WAVE, LINE = padded(L), HQ = L / 2 + 1;
    __shared__ float2 tw[L];
    __shared__ float2 lines[WAVES * LINE];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const SpecWork wk = work[blockIdx.y];
    const int kx = blockIdx.x * WAVES + wave;
    for (int i = threadIdx.x; i < L; i += NT) tw[i] = twg[i];
    __syncthreads();
    if (kx >= Wh) return;  // after the only block-wide barrier
    float2* line = lines + wave * LINE;
    // blocked column-major element (kx, y): (y / 8) * Wh * 8 + kx * 8 + y % 8; lane + 64 i = y
    const long long col = (long long)(lane >> 3) * Wh * ROWS + (long long)kx * ROWS + (lane & 7);
    const long long ystep = (long long)(WAVE / ROWS) * Wh * ROWS;  // 64 rows further
    const float2* xs = XT + (long long)wk.src_slot * spec + col;
    float2 f[E];
#pragma unroll
    for (int i = 0; i < E; ++i) f[i] = xs[i * ystep];
    wave_fft<L, -1>(f, line, tw);
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = klist[wk.k_first + j];
        const float* kq = KqT + (long long)k * kq_stride + (long long)kx * HQ;
        float2 e[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int ky = lane + WAVE * i;
            const float kv = kq[ky <= L / 2 ? ky : L - ky];
            e[i] = make_float2(f[i].x * kv, f[i].y * kv);
        }
        wave_fft<L, +1>(e, line, tw);
        float2* ys = YT + (long long)k * spec + col;
#pragma unroll
        for (int i = 0; i < E; ++i) ys[i * ystep] = e[i];
    }
}

// ---- pass 3: C2R of 8 rows per kernel + growth + weighted accumulate ------------------------
// grid (H / 8, nwork).
template <int L>
__global__ __launch_bounds__(NT) void k_rows_inv_t(const float2* __restrict__ YT, float* __restrict__ U,
                                                   const TgtWork* __restrict__ work, const int* __restrict__ klist,
                                                   const GrowthParams* __restrict__ gp, int H, long long N,
                                                   long long spec, float scale, int nblocks,
                                                   const float2* __restrict__ twg) {
    constexpr int E = L / WAVE, LINE = padded(L), LH = L / 2 + 1;
    __shared__ float2 tw[L];
    // the staging tile and the per-wave FFT lines share memory: the tile is read completely
    // into registers (and a barrier passed) before any wave starts transforming
    __shared__ float2 buf[WAVES * LINE > LH * ROWS ? WAVES * LINE : LH * ROWS];
    float2* tile = buf;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int bx, wi;
    if (!xcd_block(H / ROWS, nblocks, bx, wi)) return;
    const TgtWork wk = work[wi];
    const int ybase = bx * ROWS;
    for (int i = threadIdx.x; i < L; i += NT) tw[i] = twg[i];
    float acc0[E], acc1[E];
#pragma unroll
    for (int i = 0; i < E; ++i) acc0[i] = acc1[i] = 0.0f;
    float2* line = buf + wave * LINE;
    const int r0 = 2 * wave;
    // the next kernel's tile is loaded into registers while the current one is transformed
    constexpr int TPT = (LH * ROWS + NT - 1) / NT;
    float2 pre[TPT];
    auto fetch = [&](int k) {  // the block's tile is one contiguous run (blocked column-major)
        const float2* ys = YT + (long long)k * spec + (long long)bx * LH * ROWS;
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int t = threadIdx.x + q * NT;
            if (t < LH * ROWS) pre[q] = ys[t];
        }
    };
    if (wk.k_count > 0) fetch(klist[wk.k_first]);
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = klist[wk.k_first + j];
        __syncthreads();  // previous tile fully consumed (and twiddles staged)
#pragma unroll
        for (int q = 0; q < TPT; ++q) {
            const int t = threadIdx.x + q * NT;
            if (t < LH * ROWS) tile[t] = pre[q];
        }
        __syncthreads();
        if (j + 1 < wk.k_count) fetch(klist[wk.k_first + j + 1]);
        float2 e[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            // Hermitian completion of the two half rows; packed line Z = X0 + i X1
            const int n = lane + WAVE * i;
            if (n <= L / 2) {
                const float2 a = tile[n * ROWS + r0], b = tile[n * ROWS + r0 + 1];
                e[i] = make_float2(a.x - b.y, a.y + b.x);
            } else {
                const float2 a = tile[(L - n) * ROWS + r0], b = tile[(L - n) * ROWS + r0 + 1];
                e[i] = make_float2(a.x + b.y, b.x - a.y);
            }
        }
        __syncthreads();  // every wave holds its rows: the tile memory becomes FFT lines
        wave_fft<L, +1>(e, line, tw);
        const GrowthParams p = gp[k];
#pragma unroll
        for (int i = 0; i < E; ++i) {
#pragma clang fp contract(off)
            // growth_accumulate (growth.rs:51-57): g = 2 exp(-d*d*inv) - 1; t += w*g
            const float c0 = e[i].x * scale, c1 = e[i].y * scale;
            const float d0 = c0 - p.mu, d1 = c1 - p.mu;
            const float g0 = 2.0f * expf(-d0 * d0 * p.inv) - 1.0f;
            const float g1 = 2.0f * expf(-d1 * d1 * p.inv) - 1.0f;
            acc0[i] += p.weight * g0;
            acc1[i] += p.weight * g1;
        }
    }
    float* u0 = U + (long long)wk.target * N + (long long)(ybase + r0) * L;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        u0[lane + WAVE * i] = acc0[i];
        u0[L + lane + WAVE * i] = acc1[i];
    }
}

template <int L>
void rows_fwd_t(hipStream_t st, const float* A, float2* XT, const int* src_chan, int nsrc, const Geo& g,
                const float2* tw) {
    const int nblocks = g.H / ROWS * nsrc;
    hipLaunchKernelGGL(k_rows_fwd_t<L>, dim3((nblocks + 7) / 8 * 8), dim3(NT), 0, st, A, XT, src_chan, g.H, g.N, g.spec,
                       nblocks, tw);
}

template <int L>
void spectral_t(hipStream_t st, const float2* XT, float2* YT, const float* KqT, const SpecWork* work, int nwork,
                const int* klist, const Geo& g, const float2* tw) {
    const long long kq_stride = (long long)g.Wh * g.Hq;
    hipLaunchKernelGGL(k_spectral_t<L>, dim3((g.Wh + WAVES - 1) / WAVES, nwork), dim3(NT), 0, st, XT, YT, KqT, work,
                       klist, g.Wh, g.spec, kq_stride, tw);
}

template <int L>
void rows_inv_t(hipStream_t st, const float2* YT, float* U, const TgtWork* work, int nwork, const int* klist,
                const GrowthParams* gp, const Geo& g, const float2* tw) {
    const float scale = 1.0f / (float)(g.N);  // fft.rs:389: 1 / (W*H) as f32
    const int nblocks = g.H / ROWS * nwork;
    hipLaunchKernelGGL(k_rows_inv_t<L>, dim3((nblocks + 7) / 8 * 8), dim3(NT), 0, st, YT, U, work, klist, gp, g.H, g.N,
                       g.spec, scale, nblocks, tw);
}

}  // namespace wv

bool wave_fft_supported(int W, int H) {
    auto ok = [](int n) { return n == 256 || n == 512 || n == 1024 || n == 2048; };
    return ok(W) && ok(H);
}

#define HP_WAVE_SIZES(X) X(256) X(512) X(1024) X(2048)

bool launch_rows_fwd_t(hipStream_t st, const float* A, float2* XT, const int* src_chan, int nsrc, const Geo& g,
                       const float2* tw) {
#define X_(n) if (g.W == n) { wv::rows_fwd_t<n>(st, A, XT, src_chan, nsrc, g, tw); return true; }
    HP_WAVE_SIZES(X_)
#undef X_
    return false;
}

bool launch_spectral_t(hipStream_t st, const float2* XT, float2* YT, const float* KqT, const SpecWork* work,
                       int nwork, const int* klist, const Geo& g, const float2* tw) {
#define X_(n) if (g.H == n) { wv::spectral_t<n>(st, XT, YT, KqT, work, nwork, klist, g, tw); return true; }
    HP_WAVE_SIZES(X_)
#undef X_
    return false;
}

bool launch_rows_inv_t(hipStream_t st, const float2* YT, float* U, const TgtWork* work, int nwork, const int* klist,
                       const GrowthParams* gp, const Geo& g, const float2* tw) {
#define X_(n) if (g.W == n) { wv::rows_inv_t<n>(st, YT, U, work, nwork, klist, gp, g, tw); return true; }
    HP_WAVE_SIZES(X_)
#undef X_
    return false;
}

}  // namespace hp
