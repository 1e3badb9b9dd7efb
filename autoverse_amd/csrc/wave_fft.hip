// wave_fft.hip -- 2D convolution passes for grids whose sides are 256, 512, 1024 or 2048.
//
// Same mathematics as fft_kernels.hip (R2C rows, spectral product along y, C2R rows with
// the growth epilogue; CachedConvolver::convolve_with_kernel_scratch, fft.rs:288-324,
// and growth_accumulate, growth.rs:44-59), laid out for the GPU differently:
//
//   * the transforms are wavefront FFTs: lane l holds the E = n/64 elements l, l+64, l+128,
//     ... in registers, the radix-16/8/4/2 Stockham stages run in registers and exchange data
//     through a per-wave LDS line (no workgroup barriers), and the first stage reads / last
//     stage writes that same register layout, so a forward transform's output feeds the
//     kernel-spectrum product and the inverse transform directly;
//   * a row of L reals is transformed by one wavefront as an L/2-point complex FFT of its
//     (even, odd) pairs, followed by the split into the half spectrum (and the reverse for
//     the inverse): one wavefront per row, each with half the work and half the LDS line of
//     a two-rows-per-FFT packing;
//   * a kx column of H is transformed by two wavefronts, one H/2-point FFT each, joined by a
//     radix-2 butterfly exchanged through LDS;
//   * spectra are stored blocked column-major, XT[slot][y / 4][kx][y % 4]: the 4-row tile of
//     a row-pass workgroup is one contiguous run, and a y-pass workgroup (2 adjacent kx
//     columns, 2 wavefronts each) reads and writes 64-byte runs per 4 rows, its XCD
//     neighbour the other half of each 128-byte line.
#include <vector>

#include "hp_common.hpp"

namespace hp {
namespace wv {

constexpr int WAVE = 64;

// Radix plans of the wave FFT, by length: every stage's radix divides the 64-lane element
// count E = n / 64.
__host__ __device__ constexpr int stages(int L) { return L == 128 ? 7 : L == 256 ? 4 : 3; }
__host__ __device__ constexpr int radix(int L, int s) {
    return L == 128 ? 2 : L == 256 ? 4 : L == 512 ? 8 : L == 1024 ? (s < 2 ? 16 : 4) : (s < 2 ? 16 : 8);
}
__host__ __device__ constexpr int ns_before(int L, int s) {
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= radix(L, i);
    return ns;
}
__host__ __device__ constexpr int padded(int L) { return L + L / 8; }
__device__ __forceinline__ int pad(int i) { return i + (i >> 3); }

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
    if constexpr (R == 2) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
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

// Per-stage twiddle tables.  Stage S >= 1 multiplies input r of butterfly k by
// W_L^(r k TSTEP); its table holds those values at [(r - 1) NS + k] (copies of entries of the
// full table exp(-2 pi i m / L), so the arithmetic is unchanged), which a wave reads at
// consecutive addresses instead of strided ones.
__host__ __device__ constexpr int tw_stage_size(int L, int s) { return s == 0 ? 0 : (radix(L, s) - 1) * ns_before(L, s); }
__host__ __device__ constexpr int tw_stage_off(int L, int s) {
    int o = 0;
    for (int i = 0; i < s; ++i) o += tw_stage_size(L, i);
    return o;
}
__host__ __device__ constexpr int tw_total(int L) { return tw_stage_off(L, stages(L)); }

// Entry idx of the per-stage tables of length L holds exp(-2 pi i m / L) for the m returned here.
__host__ __device__ constexpr int stage_twiddle_src(int L, int idx) {
    int s = 1;
    while (s + 1 < stages(L) && idx >= tw_stage_off(L, s + 1)) ++s;
    const int ns = ns_before(L, s), tstep = L / (ns * radix(L, s));
    const int rel = idx - tw_stage_off(L, s);
    const int r = rel / ns + 1, k = rel - (r - 1) * ns;
    return r * k * tstep;
}

// LDS entries of the per-stage tables of length L staged by NTH threads: tw_total rounded up to
// a multiple of NTH, so every thread writes every trip (the entries past tw_total hold copies
// of the last one and are never read).
__host__ __device__ constexpr int tw_alloc(int L, int NTH) { return (tw_total(L) + NTH - 1) / NTH * NTH; }

// Copy the per-stage tables of length L (built once on the host) into LDS, workgroup-cooperative,
// in two halves: load() issues the reads into registers, store() writes them to LDS.  A kernel
// calls load() first, its own reads next and store() after them, so no read waits for the
// twiddles' LDS writes and the twiddles' reads share one memory latency with the kernel's.
template <int L, int NTH>
struct TwiddleStage {
    static constexpr int T = tw_total(L), IT = (T + NTH - 1) / NTH;
    float2 v[IT];
    __device__ __forceinline__ void load(const float2* __restrict__ tws) {
#pragma unroll
        for (int i = 0; i < IT; ++i) v[i] = tws[min((int)threadIdx.x + i * NTH, T - 1)];
    }
    __device__ __forceinline__ void store(float2* tw) const {  // tw holds tw_alloc(L, NTH) entries
#pragma unroll
        for (int i = 0; i < IT; ++i) tw[threadIdx.x + i * NTH] = v[i];
    }
};

// One wavefront transforms one line of L elements held as e[i] = x[lane + 64 i].
// `line` is this wave's private LDS line (padded(L) float2), `tw` the per-stage twiddle
// tables.  Output in the same layout.  Unnormalised; DIR -1 forward.
template <int L, int S, int DIR>
__device__ __forceinline__ void wave_stage(float2 (&e)[L / WAVE], float2* line, const float2* tw) {
    constexpr int E = L / WAVE;
    constexpr int R = radix(L, S);
    constexpr int Q = E / R;
    constexpr int NS = ns_before(L, S);
    constexpr int NB = L / R;
    constexpr int TOFF = tw_stage_off(L, S);
    constexpr bool FIRST = S == 0, LAST = S == stages(L) - 1;
    static_assert(Q * R == E, "radix must divide the per-lane element count");
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
                float2 w = tw[TOFF + (r - 1) * NS + k];
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

template <int L, int S, int DIR>
__device__ __forceinline__ void wave_stages(float2 (&e)[L / WAVE], float2* line, const float2* tw) {
    if constexpr (S < stages(L)) {
        wave_stage<L, S, DIR>(e, line, tw);
        wave_stages<L, S + 1, DIR>(e, line, tw);
    }
}

template <int L, int DIR>
__device__ __forceinline__ void wave_fft(float2 (&e)[L / WAVE], float2* line, const float2* tw) {
    wave_stages<L, 0, DIR>(e, line, tw);
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

constexpr int WAVES = 4;      // waves per workgroup
constexpr int NT = WAVES * WAVE;
constexpr int RB = WAVES;     // rows per spectrum block = rows per row-pass workgroup (one per wave)
constexpr int TP = RB + 1;    // LDS pitch of a row tile's kx entries: lanes at consecutive kx hit distinct banks

// ---- pass 1: R2C of 4 rows (one per wave), written as one blocked column-major tile --------
// grid (H / 4 * nsrc).  XT[slot][y / 4][kx][y % 4], kx < W/2+1.  twg: the length-L table.
//
// Row x (L reals) as z[n] = x[2n] + i x[2n+1], n < M = L/2; Z = FFT_M(z).  With
// E[k] = (Z[k] + conj Z[M-k]) / 2 and O[k] = -i (Z[k] - conj Z[M-k]) / 2 (the transforms of the
// even and odd samples), X[k] = E[k] + W_L^k O[k] for k < M and X[M] = Re Z[0] - Im Z[0].
template <int L>
__global__ __launch_bounds__(NT) void k_rows_fwd_t(const float* __restrict__ A, float2* __restrict__ XT,
                                                   const int* __restrict__ src_chan, int H, long long N,
                                                   long long spec, int nblocks, const float2* __restrict__ twg,
                                                   const float2* __restrict__ tws) {
    constexpr int M = L / 2, E = M / WAVE, LH = L / 2 + 1;
    constexpr int LINE = padded(M);
    __shared__ float2 tw[tw_alloc(M, NT)];
    __shared__ float2 lines[WAVES * LINE > LH * TP ? WAVES * LINE : LH * TP];  // FFT lines, then the tile
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int bx, s;
    if (!xcd_block(H / RB, nblocks, bx, s)) return;
    TwiddleStage<M, NT> stage;
    stage.load(tws);
    const float2* row =
        reinterpret_cast<const float2*>(A + (long long)src_chan[s] * N + (long long)(bx * RB + wave) * L);
    float2 e[E], w[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        e[i] = row[lane + WAVE * i];
        w[i] = twg[lane + WAVE * i];  // W_L^k, k = lane + 64 i
    }
    stage.store(tw);
    __syncthreads();  // twiddles staged
    float2* line = lines + wave * LINE;
    wave_fft<M, -1>(e, line, tw);
#pragma unroll
    for (int i = 0; i < E; ++i) line[pad(lane + WAVE * i)] = e[i];
    wave_sync();
    float2 xk[E];
#pragma unroll
    for (int i = 0; i < E; ++i) {
        const int k = lane + WAVE * i;
        const float2 zk = e[i], zm = line[pad(k == 0 ? 0 : M - k)];
        const float2 ev = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
        const float2 od = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
        xk[i] = cadd(ev, cmul(w[i], od));
    }
    __syncthreads();  // all waves are done with their lines before the tile overwrites them
    float2* tile = lines;  // [kx][TP], rows 0..RB-1 used
#pragma unroll
    for (int i = 0; i < E; ++i) tile[(lane + WAVE * i) * TP + wave] = xk[i];
    if (lane == 0) tile[M * TP + wave] = make_float2(e[0].x - e[0].y, 0.0f);
    __syncthreads();
    // blocked column-major: the tile [kx][4 rows] is one contiguous run of the spectrum
    float2* xs = XT + (long long)s * spec + (long long)bx * LH * RB;
    for (int t = threadIdx.x; t < LH * RB; t += NT) xs[t] = tile[(t / RB) * TP + t % RB];
}

// ---- pass 2: y FFT of one kx column per wave pair, x folded kernel spectrum, inverse y FFT ----
// grid ceil(Wh / 2) * nwork (rounded up to a multiple of 8); a workgroup is two columns x two halves.  KqT[k][kx][kyf],
// kyf < H/2+1.
//
// A column of L = 2M is split over two wavefronts.  Forward (radix-2 decimation in frequency):
// a[n] = x[n] + x[n+M], b[n] = (x[n] - x[n+M]) W_L^n, and FFT_M(a)[k] = X[2k],
// FFT_M(b)[k] = X[2k+1].  Inverse (decimation in time) of the product P: with
// e = IFFT_M(P[2k]) and o = IFFT_M(P[2k+1]) W_L^-n (unnormalised), y[n] = e[n] + o[n] and
// y[n+M] = e[n] - o[n].  Each butterfly exchanges one half through LDS.
template <int L>
__global__ __launch_bounds__(NT) void k_spectral_t(const float2* __restrict__ XT, float2* __restrict__ YT,
                                                   const float* __restrict__ KqT, const SpecWork* __restrict__ work,
                                                   const int* __restrict__ klist, int Wh, long long spec,
                                                   long long kq_stride, int ncp, int nblocks,
                                                   const float2* __restrict__ twg, const float2* __restrict__ tws) {
    constexpr int M = L / 2, E = M / WAVE, LINE = padded(M), HQ = L / 2 + 1;
    __shared__ float2 tw[tw_alloc(M, NT)];
    __shared__ float2 lines[WAVES * LINE];
    // the wave index is uniform over the wavefront: held in a scalar register, the half test
    // below is a scalar branch, not a per-lane mask around every LDS read
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / WAVE), lane = threadIdx.x % WAVE;
    const int h = wave & 1;                    // half: 0 -> rows/outputs n, even k; 1 -> n + M, odd k
    // XCD-aware order: the workgroups that share the 128-byte lines of a row block (adjacent
    // column pairs) run on the same XCD, so the lines are completed in one L2
    int cp, wi;
    if (!xcd_block(ncp, nblocks, cp, wi)) return;  // before any barrier: the whole workgroup leaves
    const SpecWork wk = work[wi];
    const int kx = cp * (WAVES / 2) + (wave >> 1);
    TwiddleStage<M, NT> stage;
    stage.load(tws);
    float2* line = lines + wave * LINE;
    const float2* other = lines + (wave ^ 1) * LINE;
    // blocked column-major element (kx, y): (y / 4) * Wh * 4 + kx * 4 + y % 4; y = h M + lane + 64 i
    // W/2 + 1 columns is odd, so the last pair's second column (kx = Wh) does not exist: its two
    // waves transform column Wh - 1 again and store the same values to the same addresses as
    // that column's own waves.  Every wave then stores on every trip, which lets the compiler
    // count the stores still in flight when it waits for the next kernel's spectrum values.
    const int kxc = min(kx, Wh - 1);
    const long long col = (long long)((h * M) / RB + lane / RB) * Wh * RB + (long long)kxc * RB + (lane % RB);
    const long long ystep = (long long)(WAVE / RB) * Wh * RB;  // 64 rows further
    float2 f[E], w[E];
    const float2* xs = XT + (long long)wk.src_slot * spec + col;
#pragma unroll
    for (int i = 0; i < E; ++i) {
        f[i] = xs[i * ystep];
        w[i] = twg[lane + WAVE * i];  // W_L^n, n = lane + 64 i
    }
    // the next kernel's spectrum values (ky = 2k + h) are loaded while the current one works
    float kvn[E];
    auto fetch = [&](int k) {
        const float* kq = KqT + (long long)k * kq_stride + (long long)kxc * HQ;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int ky = 2 * (lane + WAVE * i) + h;
            kvn[i] = kq[ky <= L / 2 ? ky : L - ky];
        }
    };
    if (wk.k_count > 0) fetch(klist[wk.k_first]);
    else {
#pragma unroll
        for (int i = 0; i < E; ++i) kvn[i] = 0.0f;
    }
    stage.store(tw);
    // forward butterfly across the halves
#pragma unroll
    for (int i = 0; i < E; ++i) line[lane + WAVE * i] = f[i];
    __syncthreads();  // both halves (and the twiddle tables) are in LDS
    {
        float2 o[E];
#pragma unroll
        for (int i = 0; i < E; ++i) o[i] = other[lane + WAVE * i];
        if (h == 0) {
#pragma unroll
            for (int i = 0; i < E; ++i) f[i] = cadd(f[i], o[i]);
        } else {
#pragma unroll
            for (int i = 0; i < E; ++i) f[i] = cmul(csub(o[i], f[i]), w[i]);
        }
    }
    __syncthreads();  // the partner has read this line before the FFT reuses it
    wave_fft<M, -1>(f, line, tw);
    // The first kernel's values arrived during the forward transform; marking them consumed
    // here (an empty asm that takes and returns them) leaves no load of them pending on the
    // loop's entry path.  The loop head then waits only as the back edge needs: for the
    // values fetched in the trip before, with that trip's stores still in flight.
#pragma unroll
    for (int i = 0; i < E; ++i) asm volatile("" : "+v"(kvn[i]));
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = klist[wk.k_first + j];
        float2 e[E];
#pragma unroll
        for (int i = 0; i < E; ++i) e[i] = make_float2(f[i].x * kvn[i], f[i].y * kvn[i]);
        // unconditional (the last trip refetches the current kernel): a conditional fetch
        // leaves a copy into the loop's registers at the back edge, which waits for this
        // trip's stores
        fetch(klist[wk.k_first + min(j + 1, wk.k_count - 1)]);
        wave_fft<M, +1>(e, line, tw);  // own line only: the partner finished reading it
        if (h == 1) {
#pragma unroll
            for (int i = 0; i < E; ++i) e[i] = cmul(e[i], make_float2(w[i].x, -w[i].y));
        }
#pragma unroll
        for (int i = 0; i < E; ++i) line[lane + WAVE * i] = e[i];
        __syncthreads();  // both halves written
        float2* ys = YT + (long long)k * spec + col;
        float2 o[E];
#pragma unroll
        for (int i = 0; i < E; ++i) o[i] = other[lane + WAVE * i];
        // y = e + o (half 0) or o - e (half 1), as o + sg e with sg = +-1: the product is exact
        // and the sum rounds once, fused or not, so both halves store from one straight block
        const float sg = h == 0 ? 1.0f : -1.0f;
#pragma unroll
        for (int i = 0; i < E; ++i) ys[i * ystep] = make_float2(o[i].x + sg * e[i].x, o[i].y + sg * e[i].y);
        __syncthreads();  // the partner has read this line before the next FFT reuses it
    }
}

// ---- pass 3: C2R of 4 rows (one per wave) per kernel + growth + weighted accumulate ---------
// grid (H / 4 * nwork).  Inverse of pass 1's split: with X[k], k <= M, of a real row,
// Z[k] = (X[k] + conj X[M-k]) + i (X[k] - conj X[M-k]) W_L^-k, and the unnormalised
// IFFT_M(Z)[n] = L (x[2n] + i x[2n+1]) -- the unnormalised length-L inverse, as the LDS path
// produces it.  Each wave reads its own row's X[k] and X[M-k] (k = lane + 64 i) straight from
// the blocked spectrum -- the four waves of a workgroup together cover whole lines -- and holds
// the next kernel's values in registers while the current one is transformed; the only
// workgroup barrier is the twiddle staging.
template <int L>
__global__ __launch_bounds__(NT) void k_rows_inv_t(const float2* __restrict__ YT, float* __restrict__ U,
                                                   const TgtWork* __restrict__ work, const int* __restrict__ klist,
                                                   const GrowthParams* __restrict__ gp, int H, long long N,
                                                   long long spec, float scale, int nblocks,
                                                   const float2* __restrict__ twg, const float2* __restrict__ tws) {
    constexpr int M = L / 2, E = M / WAVE, LINE = padded(M), LH = L / 2 + 1;
    __shared__ float2 tw[tw_alloc(M, NT)];
    __shared__ float2 lines[WAVES * LINE];
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    int bx, wi;
    if (!xcd_block(H / RB, nblocks, bx, wi)) return;
    const TgtWork wk = work[wi];
    {  // staged before the spectrum reads are issued: measured faster here than after them
        TwiddleStage<M, NT> stage;
        stage.load(tws);
        stage.store(tw);
    }
    float2 w[E];
#pragma unroll
    for (int i = 0; i < E; ++i) w[i] = twg[lane + WAVE * i];  // W_L^k, k = lane + 64 i
    float acc0[E], acc1[E];
#pragma unroll
    for (int i = 0; i < E; ++i) acc0[i] = acc1[i] = 0.0f;
    float2* line = lines + wave * LINE;
    float2 xa[E], xb[E];  // X[k] and X[M - k] of this wave's row, next kernel
    auto fetch = [&](int k) {
        const float2* row = YT + (long long)k * spec + (long long)bx * LH * RB + wave;  // kx at row[kx * RB]
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const int n = lane + WAVE * i;
            xa[i] = row[(long long)n * RB];
            xb[i] = row[(long long)(M - n) * RB];
        }
    };
    if (wk.k_count > 0) fetch(klist[wk.k_first]);
    __syncthreads();  // twiddles staged
    for (int j = 0; j < wk.k_count; ++j) {
        const int k = klist[wk.k_first + j];
        float2 e[E];
#pragma unroll
        for (int i = 0; i < E; ++i) {
            const float2 a = xa[i], b = xb[i];
            const float2 sm = make_float2(a.x + b.x, a.y - b.y);
            const float2 t = cmul(make_float2(a.x - b.x, a.y + b.y), make_float2(w[i].x, -w[i].y));
            e[i] = make_float2(sm.x - t.y, sm.y + t.x);
        }
        fetch(klist[wk.k_first + min(j + 1, wk.k_count - 1)]);  // the last trip refetches its own
        wave_fft<M, +1>(e, line, tw);
        const GrowthParams p = gp[k];
        // exp(-d^2 inv) as 2^(-d^2 inv log2 e) on the hardware exp2 (v_exp_f32): the growth
        // value differs from a libm expf by < 1e-7 absolute, far inside the FFT's own rounding
        const float k2 = p.inv * 1.44269504088896341f;
#pragma unroll
        for (int i = 0; i < E; ++i) {
            // growth_accumulate (growth.rs:51-57): g = 2 exp(-d*d*inv) - 1; t += w*g
            const float c0 = e[i].x * scale, c1 = e[i].y * scale;
            const float d0 = c0 - p.mu, d1 = c1 - p.mu;
            const float g0 = 2.0f * __builtin_amdgcn_exp2f(-(d0 * d0) * k2) - 1.0f;
            const float g1 = 2.0f * __builtin_amdgcn_exp2f(-(d1 * d1) * k2) - 1.0f;
            acc0[i] += p.weight * g0;
            acc1[i] += p.weight * g1;
        }
    }
    float2* u = reinterpret_cast<float2*>(U + (long long)wk.target * N + (long long)(bx * RB + wave) * L);
#pragma unroll
    for (int i = 0; i < E; ++i) u[lane + WAVE * i] = make_float2(acc0[i], acc1[i]);  // cells 2n, 2n + 1
}

template <int L>
void rows_fwd_t(hipStream_t st, const float* A, float2* XT, const int* src_chan, int nsrc, const Geo& g,
                const float2* tw, const float2* tws) {
    const int nblocks = g.H / RB * nsrc;
    hipLaunchKernelGGL(k_rows_fwd_t<L>, dim3((nblocks + 7) / 8 * 8), dim3(NT), 0, st, A, XT, src_chan, g.H, g.N, g.spec,
                       nblocks, tw, tws);
}

template <int L>
void spectral_t(hipStream_t st, const float2* XT, float2* YT, const float* KqT, const SpecWork* work, int nwork,
                const int* klist, const Geo& g, const float2* tw, const float2* tws) {
    const long long kq_stride = (long long)g.Wh * g.Hq;
    const int ncp = (g.Wh + WAVES / 2 - 1) / (WAVES / 2);  // column pairs
    const int nblocks = ncp * nwork;
    hipLaunchKernelGGL(k_spectral_t<L>, dim3((nblocks + 7) / 8 * 8), dim3(NT), 0, st, XT, YT, KqT, work, klist, g.Wh,
                       g.spec, kq_stride, ncp, nblocks, tw, tws);
}

template <int L>
void rows_inv_t(hipStream_t st, const float2* YT, float* U, const TgtWork* work, int nwork, const int* klist,
                const GrowthParams* gp, const Geo& g, const float2* tw, const float2* tws) {
    const float scale = 1.0f / (float)(g.N);  // fft.rs:389: 1 / (W*H) as f32
    const int nblocks = g.H / RB * nwork;
    hipLaunchKernelGGL(k_rows_inv_t<L>, dim3((nblocks + 7) / 8 * 8), dim3(NT), 0, st, YT, U, work, klist, gp, g.H, g.N,
                       g.spec, scale, nblocks, tw, tws);
}

}  // namespace wv

bool wave_fft_supported(int W, int H) {
    auto ok = [](int n) { return n == 256 || n == 512 || n == 1024 || n == 2048; };
    return ok(W) && ok(H);
}

#define HP_WAVE_SIZES(X) X(256) X(512) X(1024) X(2048)

std::vector<int> wave_stage_twiddle_index(int L) {
    std::vector<int> idx;
#define X_(n) if (L == n) for (int i = 0; i < wv::tw_total(n); ++i) idx.push_back(wv::stage_twiddle_src(n, i));
    X_(128) X_(256) X_(512) X_(1024) X_(2048)
#undef X_
    return idx;
}

bool launch_rows_fwd_t(hipStream_t st, const float* A, float2* XT, const int* src_chan, int nsrc, const Geo& g,
                       const float2* tw, const float2* tws) {
#define X_(n) if (g.W == n) { wv::rows_fwd_t<n>(st, A, XT, src_chan, nsrc, g, tw, tws); return true; }
    HP_WAVE_SIZES(X_)
#undef X_
    return false;
}

bool launch_spectral_t(hipStream_t st, const float2* XT, float2* YT, const float* KqT, const SpecWork* work,
                       int nwork, const int* klist, const Geo& g, const float2* tw, const float2* tws) {
#define X_(n) if (g.H == n) { wv::spectral_t<n>(st, XT, YT, KqT, work, nwork, klist, g, tw, tws); return true; }
    HP_WAVE_SIZES(X_)
#undef X_
    return false;
}

bool launch_rows_inv_t(hipStream_t st, const float2* YT, float* U, const TgtWork* work, int nwork, const int* klist,
                       const GrowthParams* gp, const Geo& g, const float2* tw, const float2* tws) {
#define X_(n) if (g.W == n) { wv::rows_inv_t<n>(st, YT, U, work, nwork, klist, gp, g, tw, tws); return true; }
    HP_WAVE_SIZES(X_)
#undef X_
    return false;
}

}  // namespace hp
