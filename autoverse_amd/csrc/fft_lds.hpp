// fft_lds.hpp -- the LDS Stockham FFT core shared by the forward and adjoint passes.
//
// In-place Stockham autosort transforms in LDS: each stage loads its radix-R
// butterfly inputs into registers, synchronises, and stores the outputs; radices
// 8/4/2/3 cover lengths 2^a * 3^b.  Twiddles come from a table exp(-2 pi i m / L)
// (host FP64, rounded to f32) staged in LDS.
#pragma once

#include "hp_common.hpp"

namespace hp {

// ---- compile-time radix plan ------------------------------------------------------
// RP selects the radix order: 0 takes radix 8 first (then 4, 2, 3; 16 splits as 4 x 4),
// 1 takes radix 4 first (more, lighter stages).  Which is faster depends on the length
// and the occupancy of the pass; the plan autotuner measures both.
__host__ __device__ constexpr int pick_radix(int m, int rp = 0) {
    return rp == 0 ? ((m % 8 == 0 && m != 16) ? 8 : (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 1)
                   : ((m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : 1);
}
__host__ __device__ constexpr int num_stages(int L, int rp = 0) {
    int n = 0;
    while (L > 1) {
        L /= pick_radix(L, rp);
        ++n;
    }
    return n;
}
__host__ __device__ constexpr int radix_at(int L, int s, int rp = 0) {
    for (int i = 0; i < s; ++i) L /= pick_radix(L, rp);
    return pick_radix(L, rp);
}
__host__ __device__ constexpr int ns_at(int L, int s, int rp = 0) {  // product of the radices before stage s
    int ns = 1;
    for (int i = 0; i < s; ++i) {
        int r = pick_radix(L, rp);
        ns *= r;
        L /= r;
    }
    return ns;
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by -i (DIR = -1, forward) or +i (DIR = +1, inverse)
template <int DIR>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// In-register DFT of size R with exponent sign DIR.
template <int R, int DIR>
__device__ __forceinline__ void dft(float2* v) {
    if constexpr (R == 2) {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 3) {
        const float c1 = -0.5f, s1 = 0.86602540378443864676f;
        float2 t = cadd(v[1], v[2]);
        float2 d = csub(v[1], v[2]);
        float2 m = make_float2(v[0].x + c1 * t.x, v[0].y + c1 * t.y);
        float2 j = mul_mi<DIR>(make_float2(s1 * d.x, s1 * d.y));  // -i*s*d (fwd)
        v[0] = cadd(v[0], t);
        v[1] = cadd(m, j);
        v[2] = csub(m, j);
    } else if constexpr (R == 4) {
        float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        float2 b0 = cadd(v[1], v[3]), b1 = mul_mi<DIR>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    } else if constexpr (R == 8) {
        float2 e[4] = {v[0], v[2], v[4], v[6]};
        float2 o[4] = {v[1], v[3], v[5], v[7]};
        dft<4, DIR>(e);
        dft<4, DIR>(o);
        const float h = 0.70710678118654752440f;
        // W8^1, W8^2, W8^3 with exponent sign DIR
        float2 o1 = cmul(o[1], make_float2(h, DIR * h));
        float2 o2 = mul_mi<DIR>(o[2]);
        float2 o3 = cmul(o[3], make_float2(-h, DIR * h));
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], o1);
        v[5] = csub(e[1], o1);
        v[2] = cadd(e[2], o2);
        v[6] = csub(e[2], o2);
        v[3] = cadd(e[3], o3);
        v[7] = csub(e[3], o3);
    }
}

// The first stage writes its butterflies' outputs R0 apart (R0 its radix): in the plain
// layout the 16 lanes of a quarter-wave then share two or four of the 32 banks.  Between the
// first and the second stage a transform's element i is kept at local index
// i ^ ((i >> 4) & (R0 - 1)) instead, a permutation inside each aligned group of R0 elements
// (so inside the transform) that spreads those writes over all banks; the second stage reads
// through the same map, and every other stage, the caller's input and the output use the plain
// layout.  Only the LDS addresses change: every element takes part in the same arithmetic.
template <int L, int RP>
__host__ __device__ constexpr int swizzle_mask() {
    constexpr int R0 = radix_at(L, 0, RP);
    return num_stages(L, RP) >= 2 && (R0 == 2 || R0 == 4 || R0 == 8) ? R0 - 1 : 0;
}

// One Stockham stage over NFFT transforms of length L at buf[f*LP + i], run by NT threads.
template <int L, int STAGE, int DIR, int NFFT, int NT, int LP, int RP>
__device__ __forceinline__ void fft_stage(float2* buf, const float2* tw) {
    constexpr int R = radix_at(L, STAGE, RP);
    constexpr int NS = ns_at(L, STAGE, RP);
    constexpr int NB = L / R;
    constexpr int TOT = NFFT * NB;
    constexpr int Q = (TOT + NT - 1) / NT;
    constexpr int TSTEP = L / (NS * R);
    constexpr int SWM = swizzle_mask<L, RP>();
    constexpr bool SW_IN = SWM != 0 && STAGE == 1, SW_OUT = SWM != 0 && STAGE == 0;
    float2 v[Q][R];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = threadIdx.x + q * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int f = b / NB, j = b - f * NB;
            const float2* src = buf + f * LP;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = j + r * NB;
                v[q][r] = src[SW_IN ? i ^ ((i >> 4) & SWM) : i];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int b = threadIdx.x + q * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int f = b / NB, j = b - f * NB;
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
            const int o = (j / NS) * NS * R + k;
            float2* dst = buf + f * LP;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = o + r * NS;
                dst[SW_OUT ? i ^ ((i >> 4) & SWM) : i] = v[q][r];
            }
        }
    }
    __syncthreads();
}

template <int L, int STAGE, int DIR, int NFFT, int NT, int LP, int RP>
__device__ __forceinline__ void fft_stages(float2* buf, const float2* tw) {
    if constexpr (STAGE < num_stages(L, RP)) {
        fft_stage<L, STAGE, DIR, NFFT, NT, LP, RP>(buf, tw);
        fft_stages<L, STAGE + 1, DIR, NFFT, NT, LP, RP>(buf, tw);
    }
}

// NFFT in-place transforms of length L (unnormalised), DIR = -1 forward, +1 inverse, radix
// order RP.  Caller synchronises before (inputs written); returns after a barrier.
template <int L, int DIR, int NFFT, int NT, int LP, int RP = 0>
__device__ __forceinline__ void lds_fft(float2* buf, const float2* tw) {
    fft_stages<L, 0, DIR, NFFT, NT, LP, RP>(buf, tw);
}

// Copy the length-L twiddle table into LDS, workgroup-cooperative: every load first, then the
// stores (in one loop the compiler reuses one register pair and waits for each load in turn).
// PT: a plain pointer (kernel argument) or gptr<const float2> (read from a member descriptor).
template <int NT, int L, class PT>
__device__ __forceinline__ void stage_twiddles(float2* tw, PT g) {
    constexpr int PER = (L + NT - 1) / NT;
    float2 v[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        if (L % NT == 0 || i < L) v[q] = ld2(g + i);
    }
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int i = threadIdx.x + q * NT;
        if (L % NT == 0 || i < L) tw[i] = v[q];
    }
}

// Launch shapes per length: threads of a row pass, columns per spectral workgroup and
// threads of a spectral workgroup.  CS splits the column block (1: up to 8192 elements per
// workgroup; 2: half as many columns, less LDS per workgroup, more workgroups).
template <int L, int CS = 1>
struct Cfg {
    static constexpr int ROW_NT = L / 8 < 64 ? 64 : (L / 8 > 256 ? 256 : L / 8);
    static constexpr int CB0 = 8192 / L < 1 ? 1 : (8192 / L > 64 ? 64 : 8192 / L);
    static constexpr int CBP = CB0 / CS < 1 ? 1 : CB0 / CS;  // columns per block before balancing
    // The 2D spectral pass: a square grid's L/2 + 1 half-spectrum columns in near-equal blocks
    // of at most CBP (65 columns as 33 + 32 rather than 64 + 1), and about 4 elements per thread
    // (whole wavefronts, up to 1024 threads): fewer elements held per thread keep registers
    // low and occupancy up (at 32 per thread the pass ran one wave per SIMD; the 256-member
    // population's step went from 703 to 544 us).
    static constexpr int WH = L / 2 + 1;
    static constexpr int NBLK = (WH + CBP - 1) / CBP;
    static constexpr int CB = (WH + NBLK - 1) / NBLK;
    static constexpr int NT_RAW = ((CB * L) / 4 + 63) / 64 * 64;
    static constexpr int NT_MAX = L % 3 == 0 ? 512 : 1024;  // radix-3 stages spill at 1024 threads
    static constexpr int COL_NT = NT_RAW < 64 ? 64 : (NT_RAW > NT_MAX ? NT_MAX : NT_RAW);
    // The 3D line passes run narrower workgroups, so that a 128^3 grid spreads over several
    // hundred to a thousand workgroups on the 256 CUs instead of 160-256 single-wave ones: the y
    // column pass at most 16 lines and a thread per radix-8 butterfly, the z spectral pass at
    // most 8 lines and a thread per radix-4 butterfly (the fastest shapes measured at 128^3).
    // A line's arithmetic does not depend on the launch shape.
    static constexpr int CBC3 = CBP < 16 ? CBP : 16;
    static constexpr int COLC3_NT = (CBC3 * L) / 8 < 64 ? 64 : ((CBC3 * L) / 8 > 256 ? 256 : (CBC3 * L) / 8);
    static constexpr int CBS3 = CBP < 8 ? CBP : 8;
    static constexpr int SPEC3_NT = (CBS3 * L) / 4 < 64 ? 64 : ((CBS3 * L) / 4 > 256 ? 256 : (CBS3 * L) / 4);
};

}  // namespace hp
