// adjoint_stencil.hip -- reverse-mode stages 2-5 of the 2D Flow Lenia step on gfx950.
//
// The forward step (propagator.rs:303-371) maps the state A and affinity U to the next
// state through the channel sum, Sobel gradients, alpha, the flow F and reintegration.
// Its transpose, for one step with lam' = dL/dA_{t+1}:
//
//   flow pass  (one thread per source cell, per channel):
//     recompute F_c exactly as the forward (same f32 operations, -ffp-contract=off), walk the
//     source's box exactly as the scatter does and read lam'_c at every destination:
//       dL/dm      = sum lam' * area / (2s)^2                     -> lam_t,c (first term)
//       dL/d(dest) = sum lam' * m * d(area)/d(dest) / (2s)^2,  dL/dF = dt * dL/d(dest)
//     then through F = (1 - alpha) grad U - alpha grad A_sum:
//       dL/d(grad U_c) = (1 - alpha) dL/dF_c,  dL/d(grad A_sum) = -alpha sum_c dL/dF_c,
//       dL/dalpha = -sum_c dL/dF_c . (grad U_c + grad A_sum),  dL/dA_sum += dL/dalpha * alpha'
//   divergence pass (one thread per cell):
//     the Sobel operator is antisymmetric along its derivative axis, so its transpose is its
//     negative: dL/dA_sum = -Sobel . dL/d(grad A_sum) + (alpha part), added to every channel;
//     dL/dU_c = -Sobel . dL/d(grad U_c).
//
// No atomics: the adjoint of the scatter is a gather over the source's own box, so every
// output cell has exactly one writer.  Subgradient conventions at the kinks:
//   d overlap / d dest = [hi < c+1] - [lo > c] on cells with positive area;
//   alpha' = 0 where the clamp is active ((m/beta)^n outside (0, 1));
//   a source below the forward's |m| < 1e-10 guard keeps its mass derivative (the step
//   without the guard) and has no flow derivative (it is m * (...) = O(1e-10));
//   the point-mass fallback ((2s)^2 < 1e-10) passes lam' of the rounded destination.
#include "stencil_common.hpp"

namespace hp {

namespace adj {
constexpr int TX = 64, TY = 4, NT = TX * TY;
constexpr int HW = TX + 2, HH = TY + 2;
}  // namespace adj

// Tiles are items in row-major tile order; a tile reads one or two rows of the tiles above and
// below as halo.  Chunks of 64 consecutive tiles (eight tile rows of a 512-wide member) run on
// one XCD (xcd_chunk_item), so most of those rows are read from memory once.
constexpr int kTileChunk = 64;

// alpha'(m) on the open interval where clamp((m/beta)^n, 0, 1) is inactive.
__device__ __forceinline__ float dalpha_of(float mass, float beta_a, float n) {
    const float r = mass / beta_a;
    float p, d;
    if (n == 1.0f) {
        p = r;
        d = 1.0f;
    } else if (n == 2.0f) {
        p = r * r;
        d = 2.0f * r;
    } else if (n == 3.0f) {
        p = r * r * r;
        d = 3.0f * r * r;
    } else if (n == 4.0f) {
        p = (r * r) * (r * r);
        d = 4.0f * r * r * r;
    } else {
        p = powf(r, n);
        d = n * powf(r, n - 1.0f);
    }
    if (!(p > 0.0f && p < 1.0f)) return 0.0f;
    return d / beta_a;
}

// items[b] = (member, tile); tiles of TX x TY cells, row-major over ceil(W/TX) columns.
// li: lam[li] holds dL/dA_{t+1}; lam[1 - li] receives the first term of dL/dA_t.
__global__ __launch_bounds__(adj::NT) void kb_adj_flow(const MemberDesc* __restrict__ mem,
                                                      const int4* __restrict__ items, int nitems, int cur, int li) {
    using namespace adj;
    __shared__ float sA[HH][HW];
    __shared__ float sU[HH][HW];
    const int lin = xcd_chunk_item<kTileChunk>(nitems);
    if (lin < 0) return;  // before any barrier: the whole workgroup leaves
    const int4 it = items[lin];
    const MemberDesc& m = mem[it.x];
    const int W = m.g.W, H = m.g.H, C = m.C;
    const long long N = m.g.N;
    const int tiles_x = (W + TX - 1) / TX;
    const int tx0 = (it.y % tiles_x) * TX, ty0 = (it.y / tiles_x) * TY;
    const auto A = gpc(m.A[cur]);
    const auto lamn = gpc(m.lam[li]);
    const auto lamo = gp(m.lam[1 - li]);
    const auto gscr = gp(m.gscr);
    const float dt = m.dt, s = m.s, beta_a = m.beta_a, nexp = m.nexp;
    const float ta = (2.0f * s) * (2.0f * s);
    const float inv_ta = 1.0f / ta;
    // The halo tile's cells this thread loads (at most two), wrapped once for every channel.
    static_assert(HH * HW <= 2 * NT, "two halo cells per thread");
    long long toff[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int t = threadIdx.x + k * NT;
        const int r = t / HW, c = t - r * HW;
        toff[k] = t < HH * HW ? (long long)wrap_near(ty0 + r - 1, H) * W + wrap_near(tx0 + c - 1, W) : 0;
    }
    float* const sAf = &sA[0][0];
    float* const sUf = &sU[0][0];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int t = threadIdx.x + k * NT;
        if (t >= HH * HW) break;
        float v = 0.0f;
        for (int ch = 0; ch < C; ++ch) v += A[(long long)ch * N + toff[k]];  // 0 + A_0 + A_1 + ...
        sAf[t] = v;
    }
    __syncthreads();
    const int lx = threadIdx.x % TX, ly = threadIdx.x / TX;
    const int x = tx0 + lx, y = ty0 + ly;
    const bool inside = x < W && y < H;
    const long long i = (long long)y * W + x;
    float gax = 0.0f, gay = 0.0f, asum = 0.0f, al = 0.0f;
    if (inside) {
        asum = sA[ly + 1][lx + 1];
        sobel9(sA[ly][lx], sA[ly][lx + 1], sA[ly][lx + 2], sA[ly + 1][lx], sA[ly + 1][lx + 2], sA[ly + 2][lx],
               sA[ly + 2][lx + 1], sA[ly + 2][lx + 2], gax, gay);
        al = alpha_of(asum, beta_a, nexp);
    }
    const float oma = 1.0f - al;
    float gGAx = 0.0f, gGAy = 0.0f, gal = 0.0f;
    for (int c = 0; c < C; ++c) {
        const auto Uc = gpc(m.U) + (long long)c * N;
        __syncthreads();  // previous channel's sU reads are done
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = threadIdx.x + k * NT;
            if (t >= HH * HW) break;
            sUf[t] = Uc[toff[k]];
        }
        __syncthreads();
        if (!inside) continue;
        float gux, guy;
        sobel9(sU[ly][lx], sU[ly][lx + 1], sU[ly][lx + 2], sU[ly + 1][lx], sU[ly + 1][lx + 2], sU[ly + 2][lx],
               sU[ly + 2][lx + 1], sU[ly + 2][lx + 2], gux, guy);
        const float fx = oma * gux - al * gax;
        const float fy = oma * guy - al * gay;
        const float mass = A[(long long)c * N + i];
        const bool active = !(fabsf(mass) < 1e-10f);
        const float dxp = (float)x + dt * fx;  // reintegration.rs:81-82
        const float dyp = (float)y + dt * fy;
        const auto ln = lamn + (long long)c * N;
        float dm = 0.0f, gx = 0.0f, gy = 0.0f;
        if (ta < 1e-10f) {  // point fallback (reintegration.rs:117-124)
            dm = ln[(long long)wrapi(sat_i32(roundf(dyp)), H) * W + wrapi(sat_i32(roundf(dxp)), W)];
        } else {
            const float xm = dxp - s, xM = dxp + s, ym = dyp - s, yM = dyp + s;
            const int ixl = max(sat_i32(floorf(xm)), -W), ixh = min(sat_i32(ceilf(xM)), 2 * W);
            const int iyl = max(sat_i32(floorf(ym)), -H), iyh = min(sat_i32(ceilf(yM)), 2 * H);
            if (ixl > ixh || iyl > iyh) {
                // an empty box range (a coordinate past 2W or 2H): no cell overlaps
            } else if (ixh - ixl <= 3 && iyh - iyl <= 3 && W > 2 && H > 2) {
                // At most three columns and three rows (every box with s <= 1): cell ixl + 3 starts
                // at or past xM, so it and every later cell have zero overlap, as do the rows past
                // iyl + 2.  The nine values of lam' are loaded at once (wrapped, always in bounds)
                // and the zero-overlap cells are skipped in the order of the loops below.  With
                // -W <= ixl <= ixh <= 2W every index lies in [-W, 3W), so wrap3 applies.
                float ow[3], dow[3], oh[3], doh[3];
                int cw[3];
                long long rw[3];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const int ix = ixl + j, iy = iyl + j;
                    const float cx0 = (float)ix, cx1 = (float)(ix + 1);
                    const float cy0 = (float)iy, cy1 = (float)(iy + 1);
                    ow[j] = j <= ixh - ixl ? overlap1(cx0, cx1, xm, xM) : 0.0f;
                    dow[j] = (float)(xM < cx1) - (float)(xm > cx0);
                    cw[j] = wrap3(ix, W);
                    oh[j] = j <= iyh - iyl ? overlap1(cy0, cy1, ym, yM) : 0.0f;
                    doh[j] = (float)(yM < cy1) - (float)(ym > cy0);
                    rw[j] = (long long)wrap3(iy, H) * W;
                }
                float l[3][3];
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int j = 0; j < 3; ++j) l[r][j] = ln[rw[r] + cw[j]];
#pragma unroll
                for (int r = 0; r < 3; ++r) {
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        const bool on = ow[j] * oh[r] > 0.0f;
                        const float a = dm + l[r][j] * (ow[j] * oh[r]);
                        const float b = gx + l[r][j] * (dow[j] * oh[r]);
                        const float d = gy + l[r][j] * (ow[j] * doh[r]);
                        dm = on ? a : dm;
                        gx = on ? b : gx;
                        gy = on ? d : gy;
                    }
                }
            } else if (ixh - ixl <= 4) {
                // At most five columns (every box with s <= 1): their overlaps, derivatives and
                // wrapped indices once, reused on every row.  Columns past ixh have zero overlap
                // and are skipped as the loop below skips them; the order is the loop's.
                float ow[5], dow[5];
                int cw[5];
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    const int ix = ixl + j;
                    const float cx0 = (float)ix, cx1 = (float)(ix + 1);
                    ow[j] = j <= ixh - ixl ? overlap1(cx0, cx1, xm, xM) : 0.0f;
                    dow[j] = (float)(xM < cx1) - (float)(xm > cx0);
                    cw[j] = wrap_near(ix, W);
                }
                for (int iy = iyl; iy <= iyh; ++iy) {
                    const float cy0 = (float)iy, cy1 = (float)(iy + 1);
                    const float oh = overlap1(cy0, cy1, ym, yM);
                    if (!(oh > 0.0f)) continue;
                    const float doh = (float)(yM < cy1) - (float)(ym > cy0);
                    const auto lrow = ln + (long long)wrap_near(iy, H) * W;
#pragma unroll
                    for (int j = 0; j < 5; ++j) {
                        if (!(ow[j] * oh > 0.0f)) continue;
                        const float l = lrow[cw[j]];
                        dm += l * (ow[j] * oh);
                        gx += l * (dow[j] * oh);
                        gy += l * (ow[j] * doh);
                    }
                }
            } else {
                for (int iy = iyl; iy <= iyh; ++iy) {
                    const float cy0 = (float)iy, cy1 = (float)(iy + 1);
                    const float oh = overlap1(cy0, cy1, ym, yM);
                    if (!(oh > 0.0f)) continue;
                    const float doh = (float)(yM < cy1) - (float)(ym > cy0);
                    const auto lrow = ln + (long long)wrap_near(iy, H) * W;
                    for (int ix = ixl; ix <= ixh; ++ix) {
                        const float cx0 = (float)ix, cx1 = (float)(ix + 1);
                        const float ow = overlap1(cx0, cx1, xm, xM);
                        if (!(ow * oh > 0.0f)) continue;
                        const float dow = (float)(xM < cx1) - (float)(xm > cx0);
                        const float l = lrow[wrap_near(ix, W)];
                        dm += l * (ow * oh);
                        gx += l * (dow * oh);
                        gy += l * (ow * doh);
                    }
                }
            }
            dm *= inv_ta;
            const float k = active ? mass * inv_ta * dt : 0.0f;
            gx *= k;
            gy *= k;
        }
        lamo[(long long)c * N + i] = dm;
        gscr[(long long)(2 * c) * N + i] = oma * gx;
        gscr[(long long)(2 * c + 1) * N + i] = oma * gy;
        gGAx -= al * gx;
        gGAy -= al * gy;
        gal -= gx * (gux + gax) + gy * (guy + gay);
    }
    if (inside) {
        gscr[(long long)(2 * C) * N + i] = gGAx;
        gscr[(long long)(2 * C + 1) * N + i] = gGAy;
        gscr[(long long)(2 * C + 2) * N + i] = gal * dalpha_of(asum, beta_a, nexp);
    }
}

// Sobel transpose at one cell: -(Sobel_x(ax) + Sobel_y(ay)) from the 3x3 neighbourhood.
__device__ __forceinline__ float sobel_adj(gptr<const float> ax, gptr<const float> ay, long long rp, long long rc,
                                           long long rn, int xp, int x, int xn) {
    float sx, sy, dummy;
    sobel9(ax[rp + xp], ax[rp + x], ax[rp + xn], ax[rc + xp], ax[rc + xn], ax[rn + xp], ax[rn + x], ax[rn + xn], sx,
           dummy);
    sobel9(ay[rp + xp], ay[rp + x], ay[rp + xn], ay[rc + xp], ay[rc + xn], ay[rn + xp], ay[rn + x], ay[rn + xn],
           dummy, sy);
    return -(sx + sy);
}

// items as kb_adj_flow.  lam[1 - li] += dL/dA_sum (every channel); lamU_c = dL/dU_c.
__global__ __launch_bounds__(adj::NT) void kb_adj_div(const MemberDesc* __restrict__ mem,
                                                     const int4* __restrict__ items, int nitems, int li) {
    using namespace adj;
    const int lin = xcd_chunk_item<kTileChunk>(nitems);
    if (lin < 0) return;
    const int4 it = items[lin];
    const MemberDesc& m = mem[it.x];
    const int W = m.g.W, H = m.g.H, C = m.C;
    const long long N = m.g.N;
    const int tiles_x = (W + TX - 1) / TX;
    const int x = (it.y % tiles_x) * TX + (int)threadIdx.x % TX;
    const int y = (it.y / tiles_x) * TY + (int)threadIdx.x / TX;
    if (x >= W || y >= H) return;
    const int xp = x == 0 ? W - 1 : x - 1, xn = x == W - 1 ? 0 : x + 1;
    const int yp = y == 0 ? H - 1 : y - 1, yn = y == H - 1 ? 0 : y + 1;
    const long long rp = (long long)yp * W, rc = (long long)y * W, rn = (long long)yn * W;
    const long long i = rc + x;
    const auto g = gpc(m.gscr);
    const float lsum = sobel_adj(g + (long long)(2 * C) * N, g + (long long)(2 * C + 1) * N, rp, rc, rn, xp, x, xn) +
                       g[(long long)(2 * C + 2) * N + i];
    const auto lamo = gp(m.lam[1 - li]);
    const auto lamU = gp(m.lamU);
    for (int c = 0; c < C; ++c) {
        lamo[(long long)c * N + i] += lsum;
        lamU[(long long)c * N + i] =
            sobel_adj(g + (long long)(2 * c) * N, g + (long long)(2 * c + 1) * N, rp, rc, rn, xp, x, xn);
    }
}

void launch_adj_flow(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur, int li) {
    hipLaunchKernelGGL(kb_adj_flow, dim3(xcd_chunk_grid<kTileChunk>(nitems)), dim3(adj::NT), 0, st, mem, items, nitems,
                       cur, li);
}

void launch_adj_div(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int li) {
    hipLaunchKernelGGL(kb_adj_div, dim3(xcd_chunk_grid<kTileChunk>(nitems)), dim3(adj::NT), 0, st, mem, items, nitems,
                       li);
}

int adj_tile_w() { return adj::TX; }
int adj_tile_h() { return adj::TY; }

// ---- tape copies and clears -----------------------------------------------------------------
// items[b] = (member, part, nparts): block b strides over the member's C*N floats.
// to_tape: tape[slot] <- A[cur]; else A[cur] <- tape[slot].
// dst[j] = src[j] for j = j0, j0 + stride, ... below n (the buffers do not overlap), with four
// loads in flight per thread: one load and its store per iteration waited for every load.
__device__ __forceinline__ void copy_strided(gptr<float> dst, gptr<const float> src, long long j, long long n,
                                             long long stride) {
    for (; j + 3 * stride < n; j += 4 * stride) {
        const float v0 = src[j], v1 = src[j + stride], v2 = src[j + 2 * stride], v3 = src[j + 3 * stride];
        dst[j] = v0;
        dst[j + stride] = v1;
        dst[j + 2 * stride] = v2;
        dst[j + 3 * stride] = v3;
    }
    for (; j < n; j += stride) dst[j] = src[j];
}

__global__ __launch_bounds__(256) void kb_tape_copy(const MemberDesc* __restrict__ mem, const int4* __restrict__ items,
                                                   int cur, int slot, int to_tape) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const long long n = (long long)m.C * m.g.N;
    float* a = m.A[cur];
    float* t = slot < 0 ? m.snap : m.tape + (long long)slot * n;
    copy_strided(gp(to_tape ? t : a), gpc(to_tape ? a : t), (long long)it.y * 256 + threadIdx.x, n,
                 (long long)it.z * 256);
}

// U of a taped step to / from its tape slot (items as kb_tape_copy).
__global__ __launch_bounds__(256) void kb_u_copy(const MemberDesc* __restrict__ mem, const int4* __restrict__ items,
                                                int slot, int to_tape) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const long long n = (long long)m.C * m.g.N;
    float* t = m.tape_u + (long long)slot * n;
    copy_strided(gp(to_tape ? t : m.U), gpc(to_tape ? m.U : t), (long long)it.y * 256 + threadIdx.x, n,
                 (long long)it.z * 256);
}

// Point each member's A[0] at tape slot `slot`, and its U and source spectra X at tape_u and
// tape_x slot `u_slot` (a negative slot leaves those pointers): a taped forward step writes its
// U and X straight into their slots, and a backward step reads its state, U and X where the
// tape keeps them (X is then overwritten in place, each slot serving one step).  One thread
// per member; the host restores the descriptors afterwards.
__global__ __launch_bounds__(256) void kb_point_slots(MemberDesc* __restrict__ mem, int n, int slot, int u_slot) {
    const int i = blockIdx.x * 256 + (int)threadIdx.x;
    if (i >= n) return;
    MemberDesc& m = mem[i];
    const long long cn = (long long)m.C * m.g.N;
    if (slot >= 0) m.A[0] = m.tape + (long long)slot * cn;
    if (u_slot >= 0) {
        m.U = m.tape_u + (long long)u_slot * cn;
        m.X = m.tape_x + (long long)u_slot * m.nsrc * m.g.spec;
    }
}

// Clear the accumulators of a backward pass: P, pacc and the kernel-tap windows.
__global__ __launch_bounds__(256) void kb_tape_clear(const MemberDesc* __restrict__ mem,
                                                    const int4* __restrict__ items) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const long long stride = (long long)it.z * 256, j0 = (long long)it.y * 256 + threadIdx.x;
    const long long np = (long long)m.K * m.g.spec;
    for (long long j = j0; j < np; j += stride) m.P[j] = make_float2(0.0f, 0.0f);
    const long long na = (long long)m.K * ((m.g.H + 1) / 2) * 3;
    for (long long j = j0; j < na; j += stride) m.pacc[j] = 0.0;
    long long nw = 0;
    for (int k = 0; k < m.K; ++k) nw += (long long)(2 * m.dkw[k].y + 1) * (2 * m.dkw[k].y + 1);
    for (long long j = j0; j < nw; j += stride) m.dkwin[j] = 0.0;
}

void launch_tape_copy(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int cur, int slot,
                      bool to_tape) {
    hipLaunchKernelGGL(kb_tape_copy, dim3(nitems), dim3(256), 0, st, mem, items, cur, slot, to_tape ? 1 : 0);
}

void launch_point_slots(hipStream_t st, MemberDesc* mem, int n, int slot, int u_slot) {
    hipLaunchKernelGGL(kb_point_slots, dim3((n + 255) / 256), dim3(256), 0, st, mem, n, slot, u_slot);
}

void launch_u_copy(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, int slot, bool to_tape) {
    hipLaunchKernelGGL(kb_u_copy, dim3(nitems), dim3(256), 0, st, mem, items, slot, to_tape ? 1 : 0);
}

void launch_tape_clear(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems) {
    hipLaunchKernelGGL(kb_tape_clear, dim3(nitems), dim3(256), 0, st, mem, items);
}

// Every member's C*N adjoint values to / from one contiguous staging array, so that a host
// transfer of the whole population is one copy.
__global__ __launch_bounds__(256) void kb_stage_lam(const MemberDesc* __restrict__ mem, const int4* __restrict__ items,
                                                   float* __restrict__ stage, int li, int to_stage) {
    const int4 it = items[blockIdx.x];
    const MemberDesc& m = mem[it.x];
    const long long n = (long long)m.C * m.g.N;
    float* l = m.lam[li];
    float* s = stage + m.stage_off;
    copy_strided(gp(to_stage ? s : l), gpc(to_stage ? l : s), (long long)it.y * 256 + threadIdx.x, n,
                 (long long)it.z * 256);
}

void launch_stage_lam(hipStream_t st, const MemberDesc* mem, const int4* items, int nitems, float* stage, int li,
                      bool to_stage) {
    hipLaunchKernelGGL(kb_stage_lam, dim3(nitems), dim3(256), 0, st, mem, items, stage, li, to_stage ? 1 : 0);
}


}  // namespace hp
