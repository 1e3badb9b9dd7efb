// host_seed.cpp -- configuration validation, seed rasterisation and ring-kernel
// construction of the Flow Lenia library (host code, f32 like the reference).
//
// Built with -ffp-contract=off and no fast-math: every expression rounds in the
// same order as the reference's Rust f32 code and calls the same glibc expf /
// sqrtf, so seeds and spatial kernels are identical to the reference's.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "host_internal.hpp"

namespace hp {

// SimulationConfig::validate (config.rs:136-164)
hp_status validate_config(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel) {
    if (!cfg) return HP_ERR_NULL;
    if (cfg->width == 0 || cfg->height == 0 || cfg->depth == 0) return HP_ERR_INVALID_DIMENSIONS;
    if (cfg->channels == 0) return HP_ERR_INVALID_CHANNELS;
    if (cfg->dt <= 0.0f) return HP_ERR_INVALID_TIME_STEP;  // NaN passes, as `self.dt <= 0.0` does
    if (cfg->kernel_radius == 0) return HP_ERR_INVALID_KERNEL_RADIUS;
    if (cfg->num_kernels && !cfg->kernels) return HP_ERR_NULL;
    for (uint32_t i = 0; i < cfg->num_kernels; i++) {
        const hp_kernel& k = cfg->kernels[i];
        uint32_t bad = k.source_channel >= cfg->channels ? k.source_channel
                       : k.target_channel >= cfg->channels ? k.target_channel
                                                            : UINT32_MAX;
        if (bad != UINT32_MAX) {
            if (bad_kernel) *bad_kernel = i;
            if (bad_channel) *bad_channel = bad;
            return HP_ERR_INVALID_CHANNEL_INDEX;
        }
        if (k.ring_count && (!cfg->rings || (uint64_t)k.ring_offset + k.ring_count > cfg->num_rings))
            return HP_ERR_SHAPE;
    }
    return HP_OK;
}

// (config.radius * max_radius as f32).round() as usize  (kernel.rs:33)
uint32_t actual_radius(float radius, uint32_t kernel_radius) {
    const float r = roundf(radius * (float)kernel_radius);
    return r > 0.0f ? (uint32_t)r : 0u;  // saturating: NaN and negatives give 0
}

// Kernel::from_config (kernel.rs:32-86) and Kernel3D::from_config (kernel3d.rs:32-97):
// ring-sum values over the (2R+1)^dim cube, row-major running f32 sum, then * (1/sum).
std::vector<float> build_ring_kernel(const hp_kernel& k, const hp_ring* rings, uint32_t kernel_radius, int dim,
                                     uint32_t* size_out) {
    const uint32_t ar = actual_radius(k.radius, kernel_radius);
    const uint32_t size = ar * 2 + 1;
    const uint32_t zs = dim == 3 ? size : 1;
    std::vector<float> data((size_t)size * size * zs, 0.0f);
    const float center = (float)ar;
    float sum = 0.0f;
    for (uint32_t z = 0; z < zs; z++)
        for (uint32_t y = 0; y < size; y++)
            for (uint32_t x = 0; x < size; x++) {
                const float dx = (float)x - center;
                const float dy = (float)y - center;
                float dist;
                if (dim == 3) {
                    const float dz = (float)z - center;
                    dist = sqrtf(dx * dx + dy * dy + dz * dz);
                } else {
                    dist = sqrtf(dx * dx + dy * dy);
                }
                const float norm_dist = dist / (float)ar;
                if (norm_dist > 1.0f) continue;
                float value = 0.0f;
                for (uint32_t r = 0; r < k.ring_count; r++) {
                    const hp_ring& ring = rings[k.ring_offset + r];
                    const float rd = fabsf(norm_dist - ring.distance);
                    value += ring.amplitude * expf(-rd * rd / (2.0f * ring.width * ring.width));
                }
                data[((size_t)z * size + y) * size + x] = value;
                sum += value;
            }
    if (sum > 0.0f) {
        const float inv = 1.0f / sum;
        for (float& v : data) v *= inv;
    }
    *size_out = size;
    return data;
}

// Folded real spectrum of the wrap-padded kernel (pad_to_size, kernel.rs:98-117):
// K(fx, fy, fz) = sum_taps k(dx, dy, dz) cos(2 pi fx dx / W) cos(2 pi fy dy / H) cos(2 pi fz dz / D),
// which is the DFT of the padded kernel because the taps are symmetric under dx -> -dx etc.
// (the sine terms cancel).  Evaluated separably in FP64, rounded once to f32.
// Output layout [fz][fy][fx], fz < D/2+1, fy < H/2+1, fx < W/2+1.
std::vector<float> folded_spectrum(const std::vector<float>& k, uint32_t size, int dim, uint32_t W, uint32_t H,
                                   uint32_t D) {
    const int R = (int)(size / 2);
    const uint32_t Wh = W / 2 + 1, Hq = H / 2 + 1, Dq = dim == 3 ? D / 2 + 1 : 1;
    const uint32_t zs = dim == 3 ? size : 1;
    auto table = [](uint32_t n) {
        std::vector<double> c(n);
        for (uint32_t m = 0; m < n; m++) c[m] = cos(2.0 * M_PI * (double)m / (double)n);
        return c;
    };
    const std::vector<double> cw = table(W), ch = table(H), cd = table(dim == 3 ? D : 1);
    auto cidx = [](uint32_t f, int d, uint32_t n) -> uint32_t {
        long long v = ((long long)f * d) % (long long)n;
        return (uint32_t)(v < 0 ? v + n : v);
    };
    // pass x: P1[z][y][fx]
    std::vector<double> p1((size_t)zs * size * Wh, 0.0);
    for (uint32_t z = 0; z < zs; z++)
        for (uint32_t y = 0; y < size; y++) {
            const float* row = &k[((size_t)z * size + y) * size];
            double* out = &p1[((size_t)z * size + y) * Wh];
            for (uint32_t fx = 0; fx < Wh; fx++) {
                double acc = 0.0;
                for (uint32_t x = 0; x < size; x++)
                    if (row[x] != 0.0f) acc += (double)row[x] * cw[cidx(fx, (int)x - R, W)];
                out[fx] = acc;
            }
        }
    // pass y: P2[z][fy][fx]
    std::vector<double> p2((size_t)zs * Hq * Wh, 0.0);
    for (uint32_t z = 0; z < zs; z++)
        for (uint32_t fy = 0; fy < Hq; fy++) {
            double* out = &p2[((size_t)z * Hq + fy) * Wh];
            for (uint32_t y = 0; y < size; y++) {
                const double c = ch[cidx(fy, (int)y - R, H)];
                const double* in = &p1[((size_t)z * size + y) * Wh];
                for (uint32_t fx = 0; fx < Wh; fx++) out[fx] += in[fx] * c;
            }
        }
    std::vector<float> res((size_t)Dq * Hq * Wh);
    if (dim != 3) {
        for (size_t i = 0; i < res.size(); i++) res[i] = (float)p2[i];
        return res;
    }
    std::vector<double> p3((size_t)Dq * Hq * Wh, 0.0);
    for (uint32_t fz = 0; fz < Dq; fz++)
        for (uint32_t z = 0; z < size; z++) {
            const double c = cd[cidx(fz, (int)z - R, D)];
            const double* in = &p2[(size_t)z * Hq * Wh];
            double* out = &p3[(size_t)fz * Hq * Wh];
            for (size_t i = 0; i < (size_t)Hq * Wh; i++) out[i] += in[i] * c;
        }
    for (size_t i = 0; i < res.size(); i++) res[i] = (float)p3[i];
    return res;
}

hp_status plan_member(const hp_config* cfg, bool allow_wave, int group, MemberPlan* out, std::string* err) {
    MemberPlan& p = *out;
    const bool is3d = cfg->depth > 1;
    const uint32_t W = cfg->width, H = cfg->height, D = cfg->depth;
    // sides the LDS FFT passes do not cover take the direct convolution
    p.direct = !fft_supported((int)W) || !fft_supported((int)H) || (is3d && !fft_supported((int)D));
    p.C = (int)cfg->channels;
    p.K = (int)cfg->num_kernels;
    p.is3d = is3d;
    Geo& g = p.g;
    g.W = (int)W;
    g.H = (int)H;
    g.D = (int)D;
    g.Wh = (int)(W / 2 + 1);
    g.Wp = (g.Wh + 15) / 16 * 16;
    g.Hq = (int)(H / 2 + 1);
    g.Dq = is3d ? (int)(D / 2 + 1) : 1;
    g.N = (long long)W * H * D;
    g.spec = (long long)D * H * g.Wp;
    p.wave = !p.direct && allow_wave && !is3d && wave_fft_supported(g.W, g.H);
    // kernel spectra (host FP64); the direct path keeps the spatial taps instead
    p.kq.assign(p.direct ? 1 : (size_t)std::max(1, p.K) * g.Dq * g.Hq * g.Wh, 0.0f);
    std::vector<int> slot_of(p.C, -1);
    p.gp.assign(std::max(1, p.K), GrowthParams{0, 0, 0, 0});
    for (int k = 0; k < p.K; k++) {
        const hp_kernel& kc = cfg->kernels[k];
        uint32_t size = 0;
        std::vector<float> data = build_ring_kernel(kc, cfg->rings, cfg->kernel_radius, is3d ? 3 : 2, &size);
        if (size > W || size > H || (is3d && size > D)) {
            *err = "kernel diameter 2R+1 exceeds a grid side";
            return HP_ERR_UNSUPPORTED_SIZE;
        }
        if (p.direct) {
            const int R = (int)size / 2, S = (int)size;
            if (R > direct_max_radius(is3d ? 3 : 2)) {
                *err = "kernel radius " + std::to_string(R) + " is beyond the direct convolution's LDS tile (sides " +
                       "outside 2^a * 3^b up to 4096 take the direct path)";
                return HP_ERR_UNSUPPORTED_SIZE;
            }
            DirectKernel d{};
            d.src = (int)kc.source_channel;
            d.tgt = (int)kc.target_channel;
            d.R = R;
            d.tap_off = (int)p.taps.size();
            d.col_off = (int)p.cols.size();
            const float sigma_sq_2 = 2.0f * kc.sigma * kc.sigma;  // growth.rs:51-52
            d.weight = kc.weight;
            d.mu = kc.mu;
            d.inv = 1.0f / sigma_sq_2;
            p.taps.insert(p.taps.end(), data.begin(), data.end());
            // per column of the other axes, the nonzero range along the last axis
            const int ncol = is3d ? S * S : S, stride = ncol;
            for (int c = 0; c < ncol; c++) {
                int lo = S, hi = -1;
                for (int sl = 0; sl < S; sl++)
                    if (data[(size_t)sl * stride + c] != 0.0f) {
                        lo = std::min(lo, sl);
                        hi = std::max(hi, sl);
                    }
                p.cols.push_back(make_int2(lo, hi));
            }
            p.dk.push_back(d);
            p.rmax = std::max(p.rmax, R);
        }
        std::vector<float> f = p.direct ? std::vector<float>() : folded_spectrum(data, size, is3d ? 3 : 2, W, H, D);
        if (p.wave) {  // column-major [kx][kyf] for the y pass
            std::vector<float> t(f.size());
            for (int fy = 0; fy < g.Hq; fy++)
                for (int fx = 0; fx < g.Wh; fx++) t[(size_t)fx * g.Hq + fy] = f[(size_t)fy * g.Wh + fx];
            f.swap(t);
        }
        if (!p.direct) std::copy(f.begin(), f.end(), p.kq.begin() + (size_t)k * f.size());
        if (slot_of[kc.source_channel] < 0) {
            slot_of[kc.source_channel] = (int)p.src_chan.size();
            p.src_chan.push_back((int)kc.source_channel);
        }
        const float sigma_sq_2 = 2.0f * kc.sigma * kc.sigma;  // growth.rs:51-52
        p.gp[k] = {kc.weight, kc.mu, 1.0f / sigma_sq_2, kc.sigma};
    }
    // spectral work: one item per source (its forward transform is shared by all its kernels)
    // or one item per kernel (more workgroups; the forward transform is redone)
    const int blocks_per_source = p.wave ? (g.Wh + 3) / 4 : 1;
    const bool grouped = group >= 0 ? group != 0 : (!p.wave || blocks_per_source * (long long)p.src_chan.size() >= 256);
    p.grouped = grouped;
    if (grouped) {
        for (size_t sl = 0; sl < p.src_chan.size(); sl++) {
            int first = (int)p.klist_spec.size(), cnt = 0;
            for (int k = 0; k < p.K; k++)
                if (slot_of[cfg->kernels[k].source_channel] == (int)sl) {
                    p.klist_spec.push_back(k);
                    cnt++;
                }
            p.spec_work.push_back({(int)sl, first, cnt, 0});
        }
    } else {
        for (int k = 0; k < p.K; k++) {
            p.spec_work.push_back({slot_of[cfg->kernels[k].source_channel], (int)p.klist_spec.size(), 1, 0});
            p.klist_spec.push_back(k);
        }
    }
    // inverse-row work: one item per target with its kernels in index order (propagator.rs:270-279)
    for (int t = 0; t < p.C; t++) {
        int first = (int)p.klist_tgt.size(), cnt = 0;
        for (int k = 0; k < p.K; k++)
            if ((int)cfg->kernels[k].target_channel == t) {
                p.klist_tgt.push_back(k);
                cnt++;
            }
        if (cnt) {
            p.tgt_work.push_back({t, first, cnt, 0});
            p.tgt_chan.push_back(t);
        }
    }
    return HP_OK;
}

// ---- seeds (seed.rs:160-524) ------------------------------------------------------------
namespace {
struct Grid {
    float* const* out;
    uint32_t w, h, d, c;
    float& at(uint32_t ch, uint32_t z, uint32_t y, uint32_t x) const {
        return out[ch][((size_t)z * h + y) * w + x];
    }
};

// apply_gaussian_2d (seed.rs:320-345); sigma^2 = (r/2)^2, value = amp * exp(-d^2 / (2 sigma^2))
void gaussian2d(const Grid& g, uint32_t ch, uint32_t z, float cx, float cy, float radius, float amp) {
    if (ch >= g.c || z >= g.d) return;
    const float half = radius / 2.0f;
    const float sigma_sq = half * half;
    for (uint32_t y = 0; y < g.h; y++)
        for (uint32_t x = 0; x < g.w; x++) {
            const float dx = (float)x - cx, dy = (float)y - cy;
            const float d2 = dx * dx + dy * dy;
            g.at(ch, z, y, x) += amp * expf(-d2 / (2.0f * sigma_sq));
        }
}

// apply_ring_2d (seed.rs:347-378)
void ring2d(const Grid& g, uint32_t ch, uint32_t z, float cx, float cy, float rin, float rout, float amp) {
    if (ch >= g.c || z >= g.d) return;
    for (uint32_t y = 0; y < g.h; y++)
        for (uint32_t x = 0; x < g.w; x++) {
            const float dx = (float)x - cx, dy = (float)y - cy;
            const float dist = sqrtf(dx * dx + dy * dy);
            if (dist >= rin && dist <= rout) {
                const float edge = (rout - rin) * 0.2f;
                const float f = fminf(fminf((dist - rin) / edge, 1.0f), fminf((rout - dist) / edge, 1.0f));
                g.at(ch, z, y, x) += amp * f;
            }
        }
}

// apply_gaussian_3d (seed.rs:381-410)
void gaussian3d(const Grid& g, uint32_t ch, float cx, float cy, float cz, float radius, float amp) {
    if (ch >= g.c) return;
    const float half = radius / 2.0f;
    const float sigma_sq = half * half;
    for (uint32_t z = 0; z < g.d; z++)
        for (uint32_t y = 0; y < g.h; y++)
            for (uint32_t x = 0; x < g.w; x++) {
                const float dx = (float)x - cx, dy = (float)y - cy, dz = (float)z - cz;
                const float d2 = dx * dx + dy * dy + dz * dz;
                g.at(ch, z, y, x) += amp * expf(-d2 / (2.0f * sigma_sq));
            }
}

// apply_shell_3d (seed.rs:412-446)
void shell3d(const Grid& g, uint32_t ch, float cx, float cy, float cz, float rin, float rout, float amp) {
    if (ch >= g.c) return;
    for (uint32_t z = 0; z < g.d; z++)
        for (uint32_t y = 0; y < g.h; y++)
            for (uint32_t x = 0; x < g.w; x++) {
                const float dx = (float)x - cx, dy = (float)y - cy, dz = (float)z - cz;
                const float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                if (dist >= rin && dist <= rout) {
                    const float edge = (rout - rin) * 0.2f;
                    const float f = fminf(fminf((dist - rin) / edge, 1.0f), fminf((rout - dist) / edge, 1.0f));
                    g.at(ch, z, y, x) += amp * f;
                }
            }
}

// apply_torus_3d (seed.rs:448-488)
void torus3d(const Grid& g, uint32_t ch, float cx, float cy, float cz, float rmaj, float rmin, float amp) {
    if (ch >= g.c) return;
    const float half = rmin / 2.0f;
    const float sigma_sq = half * half;
    for (uint32_t z = 0; z < g.d; z++)
        for (uint32_t y = 0; y < g.h; y++)
            for (uint32_t x = 0; x < g.w; x++) {
                const float dx = (float)x - cx, dy = (float)y - cy, dz = (float)z - cz;
                const float rdx = sqrtf(dx * dx + dy * dy) - rmaj;
                const float dt = sqrtf(rdx * rdx + dz * dz);
                const float v = amp * expf(-dt * dt / (2.0f * sigma_sq));
                if (v > 1e-6f) g.at(ch, z, y, x) += v;
            }
}
}  // namespace

hp_status seed_generate(const hp_seed* seed, uint32_t w, uint32_t h, uint32_t d, uint32_t c, float* const* out) {
    if (!seed || !out) return HP_ERR_NULL;
    if (w == 0 || h == 0 || d == 0) return HP_ERR_INVALID_DIMENSIONS;
    if (c == 0) return HP_ERR_INVALID_CHANNELS;
    for (uint32_t i = 0; i < c; i++)
        if (!out[i]) return HP_ERR_NULL;
    const Grid g{out, w, h, d, c};
    const size_t n = (size_t)w * h * d;
    for (uint32_t i = 0; i < c; i++) memset(out[i], 0, n * sizeof(float));
    const uint32_t mwh = w < h ? w : h;
    const float min_dim = (float)(mwh < d ? mwh : d);
    const uint32_t mid_z = d / 2;
    const uint32_t ch = (uint32_t)seed->channel;
    switch (seed->kind) {
    case HP_PATTERN_GAUSSIAN_BLOB: {
        const float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h;
        const float r = seed->radius * (float)mwh;
        if (d == 1) gaussian2d(g, ch, 0, cx, cy, r, seed->amplitude);
        else gaussian3d(g, ch, cx, cy, 0.5f * (float)d, r, seed->amplitude);
        return HP_OK;
    }
    case HP_PATTERN_MULTI_BLOB:
        if (seed->num_blobs && !seed->blobs) return HP_ERR_NULL;
        for (uint32_t i = 0; i < seed->num_blobs; i++) {
            const hp_blob& b = seed->blobs[i];
            const float cx = b.center[0] * (float)w, cy = b.center[1] * (float)h;
            const float r = b.radius * (float)mwh;
            if (d == 1) gaussian2d(g, b.channel, 0, cx, cy, r, b.amplitude);
            else gaussian3d(g, b.channel, cx, cy, 0.5f * (float)d, r, b.amplitude);
        }
        return HP_OK;
    case HP_PATTERN_RING: {
        const float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h;
        const float rin = seed->inner_radius * (float)mwh, rout = seed->outer_radius * (float)mwh;
        if (d == 1) ring2d(g, ch, 0, cx, cy, rin, rout, seed->amplitude);
        else shell3d(g, ch, cx, cy, 0.5f * (float)d, rin, rout, seed->amplitude);
        return HP_OK;
    }
    case HP_PATTERN_GAUSSIAN_SPHERE:
        gaussian3d(g, ch, seed->center[0] * (float)w, seed->center[1] * (float)h, seed->center[2] * (float)d,
                   seed->radius * min_dim, seed->amplitude);
        return HP_OK;
    case HP_PATTERN_MULTI_SPHERE:
        if (seed->num_blobs && !seed->blobs) return HP_ERR_NULL;
        for (uint32_t i = 0; i < seed->num_blobs; i++) {
            const hp_blob& b = seed->blobs[i];
            gaussian3d(g, b.channel, b.center[0] * (float)w, b.center[1] * (float)h, b.center[2] * (float)d,
                       b.radius * min_dim, b.amplitude);
        }
        return HP_OK;
    case HP_PATTERN_SHELL:
        shell3d(g, ch, seed->center[0] * (float)w, seed->center[1] * (float)h, seed->center[2] * (float)d,
                seed->inner_radius * min_dim, seed->outer_radius * min_dim, seed->amplitude);
        return HP_OK;
    case HP_PATTERN_TORUS3D:
        torus3d(g, ch, seed->center[0] * (float)w, seed->center[1] * (float)h, seed->center[2] * (float)d,
                seed->major_radius * min_dim, seed->minor_radius * min_dim, seed->amplitude);
        return HP_OK;
    case HP_PATTERN_NOISE: {
        // apply_noise (seed.rs:490-524): s = s * 6364136223846793005 + 1; (s >> 33) / 2^31
        uint64_t state = seed->noise_seed;
        const uint32_t c0 = seed->channel >= 0 ? (uint32_t)seed->channel : 0;
        const uint32_t c1 = seed->channel >= 0 ? c0 + 1 : c;
        for (uint32_t k = c0; k < c1 && k < c; k++)
            for (size_t i = 0; i < n; i++) {
                state = state * 6364136223846793005ULL + 1ULL;
                const float r = (float)(state >> 33) / (float)(1ULL << 31);
                out[k][i] += seed->amplitude * r;
            }
        return HP_OK;
    }
    case HP_PATTERN_CUSTOM:
        if (seed->num_values && !seed->values) return HP_ERR_NULL;
        for (uint32_t i = 0; i < seed->num_values; i++) {
            const hp_custom_value& v = seed->values[i];
            if (v.channel < c && v.y < h && v.x < w) g.at(v.channel, mid_z, v.y, v.x) = v.value;
        }
        return HP_OK;
    case HP_PATTERN_CUSTOM3D:
        if (seed->num_values && !seed->values) return HP_ERR_NULL;
        for (uint32_t i = 0; i < seed->num_values; i++) {
            const hp_custom_value& v = seed->values[i];
            if (v.channel < c && v.z < d && v.y < h && v.x < w) g.at(v.channel, v.z, v.y, v.x) = v.value;
        }
        return HP_OK;
    default:
        return HP_ERR_SHAPE;
    }
}

}  // namespace hp
