/*
 * flowlenia_oracle.c -- CPU oracle for the Flow Lenia time step.
 * TEST INFRASTRUCTURE ONLY: see flowlenia_oracle.h for what it restates and why.
 * Build with -ffp-contract=off and without -ffast-math (oracle/Makefile) so the
 * float path rounds exactly as the reference's Rust f32 code does.
 */
#define _GNU_SOURCE
#include "flowlenia_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ---- small helpers shared by both instantiations ----------------------------- */

/* wrap_coord (reintegration.rs:171-174). */
int orc_wrap_coord(int32_t coord, uint32_t size) {
    int32_t s = (int32_t)size;
    return (int)((uint32_t)((coord % s) + s) % size);
}

/* Rust `f as i32`: saturating, NaN -> 0. */
int32_t orc_sat_i32_f32(float v) {
    if (v != v) return 0;
    if (v >= 2147483647.0f) return INT32_MAX;
    if (v <= -2147483648.0f) return INT32_MIN;
    return (int32_t)v;
}

int32_t orc_sat_i32_f64(double v) {
    if (v != v) return 0;
    if (v >= 2147483647.0) return INT32_MAX;
    if (v <= -2147483648.0) return INT32_MIN;
    return (int32_t)v;
}

/* A minimal parallel-for over [0, n) with `threads` pthreads (the reference fans
 * out with rayon over kernels, propagator.rs:259-267, and channels, :332-371). */
typedef void (*orc_task_fn)(void* arg, uint32_t i);
typedef struct {
    orc_task_fn fn;
    void* arg;
    uint32_t n;
    uint32_t stride;
    uint32_t first;
} orc_worker;

static void* orc_worker_main(void* a) {
    orc_worker* w = (orc_worker*)a;
    for (uint32_t i = w->first; i < w->n; i += w->stride) w->fn(w->arg, i);
    return NULL;
}

static void orc_parallel_for(int threads, uint32_t n, orc_task_fn fn, void* arg) {
    if (threads <= 1 || n <= 1) {
        for (uint32_t i = 0; i < n; i++) fn(arg, i);
        return;
    }
    uint32_t t = (uint32_t)threads < n ? (uint32_t)threads : n;
    pthread_t tid[256];
    orc_worker ws[256];
    if (t > 256) t = 256;
    for (uint32_t i = 0; i < t; i++) {
        ws[i].fn = fn; ws[i].arg = arg; ws[i].n = n; ws[i].stride = t; ws[i].first = i;
        if (i > 0) pthread_create(&tid[i], NULL, orc_worker_main, &ws[i]);
    }
    orc_worker_main(&ws[0]);
    for (uint32_t i = 1; i < t; i++) pthread_join(tid[i], NULL);
}

/* ---- float instantiation ------------------------------------------------------ */
#define T float
#define S f32
#define EXP expf
#define SQRT sqrtf
#define FLOOR floorf
#define CEIL ceilf
#define ROUND roundf
#define FABS fabsf
#define POW powf
#define SAT_I32 orc_sat_i32_f32
#include "oracle_impl.inc"
#undef T
#undef S
#undef EXP
#undef SQRT
#undef FLOOR
#undef CEIL
#undef ROUND
#undef FABS
#undef POW
#undef SAT_I32

/* ---- double instantiation ----------------------------------------------------- */
#define T double
#define S f64
#define EXP exp
#define SQRT sqrt
#define FLOOR floor
#define CEIL ceil
#define ROUND round
#define FABS fabs
#define POW pow
#define SAT_I32 orc_sat_i32_f64
#include "oracle_impl.inc"
#undef T
#undef S

/* ---- config validation (config.rs:136-164) ------------------------------------ */
int orc_validate(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel) {
    if (cfg->width == 0 || cfg->height == 0 || cfg->depth == 0) return HP_ERR_INVALID_DIMENSIONS;
    if (cfg->channels == 0) return HP_ERR_INVALID_CHANNELS;
    if (!(cfg->dt > 0.0f)) {
        /* Rust `self.dt <= 0.0` is false for NaN, so NaN passes validation */
        if (cfg->dt <= 0.0f) return HP_ERR_INVALID_TIME_STEP;
    }
    if (cfg->kernel_radius == 0) return HP_ERR_INVALID_KERNEL_RADIUS;
    for (uint32_t i = 0; i < cfg->num_kernels; i++) {
        if (cfg->kernels[i].source_channel >= cfg->channels) {
            if (bad_kernel) *bad_kernel = i;
            if (bad_channel) *bad_channel = cfg->kernels[i].source_channel;
            return HP_ERR_INVALID_CHANNEL_INDEX;
        }
        if (cfg->kernels[i].target_channel >= cfg->channels) {
            if (bad_kernel) *bad_kernel = i;
            if (bad_channel) *bad_channel = cfg->kernels[i].target_channel;
            return HP_ERR_INVALID_CHANNEL_INDEX;
        }
    }
    return 0;
}

/* ---- seeds (seed.rs:160-524), f32 ---------------------------------------------- */
typedef struct {
    float* const* out;
    uint32_t w, h, d, c;
} orc_grid;

#define GIDX(g, z, y, x) (((size_t)(z) * (g)->h + (y)) * (g)->w + (x))

/* apply_gaussian_2d (seed.rs:320-345) */
static void orc_gauss2d(orc_grid* g, uint32_t ch, uint32_t z, float cx, float cy, float radius, float amp) {
    if (ch >= g->c || z >= g->d) return;
    float hr = radius / 2.0f;
    float sigma_sq = hr * hr;
    for (uint32_t y = 0; y < g->h; y++)
        for (uint32_t x = 0; x < g->w; x++) {
            float dx = (float)x - cx, dy = (float)y - cy;
            float dist_sq = dx * dx + dy * dy;
            float value = amp * expf(-dist_sq / (2.0f * sigma_sq));
            g->out[ch][GIDX(g, z, y, x)] += value;
        }
}

/* apply_ring_2d (seed.rs:347-378) */
static void orc_ring2d(orc_grid* g, uint32_t ch, uint32_t z, float cx, float cy, float rin, float rout, float amp) {
    if (ch >= g->c || z >= g->d) return;
    for (uint32_t y = 0; y < g->h; y++)
        for (uint32_t x = 0; x < g->w; x++) {
            float dx = (float)x - cx, dy = (float)y - cy;
            float dist = sqrtf(dx * dx + dy * dy);
            if (dist >= rin && dist <= rout) {
                float edge = (rout - rin) * 0.2f;
                float fi = fminf((dist - rin) / edge, 1.0f);
                float fo = fminf((rout - dist) / edge, 1.0f);
                float f = fminf(fi, fo);
                g->out[ch][GIDX(g, z, y, x)] += amp * f;
            }
        }
}

/* apply_gaussian_3d (seed.rs:381-410) */
static void orc_gauss3d(orc_grid* g, uint32_t ch, float cx, float cy, float cz, float radius, float amp) {
    if (ch >= g->c) return;
    float hr = radius / 2.0f;
    float sigma_sq = hr * hr;
    for (uint32_t z = 0; z < g->d; z++)
        for (uint32_t y = 0; y < g->h; y++)
            for (uint32_t x = 0; x < g->w; x++) {
                float dx = (float)x - cx, dy = (float)y - cy, dz = (float)z - cz;
                float dist_sq = dx * dx + dy * dy + dz * dz;
                float value = amp * expf(-dist_sq / (2.0f * sigma_sq));
                g->out[ch][GIDX(g, z, y, x)] += value;
            }
}

/* apply_shell_3d (seed.rs:412-446) */
static void orc_shell3d(orc_grid* g, uint32_t ch, float cx, float cy, float cz, float rin, float rout, float amp) {
    if (ch >= g->c) return;
    for (uint32_t z = 0; z < g->d; z++)
        for (uint32_t y = 0; y < g->h; y++)
            for (uint32_t x = 0; x < g->w; x++) {
                float dx = (float)x - cx, dy = (float)y - cy, dz = (float)z - cz;
                float dist = sqrtf(dx * dx + dy * dy + dz * dz);
                if (dist >= rin && dist <= rout) {
                    float edge = (rout - rin) * 0.2f;
                    float fi = fminf((dist - rin) / edge, 1.0f);
                    float fo = fminf((rout - dist) / edge, 1.0f);
                    float f = fminf(fi, fo);
                    g->out[ch][GIDX(g, z, y, x)] += amp * f;
                }
            }
}

/* apply_torus_3d (seed.rs:448-488) */
static void orc_torus3d(orc_grid* g, uint32_t ch, float cx, float cy, float cz, float rmaj, float rmin, float amp) {
    if (ch >= g->c) return;
    float hr = rmin / 2.0f;
    float sigma_sq = hr * hr;
    for (uint32_t z = 0; z < g->d; z++)
        for (uint32_t y = 0; y < g->h; y++)
            for (uint32_t x = 0; x < g->w; x++) {
                float dx = (float)x - cx, dy = (float)y - cy, dz = (float)z - cz;
                float dist_xy = sqrtf(dx * dx + dy * dy);
                float ring_dx = dist_xy - rmaj;
                float dtube = sqrtf(ring_dx * ring_dx + dz * dz);
                float value = amp * expf(-dtube * dtube / (2.0f * sigma_sq));
                if (value > 1e-6f) g->out[ch][GIDX(g, z, y, x)] += value;
            }
}

int orc_seed_generate(const hp_seed* seed, uint32_t w, uint32_t h, uint32_t d, uint32_t c, float* const* out) {
    orc_grid g = {out, w, h, d, c};
    size_t n = (size_t)w * h * d;
    for (uint32_t i = 0; i < c; i++) memset(out[i], 0, sizeof(float) * n);
    uint32_t mwh = w < h ? w : h;
    float min_dim = (float)(mwh < d ? mwh : d);
    uint32_t mid_z = d / 2;
    switch (seed->kind) {
    case HP_PATTERN_GAUSSIAN_BLOB: {
        float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h;
        float r = seed->radius * (float)mwh;
        if (d == 1) orc_gauss2d(&g, (uint32_t)seed->channel, 0, cx, cy, r, seed->amplitude);
        else orc_gauss3d(&g, (uint32_t)seed->channel, cx, cy, 0.5f * (float)d, r, seed->amplitude);
        break;
    }
    case HP_PATTERN_MULTI_BLOB:
        for (uint32_t i = 0; i < seed->num_blobs; i++) {
            const hp_blob* b = &seed->blobs[i];
            float cx = b->center[0] * (float)w, cy = b->center[1] * (float)h;
            float r = b->radius * (float)mwh;
            if (d == 1) orc_gauss2d(&g, b->channel, 0, cx, cy, r, b->amplitude);
            else orc_gauss3d(&g, b->channel, cx, cy, 0.5f * (float)d, r, b->amplitude);
        }
        break;
    case HP_PATTERN_RING: {
        float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h;
        float rin = seed->inner_radius * (float)mwh, rout = seed->outer_radius * (float)mwh;
        if (d == 1) orc_ring2d(&g, (uint32_t)seed->channel, 0, cx, cy, rin, rout, seed->amplitude);
        else orc_shell3d(&g, (uint32_t)seed->channel, cx, cy, 0.5f * (float)d, rin, rout, seed->amplitude);
        break;
    }
    case HP_PATTERN_GAUSSIAN_SPHERE: {
        float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h, cz = seed->center[2] * (float)d;
        orc_gauss3d(&g, (uint32_t)seed->channel, cx, cy, cz, seed->radius * min_dim, seed->amplitude);
        break;
    }
    case HP_PATTERN_MULTI_SPHERE:
        for (uint32_t i = 0; i < seed->num_blobs; i++) {
            const hp_blob* b = &seed->blobs[i];
            orc_gauss3d(&g, b->channel, b->center[0] * (float)w, b->center[1] * (float)h, b->center[2] * (float)d,
                        b->radius * min_dim, b->amplitude);
        }
        break;
    case HP_PATTERN_SHELL: {
        float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h, cz = seed->center[2] * (float)d;
        orc_shell3d(&g, (uint32_t)seed->channel, cx, cy, cz, seed->inner_radius * min_dim,
                    seed->outer_radius * min_dim, seed->amplitude);
        break;
    }
    case HP_PATTERN_TORUS3D: {
        float cx = seed->center[0] * (float)w, cy = seed->center[1] * (float)h, cz = seed->center[2] * (float)d;
        orc_torus3d(&g, (uint32_t)seed->channel, cx, cy, cz, seed->major_radius * min_dim,
                    seed->minor_radius * min_dim, seed->amplitude);
        break;
    }
    case HP_PATTERN_NOISE: {
        /* apply_noise (seed.rs:490-524): LCG, (state >> 33) / 2^31 */
        uint64_t state = seed->noise_seed;
        uint32_t c0 = 0, c1 = c;
        if (seed->channel >= 0) { c0 = (uint32_t)seed->channel; c1 = c0 + 1; }
        for (uint32_t ch = c0; ch < c1; ch++) {
            if (ch >= c) continue;
            for (size_t i = 0; i < n; i++) {
                state = state * 6364136223846793005ULL + 1ULL;
                float r = (float)(state >> 33) / (float)(1ULL << 31);
                out[ch][i] += seed->amplitude * r;
            }
        }
        break;
    }
    case HP_PATTERN_CUSTOM:
        for (uint32_t i = 0; i < seed->num_values; i++) {
            const hp_custom_value* v = &seed->values[i];
            if (v->channel < c && v->y < h && v->x < w) out[v->channel][GIDX(&g, mid_z, v->y, v->x)] = v->value;
        }
        break;
    case HP_PATTERN_CUSTOM3D:
        for (uint32_t i = 0; i < seed->num_values; i++) {
            const hp_custom_value* v = &seed->values[i];
            if (v->channel < c && v->z < d && v->y < h && v->x < w)
                out[v->channel][GIDX(&g, v->z, v->y, v->x)] = v->value;
        }
        break;
    default:
        return HP_ERR_SHAPE;
    }
    return 0;
}

float orc_total_mass_f32(const float* const* channels, uint32_t c, uint64_t n) {
    float total = 0.0f;
    for (uint32_t i = 0; i < c; i++) {
        float s = 0.0f;
        for (uint64_t j = 0; j < n; j++) s += channels[i][j];
        total += s;
    }
    return total;
}

void orc_center_of_mass(const float* grid, uint32_t w, uint32_t h, float* cx, float* cy) {
    float total = 0.0f, sx = 0.0f, sy = 0.0f;
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            float m = grid[(size_t)y * w + x];
            if (m > 0.0f) {
                total += m;
                sx += (float)x * m;
                sy += (float)y * m;
            }
        }
    if (total > 1e-6f) {
        *cx = sx / total;
        *cy = sy / total;
    } else {
        *cx = (float)w / 2.0f;
        *cy = (float)h / 2.0f;
    }
}

float orc_locomotion(float x0, float y0, float x1, float y1, uint32_t w, uint32_t h) {
    float ddx = x1 - x0, ddy = y1 - y0;
    float disp = sqrtf(ddx * ddx + ddy * ddy);
    float maxd = sqrtf((float)((uint64_t)w * w + (uint64_t)h * h));
    float v = disp / maxd;
    if (v < 0.0f) v = 0.0f;
    if (v > 1.0f) v = 1.0f;
    return v;
}
