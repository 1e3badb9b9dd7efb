// host_embedded.cpp -- the embedded-parameter propagator behind the C ABI
// (EmbeddedPropagator, src/compute/embedded_propagator.rs:144-340).
//
// An hp_emb owns a 2D propagator context built on the LDS FFT passes: its state ping-pong holds
// the mass, its FFT buffers compute every kernel's convolution, its boxes serve the reintegration
// when embedding is disabled.  The hp_emb adds, all resident in HBM:
//   P[2]   per-cell parameters, [C][5][N] planes (mu, sigma, weight, beta_a, n), paired with the
//          mass ping-pong: P[cur] is the state's, P[1 - cur] the propagator's next grids (the
//          reference's pre-allocated `next_params`, which keep their content at destinations
//          that receive no mix)
//   U      [C][N] affinity; conv [K][N] convolutions; F [C][2][N] flows; maxbits [C]; err [4]
// A step is: convolutions (3 FFT passes) -> growth with per-cell parameters -> flow -> gather
// with mixing (or, embedding disabled, the general reintegration and a parameter copy).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "hp_common.hpp"
#include "host_internal.hpp"

namespace {

thread_local std::string g_emb_error;

// Work bound of one step's gathers: (2R + 1)^2 source visits per destination and pass over every
// channel; steps beyond it are refused rather than run for minutes.
constexpr double kVisitBound = 2e10;

int radius_bound(long long cells, int passes) {
    const double per = kVisitBound / ((double)cells * passes);
    const double side = std::floor(std::sqrt(std::max(per, 1.0)));
    return std::max(1, (int)((side - 1.0) / 2.0));
}

}  // namespace

struct hp_emb {
    hp_ctx* ctx = nullptr;
    hp::CtxView v;
    hp_config cfg{};
    hp_embedding emb{};
    int C = 0, K = 0;
    long long N = 0;
    float* P[2] = {nullptr, nullptr};
    float* Psnap[2] = {nullptr, nullptr};
    float* U = nullptr;
    float* conv = nullptr;
    float* F = nullptr;
    float* stage = nullptr;  // AoS staging, C*N*5 floats
    unsigned* maxbits = nullptr;
    int* err = nullptr;      // [0] mixing-gather radius bit, [1] general radius, [2] general flags
    int* ktarget = nullptr;
    // stage-entry scratch (allocated on first use)
    float* sA = nullptr;
    float* sP = nullptr;
    float* sPn = nullptr;
    int rmax_mix = 0, rmax_gen = 0;
    std::string msg;

    hp_status fail(hp_status st, const std::string& m) {
        msg = m;
        return st;
    }
};

#define EMB_HIP(e, expr)                                                                     \
    do {                                                                                     \
        hipError_t x_ = (expr);                                                              \
        if (x_ != hipSuccess)                                                                \
            return (e)->fail(HP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(x_)); \
    } while (0)

namespace {

void emb_release(hp_emb* e) {
    if (!e) return;
    void* ptrs[] = {e->P[0], e->P[1], e->Psnap[0], e->Psnap[1], e->U, e->conv, e->F, e->stage, e->maxbits,
                    e->err, e->ktarget, e->sA, e->sP, e->sPn};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    if (e->ctx) hp_destroy(e->ctx);
    delete e;
}

template <typename T>
hp_status emb_alloc(hp_emb* e, T** p, size_t count) {
    EMB_HIP(e, hipMalloc((void**)p, std::max<size_t>(count, 1) * sizeof(T)));
    return HP_OK;
}

// CellParams::default (embedding.rs:35-45) for `cells` cells, AoS.
void default_cells(std::vector<float>& v, size_t cells) {
    v.resize(cells * 5);
    for (size_t i = 0; i < cells; i++) {
        v[5 * i + 0] = 0.15f;
        v[5 * i + 1] = 0.015f;
        v[5 * i + 2] = 1.0f;
        v[5 * i + 3] = 1.0f;
        v[5 * i + 4] = 2.0f;
    }
}

hp_status check_params(hp_emb* e, const hp_cell_params* const* params) {
    if (!params) return e->fail(HP_ERR_NULL, "parameter pointer array is NULL");
    for (int c = 0; c < e->C; c++) {
        if (!params[c]) return e->fail(HP_ERR_NULL, "parameter pointer is NULL");
        const float* p = reinterpret_cast<const float*>(params[c]);
        for (long long i = 0; i < e->N * 5; i++)
            if (!std::isfinite(p[i])) return e->fail(HP_ERR_NONFINITE, "parameters hold NaN/Inf");
    }
    return HP_OK;
}

hp_status check_channels(hp_emb* e, const float* const* ch) {
    if (!ch) return e->fail(HP_ERR_NULL, "channel pointer array is NULL");
    for (int c = 0; c < e->C; c++) {
        if (!ch[c]) return e->fail(HP_ERR_NULL, "channel pointer is NULL");
        for (long long i = 0; i < e->N; i++)
            if (!std::isfinite(ch[c][i])) return e->fail(HP_ERR_NONFINITE, "input state holds NaN/Inf");
    }
    return HP_OK;
}

hp_status check_out(hp_emb* e, const void* const* p) {
    if (!p) return e->fail(HP_ERR_NULL, "output pointer array is NULL");
    for (int c = 0; c < e->C; c++)
        if (!p[c]) return e->fail(HP_ERR_NULL, "output pointer is NULL");
    return HP_OK;
}

hp_status upload_channels(hp_emb* e, const float* const* ch, float* dst) {
    for (int c = 0; c < e->C; c++)
        EMB_HIP(e, hipMemcpyAsync(dst + (size_t)c * e->N, ch[c], e->N * sizeof(float), hipMemcpyHostToDevice,
                                  e->v.stream));
    EMB_HIP(e, hipStreamSynchronize(e->v.stream));
    return HP_OK;
}

hp_status download_channels(hp_emb* e, const float* src, float* const* ch) {
    for (int c = 0; c < e->C; c++)
        EMB_HIP(e, hipMemcpyAsync(ch[c], src + (size_t)c * e->N, e->N * sizeof(float), hipMemcpyDeviceToHost,
                                  e->v.stream));
    EMB_HIP(e, hipStreamSynchronize(e->v.stream));
    return HP_OK;
}

// host AoS grids (one pointer per channel) -> planes
hp_status upload_params(hp_emb* e, const hp_cell_params* const* params, float* planes) {
    const size_t per = (size_t)e->N * 5;
    for (int c = 0; c < e->C; c++)
        EMB_HIP(e, hipMemcpyAsync(e->stage + c * per, params[c], per * sizeof(float), hipMemcpyHostToDevice,
                                  e->v.stream));
    hp::launch_emb_split(e->v.stream, e->stage, planes, e->C, e->N);
    EMB_HIP(e, hipGetLastError());
    EMB_HIP(e, hipStreamSynchronize(e->v.stream));
    return HP_OK;
}

hp_status download_params(hp_emb* e, const float* planes, hp_cell_params* const* params) {
    const size_t per = (size_t)e->N * 5;
    hp::launch_emb_join(e->v.stream, planes, e->stage, e->C, e->N);
    EMB_HIP(e, hipGetLastError());
    for (int c = 0; c < e->C; c++)
        EMB_HIP(e, hipMemcpyAsync(params[c], e->stage + c * per, per * sizeof(float), hipMemcpyDeviceToHost,
                                  e->v.stream));
    EMB_HIP(e, hipStreamSynchronize(e->v.stream));
    return HP_OK;
}

// Stage 1: the convolutions, then the growth with the targets' per-cell parameters.
hp_status enqueue_affinity(hp_emb* e, const float* A, const float* P, float* U) {
    if (e->K > 0) {
        hp_status st = hp::ctx_enqueue_conv(e->ctx, A, e->conv);
        if (st != HP_OK) return e->fail(st, hp::ctx_error(e->ctx));
    }
    hp::launch_emb_growth(e->v.stream, e->conv, P, e->ktarget, U, e->C, e->K, e->N);
    EMB_HIP(e, hipGetLastError());
    return HP_OK;
}

// Stages 2-5 from state (A, P) and affinity U into (Aout, Pout).
hp_status enqueue_stencil(hp_emb* e, const float* A, const float* P, const float* U, float* Aout, float* Pout) {
    hipStream_t st = e->v.stream;
    hp::launch_clear_ints(st, reinterpret_cast<int*>(e->maxbits), e->C);
    hp::launch_emb_flow(st, A, U, P, e->F, e->maxbits, e->C, e->v.g.W, e->v.g.H);
    const float dt = e->cfg.dt, s = e->cfg.flow.distribution_size;
    if (e->emb.enabled) {
        hp::launch_emb_advect(st, A, P, e->F, e->maxbits, Aout, Pout, e->C, e->v.g.W, e->v.g.H, dt, s,
                              e->emb.linear_mixing != 0, e->emb.mixing_temperature, e->rmax_mix, e->err);
    } else {
        // advect_mass_into (reintegration.rs:60-88) with the per-cell flows; parameters copied
        hp::launch_clear_ints(st, e->err + 1, 1);  // this step's radius; the flags in err[2] stay
        for (int c = 0; c < e->C; c++) {
            const long long o = (long long)c * e->N;
            hp::launch_boxes_from_flow(st, A + o, e->F + 2 * o, e->F + 2 * o + e->N, nullptr, e->v.g, dt, s,
                                       e->v.box_lo + o, e->v.box_hi + o, e->err + 1);
        }
        hp::launch_gather_general(st, e->v.box_lo, e->v.box_hi, Aout, e->C, e->v.g, s, 0, false, e->err + 1,
                                  e->rmax_gen);
        EMB_HIP(e, hipMemcpyAsync(Pout, P, (size_t)e->C * 5 * e->N * sizeof(float), hipMemcpyDeviceToDevice, st));
    }
    EMB_HIP(e, hipGetLastError());
    return HP_OK;
}

hp_status enqueue_step(hp_emb* e) {
    const int cur = *e->v.cur;
    hp_status st = enqueue_affinity(e, e->v.A[cur], e->P[cur], e->U);
    if (st != HP_OK) return st;
    st = enqueue_stencil(e, e->v.A[cur], e->P[cur], e->U, e->v.A[1 - cur], e->P[1 - cur]);
    if (st != HP_OK) return st;
    *e->v.cur = 1 - cur;
    *e->v.time += e->cfg.dt;
    *e->v.step += 1;
    return HP_OK;
}

// The flags a run left: 0 when every step ran.
hp_status read_flags(hp_emb* e) {
    int h[3] = {0, 0, 0};
    EMB_HIP(e, hipMemcpyAsync(h, e->err, sizeof(h), hipMemcpyDeviceToHost, e->v.stream));
    EMB_HIP(e, hipStreamSynchronize(e->v.stream));
    if (h[0] & 1)
        return e->fail(HP_ERR_UNSUPPORTED_SIZE, "flow displacement beyond the mixing gather's work bound (search radius "
                                                ">= " + std::to_string(e->rmax_mix) + ")");
    if (h[2] & 2) return e->fail(HP_ERR_NONFINITE, "state or flow holds NaN/Inf");
    if (h[2] & 4)
        return e->fail(HP_ERR_UNSUPPORTED_SIZE, "flow displacement beyond the general gather's work bound (radius > " +
                                                    std::to_string(e->rmax_gen) + ")");
    return HP_OK;
}

// `steps` steps from the resident state; on a refused step the state before the run is restored.
hp_status advance(hp_emb* e, uint64_t steps) {
    if (steps == 0) return HP_OK;
    const size_t abytes = (size_t)e->C * e->N * sizeof(float), pbytes = abytes * 5;
    const int cur0 = *e->v.cur;
    const float t0 = *e->v.time;
    const uint64_t s0 = *e->v.step;
    hipStream_t st = e->v.stream;
    EMB_HIP(e, hipMemcpyAsync(e->v.snapshot, e->v.A[cur0], abytes, hipMemcpyDeviceToDevice, st));
    EMB_HIP(e, hipMemcpyAsync(e->Psnap[0], e->P[0], pbytes, hipMemcpyDeviceToDevice, st));
    EMB_HIP(e, hipMemcpyAsync(e->Psnap[1], e->P[1], pbytes, hipMemcpyDeviceToDevice, st));
    hp::launch_clear_ints(st, e->err, 4);
    for (uint64_t i = 0; i < steps; i++) {
        hp_status r = enqueue_step(e);
        if (r != HP_OK) return r;
    }
    hp_status r = read_flags(e);
    if (r == HP_OK) return HP_OK;
    *e->v.cur = cur0;
    *e->v.time = t0;
    *e->v.step = s0;
    EMB_HIP(e, hipMemcpyAsync(e->v.A[cur0], e->v.snapshot, abytes, hipMemcpyDeviceToDevice, st));
    EMB_HIP(e, hipMemcpyAsync(e->P[0], e->Psnap[0], pbytes, hipMemcpyDeviceToDevice, st));
    EMB_HIP(e, hipMemcpyAsync(e->P[1], e->Psnap[1], pbytes, hipMemcpyDeviceToDevice, st));
    EMB_HIP(e, hipStreamSynchronize(st));
    return r;
}

hp_status stage_scratch(hp_emb* e) {
    if (e->sA) return HP_OK;
    hp_status st = emb_alloc(e, &e->sA, (size_t)e->C * e->N * 2);
    if (st == HP_OK) st = emb_alloc(e, &e->sP, (size_t)e->C * e->N * 5);
    if (st == HP_OK) st = emb_alloc(e, &e->sPn, (size_t)e->C * e->N * 5);
    return st;
}

template <typename F>
hp_status guarded(hp_emb* e, F&& f) {
    if (!e) return HP_ERR_NULL;
    try {
        if (hipSetDevice(e->v.device) != hipSuccess) return e->fail(HP_ERR_HIP, "hipSetDevice failed");
        return f();
    } catch (const std::exception& ex) {
        return e->fail(HP_ERR_HIP, std::string("host exception: ") + ex.what());
    }
}

}  // namespace

extern "C" {

hp_status hp_emb_default_params(const hp_config* cfg, hp_cell_params* out) {
    if (!cfg || !out) return HP_ERR_NULL;
    // EmbeddedState::from_seed (embedded_propagator.rs:50-57): the first kernel's growth
    // parameters (or 0.15 / 0.015 / 1.0 without kernels) and the flow's beta_a and n
    const bool has = cfg->num_kernels > 0 && cfg->kernels;
    out->mu = has ? cfg->kernels[0].mu : 0.15f;
    out->sigma = has ? cfg->kernels[0].sigma : 0.015f;
    out->weight = has ? cfg->kernels[0].weight : 1.0f;
    out->beta_a = cfg->flow.beta_a;
    out->n = cfg->flow.n;
    return HP_OK;
}

hp_status hp_emb_create(const hp_config* cfg, const hp_embedding* emb, int device, hp_emb** out) {
    if (!cfg || !out) return HP_ERR_NULL;
    *out = nullptr;
    // the embedded state is 2D (embedded_propagator.rs:42-43): depth is not read
    hp_config c2 = *cfg;
    c2.depth = 1;
    hp_emb* e = new hp_emb();
    hp_status st = hp::ctx_create_lds(&c2, device, &e->ctx);
    if (st != HP_OK) {
        g_emb_error = hp_last_global_error();
        delete e;
        return st;
    }
    hp::ctx_view(e->ctx, &e->v);
    e->cfg = c2;
    e->emb = emb ? *emb : hp_embedding{0, 1.0f, 0};
    e->C = e->v.C;
    e->K = e->v.K;
    e->N = e->v.g.N;
    const size_t CN = (size_t)e->C * e->N;
#define TRY(x)                          \
    do {                                \
        hp_status s_ = (x);             \
        if (s_ != HP_OK) {              \
            g_emb_error = e->msg;       \
            emb_release(e);             \
            return s_;                  \
        }                               \
    } while (0)
    if (hipSetDevice(device) != hipSuccess) {
        emb_release(e);
        g_emb_error = "hipSetDevice failed";
        return HP_ERR_HIP;
    }
    TRY(emb_alloc(e, &e->P[0], CN * 5));
    TRY(emb_alloc(e, &e->P[1], CN * 5));
    TRY(emb_alloc(e, &e->Psnap[0], CN * 5));
    TRY(emb_alloc(e, &e->Psnap[1], CN * 5));
    TRY(emb_alloc(e, &e->U, CN));
    TRY(emb_alloc(e, &e->conv, (size_t)std::max(1, e->K) * e->N));
    TRY(emb_alloc(e, &e->F, CN * 2));
    TRY(emb_alloc(e, &e->stage, CN * 5));
    TRY(emb_alloc(e, &e->maxbits, (size_t)e->C));
    TRY(emb_alloc(e, &e->err, 4));
    TRY(emb_alloc(e, &e->ktarget, (size_t)std::max(1, e->K)));
    std::vector<int> kt(std::max(1, e->K), 0);
    for (int k = 0; k < e->K; k++) kt[k] = (int)c2.kernels[k].target_channel;
    std::vector<float> dflt;
    default_cells(dflt, CN);
    e->rmax_mix = radius_bound((long long)e->C * e->N, e->emb.linear_mixing ? 2 : 3);
    e->rmax_gen = radius_bound((long long)e->C * e->N, 1);
    auto init = [&]() -> hp_status {
        EMB_HIP(e, hipMemcpy(e->ktarget, kt.data(), kt.size() * sizeof(int), hipMemcpyHostToDevice));
        // both parameter grids start as CellParams::default(), as EmbeddedPropagator::new's
        // next grids do; the state's are set with the state
        EMB_HIP(e, hipMemcpy(e->stage, dflt.data(), dflt.size() * sizeof(float), hipMemcpyHostToDevice));
        hp::launch_emb_split(e->v.stream, e->stage, e->P[0], e->C, e->N);
        hp::launch_emb_split(e->v.stream, e->stage, e->P[1], e->C, e->N);
        EMB_HIP(e, hipMemsetAsync(e->err, 0, 4 * sizeof(int), e->v.stream));
        EMB_HIP(e, hipMemsetAsync(e->U, 0, CN * sizeof(float), e->v.stream));
        EMB_HIP(e, hipStreamSynchronize(e->v.stream));
        return HP_OK;
    };
    TRY(init());
#undef TRY
    *out = e;
    return HP_OK;
}

void hp_emb_destroy(hp_emb* e) {
    if (!e) return;
    (void)hipSetDevice(e->v.device);
    (void)hipStreamSynchronize(e->v.stream);
    emb_release(e);
}

const char* hp_emb_last_error(const hp_emb* e) { return e ? e->msg.c_str() : g_emb_error.c_str(); }

hp_status hp_emb_set_state(hp_emb* e, const float* const* channels, const hp_cell_params* const* params, float time,
                           uint64_t step) {
    return guarded(e, [&]() -> hp_status {
        hp_status st = check_channels(e, channels);
        if (st == HP_OK) st = check_params(e, params);
        if (st != HP_OK) return st;
        const int cur = *e->v.cur;
        st = upload_channels(e, channels, e->v.A[cur]);
        if (st == HP_OK) st = upload_params(e, params, e->P[cur]);
        if (st != HP_OK) return st;
        *e->v.time = time;
        *e->v.step = step;
        return HP_OK;
    });
}

hp_status hp_emb_get_state(hp_emb* e, float* const* channels, hp_cell_params* const* params, float* time,
                           uint64_t* step) {
    return guarded(e, [&]() -> hp_status {
        hp_status st = check_out(e, (const void* const*)channels);
        if (st == HP_OK && params) st = check_out(e, (const void* const*)params);
        if (st != HP_OK) return st;
        const int cur = *e->v.cur;
        st = download_channels(e, e->v.A[cur], channels);
        if (st == HP_OK && params) st = download_params(e, e->P[cur], params);
        if (st != HP_OK) return st;
        if (time) *time = *e->v.time;
        if (step) *step = *e->v.step;
        return HP_OK;
    });
}

hp_status hp_emb_advance(hp_emb* e, uint64_t steps) {
    return guarded(e, [&]() -> hp_status { return advance(e, steps); });
}

hp_status hp_emb_run(hp_emb* e, float* const* channels, hp_cell_params* const* params, float* time, uint64_t* step,
                     uint64_t steps) {
    hp_status st = hp_emb_set_state(e, channels, params, time ? *time : 0.0f, step ? *step : 0);
    if (st != HP_OK) return st;
    st = hp_emb_advance(e, steps);
    if (st != HP_OK) return st;
    return hp_emb_get_state(e, channels, params, time, step);
}

hp_status hp_emb_step(hp_emb* e, float* const* channels, hp_cell_params* const* params, float* time,
                      uint64_t* step) {
    return hp_emb_run(e, channels, params, time, step, 1);
}

hp_status hp_emb_mass_f64(hp_emb* e, double* per_channel, double* total) {
    if (!e) return HP_ERR_NULL;
    const hp_status st = hp_mass_f64(e->ctx, per_channel, total);
    if (st != HP_OK) e->msg = hp_last_error(e->ctx);
    return st;
}

hp_status hp_emb_get_next_params(hp_emb* e, hp_cell_params* const* params) {
    return guarded(e, [&]() -> hp_status {
        hp_status st = check_out(e, (const void* const*)params);
        if (st != HP_OK) return st;
        return download_params(e, e->P[1 - *e->v.cur], params);
    });
}

hp_status hp_emb_set_next_params(hp_emb* e, const hp_cell_params* const* params) {
    return guarded(e, [&]() -> hp_status {
        hp_status st = check_params(e, params);
        if (st != HP_OK) return st;
        return upload_params(e, params, e->P[1 - *e->v.cur]);
    });
}

hp_status hp_emb_last_radius(hp_emb* e, int32_t* radius) {
    return guarded(e, [&]() -> hp_status {
        if (!radius) return e->fail(HP_ERR_NULL, "radius is NULL");
        std::vector<unsigned> bits(e->C);
        EMB_HIP(e, hipMemcpyAsync(bits.data(), e->maxbits, bits.size() * sizeof(unsigned), hipMemcpyDeviceToHost,
                                  e->v.stream));
        EMB_HIP(e, hipStreamSynchronize(e->v.stream));
        for (int c = 0; c < e->C; c++) {
            float m;
            memcpy(&m, &bits[c], sizeof(float));
            // param_advection.rs:99-100: ceil(max |F| dt + s) as i32 + 1 (f32, saturating cast)
            const float v = ceilf(m * e->cfg.dt + e->cfg.flow.distribution_size);
            const int32_t r = v != v ? 0 : v >= 2147483647.0f ? 2147483647 : v <= -2147483648.0f ? INT32_MIN : (int32_t)v;
            radius[c] = e->emb.enabled ? (r == 2147483647 ? r : r + 1) : 0;
        }
        return HP_OK;
    });
}

hp_status hp_emb_affinity(hp_emb* e, const float* const* channels, const hp_cell_params* const* params,
                          float* const* affinity) {
    return guarded(e, [&]() -> hp_status {
        hp_status st = check_channels(e, channels);
        if (st == HP_OK) st = check_params(e, params);
        if (st == HP_OK) st = check_out(e, (const void* const*)affinity);
        if (st == HP_OK) st = stage_scratch(e);
        if (st == HP_OK) st = upload_channels(e, channels, e->sA);
        if (st == HP_OK) st = upload_params(e, params, e->sP);
        if (st == HP_OK) st = enqueue_affinity(e, e->sA, e->sP, e->U);
        if (st == HP_OK) st = download_channels(e, e->U, affinity);
        return st;
    });
}

hp_status hp_emb_stencil(hp_emb* e, const float* const* channels, const hp_cell_params* const* params,
                         const float* const* affinity, float* const* next, hp_cell_params* const* next_params) {
    return guarded(e, [&]() -> hp_status {
        hp_status st = check_channels(e, channels);
        if (st == HP_OK) st = check_params(e, params);
        if (st == HP_OK) st = check_channels(e, affinity);
        if (st == HP_OK) st = check_out(e, (const void* const*)next);
        if (st == HP_OK) st = check_params(e, next_params);
        if (st == HP_OK) st = stage_scratch(e);
        if (st == HP_OK) st = upload_channels(e, channels, e->sA);
        if (st == HP_OK) st = upload_channels(e, affinity, e->U);
        if (st == HP_OK) st = upload_params(e, params, e->sP);
        if (st == HP_OK) st = upload_params(e, next_params, e->sPn);
        if (st != HP_OK) return st;
        hp::launch_clear_ints(e->v.stream, e->err, 4);
        float* out = e->sA + (size_t)e->C * e->N;
        st = enqueue_stencil(e, e->sA, e->sP, e->U, out, e->sPn);
        if (st == HP_OK) st = read_flags(e);
        if (st == HP_OK) st = download_channels(e, out, next);
        if (st == HP_OK) st = download_params(e, e->sPn, next_params);
        return st;
    });
}

// advect_mass_and_params_into (param_advection.rs:83-166) on host buffers, one channel.
hp_status hp_advect_mass_and_params(const float* mass, const hp_cell_params* params, const float* fx,
                                    const float* fy, const hp_embedding* emb, float dt, float s, uint32_t w,
                                    uint32_t h, float* next_mass, hp_cell_params* next_params, int device) {
    if (!mass || !params || !fx || !fy || !next_mass || !next_params) return HP_ERR_NULL;
    if (w == 0 || h == 0) return HP_ERR_INVALID_DIMENSIONS;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        g_emb_error = "no HIP device";
        return HP_ERR_NO_DEVICE;
    }
    if (hipSetDevice(device) != hipSuccess) return HP_ERR_HIP;
    const long long N = (long long)w * h;
    const hp_embedding em = emb ? *emb : hp_embedding{1, 1.0f, 0};
    float *dA = nullptr, *dF = nullptr, *dP = nullptr, *dPn = nullptr, *dAo = nullptr, *dS = nullptr;
    unsigned* dmax = nullptr;
    int* derr = nullptr;
    hp_status result = HP_OK;
    auto ok = [&](hipError_t x, const char* what) {
        if (x != hipSuccess && result == HP_OK) {
            result = HP_ERR_HIP;
            g_emb_error = std::string(what) + ": " + hipGetErrorString(x);
        }
        return result == HP_OK;
    };
    if (ok(hipMalloc((void**)&dA, N * 4), "hipMalloc") && ok(hipMalloc((void**)&dF, 2 * N * 4), "hipMalloc") &&
        ok(hipMalloc((void**)&dP, 5 * N * 4), "hipMalloc") && ok(hipMalloc((void**)&dPn, 5 * N * 4), "hipMalloc") &&
        ok(hipMalloc((void**)&dAo, N * 4), "hipMalloc") && ok(hipMalloc((void**)&dS, 5 * N * 4), "hipMalloc") &&
        ok(hipMalloc((void**)&dmax, 4), "hipMalloc") && ok(hipMalloc((void**)&derr, 16), "hipMalloc")) {
        const int rmax = radius_bound(N, em.linear_mixing ? 2 : 3);
        if (ok(hipMemcpy(dA, mass, N * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
            ok(hipMemcpy(dF, fx, N * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
            ok(hipMemcpy(dF + N, fy, N * 4, hipMemcpyHostToDevice), "hipMemcpy") &&
            ok(hipMemcpy(dS, params, 5 * N * 4, hipMemcpyHostToDevice), "hipMemcpy")) {
            hp::launch_emb_split(nullptr, dS, dP, 1, N);
            ok(hipMemcpy(dS, next_params, 5 * N * 4, hipMemcpyHostToDevice), "hipMemcpy");
            hp::launch_emb_split(nullptr, dS, dPn, 1, N);
            ok(hipMemset(dmax, 0, 4), "hipMemset");
            ok(hipMemset(derr, 0, 16), "hipMemset");
            hp::launch_emb_maxmag(nullptr, dF, dmax, 1, N);
            hp::launch_emb_advect(nullptr, dA, dP, dF, dmax, dAo, dPn, 1, (int)w, (int)h, dt, s, em.linear_mixing != 0,
                                  em.mixing_temperature, rmax, derr);
            ok(hipGetLastError(), "launch");
            int herr = 0;
            if (ok(hipMemcpy(&herr, derr, 4, hipMemcpyDeviceToHost), "hipMemcpy") && (herr & 1)) {
                result = HP_ERR_UNSUPPORTED_SIZE;
                g_emb_error = "flow displacement beyond the mixing gather's work bound";
            }
            if (result == HP_OK) {
                hp::launch_emb_join(nullptr, dPn, dS, 1, N);
                ok(hipMemcpy(next_mass, dAo, N * 4, hipMemcpyDeviceToHost), "hipMemcpy");
                ok(hipMemcpy(next_params, dS, 5 * N * 4, hipMemcpyDeviceToHost), "hipMemcpy");
            }
        }
    }
    void* ptrs[] = {dA, dF, dP, dPn, dAo, dS, dmax, derr};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    return result;
}

}  // extern "C"
