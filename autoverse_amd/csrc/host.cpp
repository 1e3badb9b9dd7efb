// host.cpp -- the HipPropagator context behind the C ABI of include/flowlenia_hip.h.
//
// One context = one simulation on one GPU (CpuPropagator::new, propagator.rs:157-236).
// Device memory, all resident in HBM for the life of the context:
//   A[2]   state ping-pong, [C][D*H*W] f32 (the reference's state.channels / next_channels)
//   U      affinity [C][N] f32 (the reference's `affinity`)
//   X      half spectra of the distinct source channels, [nsrc][D*H][Wp] complex64
//   Y      per-kernel filtered half spectra, [K][D*H][Wp] complex64
//   Kq     folded real kernel spectra, [K][D/2+1][H/2+1][W/2+1] f32
//   box    general-path source boxes, 2 x [C][N] float4 (only touched on the rare path)
// A step is: rows_fwd -> (3D: cols_fwd) -> spectral -> (3D: cols_inv) -> rows_inv ->
// stencil; steps replay from two captured hipGraphs (A0->A1 and A1->A0).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <mutex>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "hp_common.hpp"
#include "host_internal.hpp"

namespace {

thread_local std::string g_global_error;

struct Launch {
    int cls;
    hipEvent_t a, b;
    int n;  // launches inside the bracket
};

}  // namespace

namespace hp {
void set_global_error(const std::string& msg) { g_global_error = msg; }
}  // namespace hp

struct hp_ctx {
    hp_config cfg{};
    std::vector<hp_kernel> kernels;
    std::vector<hp_ring> rings;
    int device = 0;
    hipStream_t stream = nullptr;
    hp::Geo g{};
    int C = 0, K = 0;
    bool is3d = false;
    bool wave = false;  // 2D sides in {256..2048}: column-major spectra, wave-level FFTs
    std::vector<int> src_chan;  // source slot -> channel
    std::vector<hp::SpecWork> spec_work;  // host copy of the spectral work list
    std::vector<int> tgt_chan;  // target work item -> channel
    float* A[2] = {nullptr, nullptr};
    float* U = nullptr;
    float* h_stage = nullptr;  // pinned host staging of one state (C*N floats), made on first use
    float2* X = nullptr;
    float2* Y = nullptr;
    float* Kq = nullptr;
    std::map<int, float2*> tw;   // FFT length -> twiddle table
    std::map<int, float2*> tws;  // wave path: line length L -> per-stage tables of the L/2-point FFTs
    int* d_src_chan = nullptr;
    hp::SpecWork* d_spec_work = nullptr;
    int n_spec_work = 0;
    int* d_klist_spec = nullptr;
    hp::TgtWork* d_tgt_work = nullptr;
    int n_tgt_work = 0;
    int* d_klist_tgt = nullptr;
    hp::GrowthParams* d_gp = nullptr;
    // sides the FFT passes do not cover: the direct convolution's kernels, taps, column ranges
    bool direct = false;
    int direct_rmax = 0;
    float* d_conv = nullptr;  // [K][N] convolutions of the direct path
    hp::DirectKernel* d_dk = nullptr;
    float* d_taps = nullptr;
    int2* d_cols = nullptr;
    int* d_flags = nullptr;  // [0] stencil flag bits, [1] needed radius, [2] non-finite bits
    float4* box_lo = nullptr;
    float4* box_hi = nullptr;
    double* d_partials = nullptr;
    double* d_mass = nullptr;
    int mass_parts = 0;
    float* snapshot = nullptr;
    int cur = 0;
    float time = 0.0f;
    uint64_t step = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};       // one step from parity b
    hipGraphExec_t graph_run[2] = {nullptr, nullptr};   // kGraphSteps steps from parity b
    bool profiling = false;
    int prof_reps = 1;  // back-to-back repeats of an out-of-place pass inside its event bracket
    std::vector<Launch> launches;
    std::vector<hipEvent_t> event_pool;
    double prof_ms[HP_K_NUM_CLASSES] = {};
    uint64_t prof_n[HP_K_NUM_CLASSES] = {};
    uint32_t max_radius = 0;
    uint64_t general_steps = 0;
    hp_plan plan{};        // resolved execution plan
    hp::FftVariant fv;     // LDS FFT variant of the plan
    int tile = 0;          // fused stencil tile shape of the plan
    bool exact = false;    // exact-conservation reintegration (HP_CONSERVATION_EXACT)
    std::string err;

    hp_status fail(hp_status st, const std::string& msg) {
        err = msg;
        return st;
    }
};

#define HIP_TRY(ctx, expr)                                                                   \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return (ctx)->fail(HP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

hipEvent_t take_event(hp_ctx* c) {
    if (!c->event_pool.empty()) {
        hipEvent_t e = c->event_pool.back();
        c->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

// Brackets one kernel launch with events when profiling is on.
struct Timed {
    hp_ctx* c;
    int cls;
    hipEvent_t a = nullptr;
    Timed(hp_ctx* ctx, int k) : c(ctx), cls(k) {
        if (c->profiling) {
            a = take_event(c);
            (void)hipEventRecord(a, c->stream);
        }
    }
    ~Timed() {
        if (c->profiling) {
            hipEvent_t b = take_event(c);
            (void)hipEventRecord(b, c->stream);
            c->launches.push_back({cls, a, b, 1});
        }
    }
};

// Profiling mode brackets a pass with events.  A pass that writes only buffers its inputs do
// not alias (the wave FFT passes and the fused stencils) is launched prof_reps times back to
// back inside the bracket: each repeat recomputes the same output, and the event overhead is
// spread over the repeats.  Outside profiling the pass is launched once.
template <class F>
void timed_pass(hp_ctx* c, int cls, bool repeatable, F&& launch) {
    if (!c->profiling) {
        launch();
        return;
    }
    const int n = repeatable ? std::max(1, c->prof_reps) : 1;
    hipEvent_t a = take_event(c), b = take_event(c);
    (void)hipEventRecord(a, c->stream);
    for (int i = 0; i < n; i++) launch();
    (void)hipEventRecord(b, c->stream);
    c->launches.push_back({cls, a, b, n});
}

void collect_profile(hp_ctx* c) {
    for (auto& l : c->launches) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, l.a, l.b) == hipSuccess) {
            c->prof_ms[l.cls] += ms;
            c->prof_n[l.cls] += (uint64_t)l.n;
        }
        c->event_pool.push_back(l.a);
        c->event_pool.push_back(l.b);
    }
    c->launches.clear();
}

template <typename T>
hp_status dalloc(hp_ctx* c, T** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)p, count * sizeof(T));
    if (e != hipSuccess) return c->fail(HP_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    return HP_OK;
}

template <typename T>
hp_status upload(hp_ctx* c, T* dst, const std::vector<T>& src) {
    if (src.empty()) return HP_OK;
    HIP_TRY(c, hipMemcpy(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
    return HP_OK;
}

hp_status make_twiddles(hp_ctx* c, int L) {
    if (c->tw.count(L)) return HP_OK;
    std::vector<float2> t(L);
    for (int m = 0; m < L; m++) {
        const double a = -2.0 * M_PI * (double)m / (double)L;
        t[m] = make_float2((float)cos(a), (float)sin(a));
    }
    float2* d = nullptr;
    hp_status st = dalloc(c, &d, L);
    if (st != HP_OK) return st;
    c->tw[L] = d;
    return upload(c, d, t);
}

// The wave path's per-stage twiddle tables for the L/2-point transforms of a length-L line,
// each entry the length-L table's value, computed by the same expression.
hp_status make_stage_twiddles(hp_ctx* c, int L) {
    if (c->tws.count(L)) return HP_OK;
    const std::vector<int> idx = hp::wave_stage_twiddle_index(L / 2);
    std::vector<float2> t(idx.size());
    for (size_t i = 0; i < idx.size(); i++) {
        const int m = 2 * idx[i];
        const double a = -2.0 * M_PI * (double)m / (double)L;
        t[i] = make_float2((float)cos(a), (float)sin(a));
    }
    float2* d = nullptr;
    hp_status st = dalloc(c, &d, t.size());
    if (st != HP_OK) return st;
    c->tws[L] = d;
    return upload(c, d, t);
}

// ---- one step, enqueued on c->stream: A[src] -> A[1-src]; the FFT path leaves U filled ----
hp_status enqueue_affinity(hp_ctx* c, const float* A) {
    if (c->K == 0) return HP_OK;  // U stays zero (propagator.rs:245-247)
    const hp::Geo& g = c->g;
    if (c->direct) {
        Timed t(c, HP_K_ROWS_INV);
        if (!hp::launch_direct_affinity(c->stream, A, c->U, c->d_conv, c->d_dk, c->d_taps, c->d_cols, c->K, c->C, g,
                                        c->direct_rmax))
            return c->fail(HP_ERR_UNSUPPORTED_SIZE, "kernel radius beyond the direct convolution's LDS tile");
        HIP_TRY(c, hipGetLastError());
        return HP_OK;
    }
    if (c->wave) {  // A -> X -> Y -> U, each pass out of place
        timed_pass(c, HP_K_ROWS_FWD, true, [&] {
            hp::launch_rows_fwd_t(c->stream, A, c->X, c->d_src_chan, (int)c->src_chan.size(), g, c->tw[g.W], c->tws[g.W]);
        });
        timed_pass(c, HP_K_SPECTRAL, true, [&] {
            hp::launch_spectral_t(c->stream, c->X, c->Y, c->Kq, c->d_spec_work, c->n_spec_work, c->d_klist_spec, g,
                                  c->tw[g.H], c->tws[g.H]);
        });
        timed_pass(c, HP_K_ROWS_INV, true, [&] {
            hp::launch_rows_inv_t(c->stream, c->Y, c->U, c->d_tgt_work, c->n_tgt_work, c->d_klist_tgt, c->d_gp, g,
                                  c->tw[g.W], c->tws[g.W]);
        });
        HIP_TRY(c, hipGetLastError());
        return HP_OK;
    }
    {
        Timed t(c, HP_K_ROWS_FWD);
        hp::launch_rows_fwd(g.W, c->stream, A, c->X, c->d_src_chan, (int)c->src_chan.size(), g, c->tw[g.W], c->fv);
    }
    if (c->is3d) {
        Timed t(c, HP_K_COLS_FWD);
        hp::launch_cols(g.H, c->stream, -1, c->X, (int)c->src_chan.size(), g, c->tw[g.H], c->fv);
    }
    {
        Timed t(c, HP_K_SPECTRAL);
        const int L = c->is3d ? g.D : g.H;
        hp::launch_spectral(L, c->is3d ? 3 : 2, c->stream, c->X, c->Y, c->Kq, c->d_spec_work, c->n_spec_work,
                            c->d_klist_spec, g, c->tw[L], c->fv);
    }
    if (c->is3d) {
        Timed t(c, HP_K_COLS_INV);
        hp::launch_cols(g.H, c->stream, +1, c->Y, c->K, g, c->tw[g.H], c->fv);
    }
    {
        Timed t(c, HP_K_ROWS_INV);
        hp::launch_rows_inv(g.W, c->stream, c->Y, c->U, c->d_tgt_work, c->n_tgt_work, c->d_klist_tgt, c->d_gp, g,
                            c->tw[g.W], c->fv);
    }
    HIP_TRY(c, hipGetLastError());
    return HP_OK;
}

bool fast_stencil_ok(const hp_ctx* c) {
    return c->is3d ? hp::stencil3d_supported(c->g, c->cfg.flow.distribution_size)
                   : hp::stencil2d_supported(c->g, c->cfg.flow.distribution_size, c->tile);
}

hp_status enqueue_fast_stencil(hp_ctx* c, const float* A, const float* U, float* out) {
    timed_pass(c, HP_K_STENCIL, true, [&] {  // A, U -> out (and an idempotent flag OR)
        if (c->is3d)
            hp::launch_stencil3d(c->stream, A, U, out, c->C, c->g, c->cfg.dt, c->cfg.flow.beta_a, c->cfg.flow.n,
                                 c->cfg.flow.distribution_size, c->d_flags, c->exact);
        else
            hp::launch_stencil2d(c->stream, A, U, out, c->C, c->g, c->cfg.dt, c->cfg.flow.beta_a, c->cfg.flow.n,
                                 c->cfg.flow.distribution_size, c->d_flags, c->tile, c->exact);
    });
    HIP_TRY(c, hipGetLastError());
    return HP_OK;
}

// General path for one step whose affinity is in `U`: boxes, sync, radius, gather.
hp_status general_stencil(hp_ctx* c, const float* A, const float* U, float* out) {
    hp::launch_clear_ints(c->stream, c->d_flags + 1, 2);
    {
        Timed t(c, HP_K_GENERAL);
        hp::launch_boxes(c->stream, A, U, c->C, c->g, c->cfg.dt, c->cfg.flow.beta_a, c->cfg.flow.n,
                         c->cfg.flow.distribution_size, c->box_lo, c->box_hi, c->d_flags + 1);
    }
    int h[2] = {0, 0};
    HIP_TRY(c, hipMemcpyAsync(h, c->d_flags + 1, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (h[1]) return c->fail(HP_ERR_NONFINITE, "state or flow holds NaN/Inf");
    const int R = std::max(h[0], 1);
    c->max_radius = std::max<uint32_t>(c->max_radius, (uint32_t)R);
    // the gather visits (2R+1)^d sources per destination (capped at the grid): refuse work that
    // would keep the GPU busy for minutes rather than run it
    auto win = [](int R_, int n) { return (double)std::min(2 * R_ + 1, n); };
    const double visits = win(R, c->g.W) * win(R, c->g.H) * (c->is3d ? win(R, c->g.D) : 1.0) * c->C * c->g.N;
    if (visits > 2e10)
        return c->fail(HP_ERR_UNSUPPORTED_SIZE, "flow displacement of " + std::to_string(R) +
                                                    " cells is beyond the general gather's work bound");
    {
        Timed t(c, HP_K_GENERAL);
        hp::launch_gather_general(c->stream, c->box_lo, c->box_hi, out, c->C, c->g, c->cfg.flow.distribution_size, R,
                                  c->exact);
    }
    HIP_TRY(c, hipGetLastError());
    c->general_steps++;
    return HP_OK;
}

// One step with a host check of the speculative radius (used by hp_step and by the replay
// after a flagged run).
hp_status safe_step(hp_ctx* c) {
    float* A = c->A[c->cur];
    float* out = c->A[1 - c->cur];
    hp_status st = enqueue_affinity(c, A);
    if (st != HP_OK) return st;
    bool need_general = !fast_stencil_ok(c);
    if (!need_general) {
        hp::launch_clear_ints(c->stream, c->d_flags, 1);
        st = enqueue_fast_stencil(c, A, c->U, out);
        if (st != HP_OK) return st;
        int f = 0;
        HIP_TRY(c, hipMemcpyAsync(&f, c->d_flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (f & 2) return c->fail(HP_ERR_NONFINITE, "state or flow holds NaN/Inf");
        if (f & 1) need_general = true;
        else c->max_radius = std::max<uint32_t>(c->max_radius, 2);
    }
    if (need_general) {
        st = general_stencil(c, A, c->U, out);
        if (st != HP_OK) return st;
    }
    c->cur = 1 - c->cur;
    c->time += c->cfg.dt;
    c->step += 1;
    return HP_OK;
}

// Steps per replay of the long graphs: an even count, so a replay returns to its parity and
// long runs pay one graph launch per kGraphSteps steps.
constexpr int kGraphSteps = 16;

hp_status capture_steps(hp_ctx* c, int b, int n, hipGraphExec_t* out) {
    hipGraph_t graph = nullptr;
    HIP_TRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hp_status st = HP_OK;
    for (int i = 0; i < n && st == HP_OK; i++) {
        const int p = b ^ (i & 1);
        st = enqueue_affinity(c, c->A[p]);
        if (st == HP_OK) st = enqueue_fast_stencil(c, c->A[p], c->U, c->A[1 - p]);
    }
    hipError_t ec = hipStreamEndCapture(c->stream, &graph);
    if (st != HP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
    }
    if (ec != hipSuccess) return c->fail(HP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    hipError_t e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return c->fail(HP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return HP_OK;
}

hp_status build_graphs(hp_ctx* c) {
    for (int b = 0; b < 2; b++) {
        hp_status st = HP_OK;
        if (!c->graph[b]) st = capture_steps(c, b, 1, &c->graph[b]);
        if (st == HP_OK && !c->graph_run[b]) st = capture_steps(c, b, kGraphSteps, &c->graph_run[b]);
        if (st != HP_OK) return st;
    }
    return HP_OK;
}

// Switch the reintegration arithmetic; the replay graphs are captured again for it.
hp_status set_exact(hp_ctx* c, bool exact) {
    if (c->exact == exact) return HP_OK;
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    for (int b = 0; b < 2; b++) {
        if (c->graph[b]) (void)hipGraphExecDestroy(c->graph[b]);
        if (c->graph_run[b]) (void)hipGraphExecDestroy(c->graph_run[b]);
        c->graph[b] = c->graph_run[b] = nullptr;
    }
    c->exact = exact;
    return fast_stencil_ok(c) ? build_graphs(c) : HP_OK;
}

hp_status advance(hp_ctx* c, uint64_t steps) {
    if (steps == 0) return HP_OK;
    if (!fast_stencil_ok(c)) {
        for (uint64_t i = 0; i < steps; i++) {
            hp_status st = safe_step(c);
            if (st != HP_OK) return st;
        }
        if (c->profiling) collect_profile(c);
        return HP_OK;
    }
    // Speculative run: every step uses the fused stencil; a sticky flag records whether any
    // source left the fused radius.  If it did, restore the snapshot and replay step by step.
    const size_t bytes = (size_t)c->C * c->g.N * sizeof(float);
    const int start_cur = c->cur;
    const float start_time = c->time;
    const uint64_t start_step = c->step;
    HIP_TRY(c, hipMemcpyAsync(c->snapshot, c->A[c->cur], bytes, hipMemcpyDeviceToDevice, c->stream));
    hp::launch_clear_ints(c->stream, c->d_flags, 1);
    if (!c->profiling) {
        hp_status st = build_graphs(c);
        if (st != HP_OK) return st;
    }
    for (uint64_t i = 0; i < steps;) {
        if (c->profiling) {
            float* A = c->A[c->cur];
            hp_status st = enqueue_affinity(c, A);
            if (st == HP_OK) st = enqueue_fast_stencil(c, A, c->U, c->A[1 - c->cur]);
            if (st != HP_OK) return st;
        } else if (steps - i >= (uint64_t)kGraphSteps) {
            HIP_TRY(c, hipGraphLaunch(c->graph_run[c->cur], c->stream));
            for (int k = 0; k < kGraphSteps; k++) c->time += c->cfg.dt;  // f32, one dt per step
            c->step += kGraphSteps;
            i += kGraphSteps;
            continue;  // an even count of steps: the parity is unchanged
        } else {
            HIP_TRY(c, hipGraphLaunch(c->graph[c->cur], c->stream));
        }
        c->cur = 1 - c->cur;
        c->time += c->cfg.dt;
        c->step += 1;
        i += 1;
    }
    int f = 0;
    HIP_TRY(c, hipMemcpyAsync(&f, c->d_flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->profiling) collect_profile(c);
    if (f == 0) {
        c->max_radius = std::max<uint32_t>(c->max_radius, 2);
        return HP_OK;
    }
    // replay from the snapshot with a per-step check
    c->cur = start_cur;
    c->time = start_time;
    c->step = start_step;
    HIP_TRY(c, hipMemcpyAsync(c->A[c->cur], c->snapshot, bytes, hipMemcpyDeviceToDevice, c->stream));
    for (uint64_t i = 0; i < steps; i++) {
        hp_status st = safe_step(c);
        if (st != HP_OK) return st;
    }
    if (c->profiling) collect_profile(c);
    return HP_OK;
}

hp_status check_host_state(hp_ctx* c, const float* const* ch) {
    if (!ch) return c->fail(HP_ERR_NULL, "channel pointer array is NULL");
    for (int i = 0; i < c->C; i++)
        if (!ch[i]) return c->fail(HP_ERR_NULL, "channel pointer is NULL");
    return HP_OK;
}

// The host side of a state transfer: the channels move between the caller's (pageable) arrays
// and a pinned staging buffer of the context on a few host threads, and between the staging
// buffer and the device in one DMA copy.  Host threads for states of at least 1 Mi values.
hp_status ensure_host_stage(hp_ctx* c) {
    if (c->h_stage) return HP_OK;
    HIP_TRY(c, hipHostMalloc((void**)&c->h_stage, std::max<size_t>(1, (size_t)c->C * c->g.N) * sizeof(float),
                             hipHostMallocDefault));
    return HP_OK;
}

// f(lo, hi) over [0, n) in contiguous parts on up to 8 threads (the caller runs the first part).
template <class F>
void host_parallel(size_t n, const F& f) {
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = n < (1u << 20) ? 1 : std::min<size_t>(std::min(8u, hw), n >> 18);
    if (T <= 1) {
        f((size_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    const size_t part = (n + T - 1) / T;
    for (size_t t = 1; t < T; t++) {
        const size_t lo = std::min(n, t * part), hi = std::min(n, lo + part);
        try {
            th.emplace_back([&f, lo, hi] { f(lo, hi); });
        } catch (...) {  // no thread: run the part here
            f(lo, hi);
        }
    }
    f((size_t)0, std::min(n, part));
    for (auto& t : th) t.join();
}

// Every channel is checked (in the same pass that stages it) before anything is copied to the
// device, so a refused state leaves the device state as it was.
hp_status upload_state(hp_ctx* c, const float* const* ch, float* dst) {
    hp_status st = ensure_host_stage(c);
    if (st != HP_OK) return st;
    const size_t N = (size_t)c->g.N, total = (size_t)c->C * N;
    std::vector<char> bad(64, 0);  // one flag per possible part, written by its thread only
    std::atomic<int> part_id{0};
    host_parallel(total, [&](size_t lo, size_t hi) {
        const int me = part_id.fetch_add(1);
        unsigned acc = 0;
        for (size_t k = lo; k < hi;) {
            const size_t i = k / N, j0 = k - i * N, j1 = std::min(N, j0 + (hi - k));
            const float* s = ch[i] + j0;
            float* d = c->h_stage + k;
            const size_t m = j1 - j0;
            for (size_t j = 0; j < m; j++) {
                const float v = s[j];
                d[j] = v;
                unsigned u;
                memcpy(&u, &v, 4);
                acc |= (unsigned)((u & 0x7f800000u) == 0x7f800000u);  // Inf or NaN
            }
            k += m;
        }
        bad[me & 63] = (char)acc;
    });
    for (char b : bad)
        if (b) return c->fail(HP_ERR_NONFINITE, "input state holds NaN/Inf");
    HIP_TRY(c, hipMemcpyAsync(dst, c->h_stage, total * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    return HP_OK;
}

hp_status download(hp_ctx* c, const float* src, float* const* ch) {
    hp_status st = ensure_host_stage(c);
    if (st != HP_OK) return st;
    const size_t N = (size_t)c->g.N, total = (size_t)c->C * N;
    HIP_TRY(c, hipMemcpyAsync(c->h_stage, src, total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    host_parallel(total, [&](size_t lo, size_t hi) {
        for (size_t k = lo; k < hi;) {
            const size_t i = k / N, j0 = k - i * N, m = std::min(N - j0, hi - k);
            memcpy(ch[i] + j0, c->h_stage + k, m * sizeof(float));
            k += m;
        }
    });
    return HP_OK;
}

void release(hp_ctx* c) {
    if (!c) return;
    for (int b = 0; b < 2; b++) {
        if (c->graph[b]) (void)hipGraphExecDestroy(c->graph[b]);
        if (c->graph_run[b]) (void)hipGraphExecDestroy(c->graph_run[b]);
    }
    for (auto& l : c->launches) {
        (void)hipEventDestroy(l.a);
        (void)hipEventDestroy(l.b);
    }
    for (auto e : c->event_pool) (void)hipEventDestroy(e);
    void* ptrs[] = {c->A[0], c->A[1], c->U, c->X, c->Y, c->Kq, c->d_src_chan, c->d_spec_work, c->d_klist_spec,
                    c->d_tgt_work, c->d_klist_tgt, c->d_gp, c->d_flags, c->box_lo, c->box_hi, c->d_partials,
                    c->d_mass, c->snapshot, c->d_dk, c->d_taps, c->d_cols, c->d_conv};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (auto& kv : c->tw) (void)hipFree(kv.second);
    for (auto& kv : c->tws) (void)hipFree(kv.second);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

// Plans the autotuner found, per configuration shape (sides, channels, each kernel's source
// and target, distribution size), for the contexts created later in this process.
std::mutex g_plan_mu;
std::map<std::string, hp_plan> g_plan_cache;

std::string plan_key(const hp_config* cfg) {
    std::string k = std::to_string(cfg->width) + "x" + std::to_string(cfg->height) + "x" +
                    std::to_string(cfg->depth) + "c" + std::to_string(cfg->channels) + "s" +
                    std::to_string(cfg->flow.distribution_size) + ":";
    for (uint32_t i = 0; i < cfg->num_kernels; i++)
        k += std::to_string(cfg->kernels[i].source_channel) + ">" + std::to_string(cfg->kernels[i].target_channel) + ",";
    return k;
}

// The tuned plan of this shape if there is one, else the built-in defaults; then the
// environment overrides, which always win: HP_SPECTRAL_GROUP (1 per source, 0 per kernel),
// HP_FFT_PATH=generic (LDS FFTs even where the wave-level path applies).
hp_plan default_plan(const hp_config* cfg) {
    hp_plan p{};
    p.fft_path = 0;
    p.spectral_group = -1;
    p.radix_order = 0;
    p.column_split = 1;
    p.stencil_tile = 0;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_plan_cache.find(plan_key(cfg));
        if (it != g_plan_cache.end()) p = it->second;
    }
    const char* grp = getenv("HP_SPECTRAL_GROUP");
    const char* path = getenv("HP_FFT_PATH");
    if (path && strcmp(path, "generic") == 0) p.fft_path = 1;
    if (grp) p.spectral_group = atoi(grp);
    return p;
}

hp_status create(const hp_config* cfg, int device, const hp_plan* req, hp_ctx** out) {
    *out = nullptr;
    hp_status st = hp::validate_config(cfg, nullptr, nullptr);
    if (st != HP_OK) {
        g_global_error = "invalid configuration";
        return st;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        g_global_error = "no HIP device";
        return HP_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        g_global_error = "device ordinal out of range";
        return HP_ERR_NO_DEVICE;
    }
    const hp_plan want = req ? *req : default_plan(cfg);
    if (want.fft_path < 0 || want.fft_path > 1 || want.spectral_group < -1 || want.spectral_group > 1 ||
        want.radix_order < 0 || want.radix_order > 1 || want.column_split < 1 || want.column_split > 2 ||
        want.stencil_tile < 0 || want.stencil_tile >= hp::stencil2d_tile_count()) {
        g_global_error = "plan field out of range";
        return HP_ERR_SHAPE;
    }
    hp::MemberPlan plan;
    std::string perr;
    st = hp::plan_member(cfg, want.fft_path == 0, want.spectral_group, &plan, &perr);
    if (st != HP_OK) {
        g_global_error = perr;
        return st;
    }
    hp_ctx* c = new hp_ctx();
    c->device = device;
    c->cfg = *cfg;
    c->kernels.assign(cfg->kernels, cfg->kernels + cfg->num_kernels);
    c->rings.assign(cfg->rings, cfg->rings + cfg->num_rings);
    c->cfg.kernels = c->kernels.data();
    c->cfg.rings = c->rings.data();
    c->C = plan.C;
    c->K = plan.K;
    c->is3d = plan.is3d;
    c->wave = plan.wave;
    c->plan = want;
    c->plan.fft_path = plan.wave ? 0 : 1;
    c->plan.spectral_group = plan.grouped ? 1 : 0;
    c->fv.rp = want.radix_order;
    c->fv.cs = want.column_split;
    c->tile = want.stencil_tile;
    // a tile whose strips need more rows than the grid has falls back to the default shape
    if (!plan.is3d && !hp::stencil2d_supported(plan.g, cfg->flow.distribution_size, c->tile) &&
        hp::stencil2d_supported(plan.g, cfg->flow.distribution_size, 0))
        c->tile = 0;
    c->plan.stencil_tile = c->tile;
    c->g = plan.g;
    const hp::Geo& g = c->g;
    const bool is3d = c->is3d;
    c->src_chan = plan.src_chan;
    c->tgt_chan = plan.tgt_chan;
    const std::vector<float>& kq = plan.kq;
    const std::vector<hp::SpecWork>& sw = plan.spec_work;
    const std::vector<int>& kls = plan.klist_spec;
    const std::vector<hp::TgtWork>& tw = plan.tgt_work;
    const std::vector<int>& klt = plan.klist_tgt;
    const std::vector<hp::GrowthParams>& gp = plan.gp;
    c->n_spec_work = (int)sw.size();
    c->spec_work = sw;
    c->n_tgt_work = (int)tw.size();
    c->mass_parts = 256;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
        release(c);
        g_global_error = "hipStreamCreate failed";
        return HP_ERR_HIP;
    }
    const size_t N = (size_t)g.N, CN = (size_t)c->C * N;
    const size_t nsrc = std::max<size_t>(1, c->src_chan.size());
#define TRY(x)                                \
    do {                                      \
        hp_status s_ = (x);                   \
        if (s_ != HP_OK) {                    \
            g_global_error = c->err;          \
            release(c);                       \
            return s_;                        \
        }                                     \
    } while (0)
    TRY(dalloc(c, &c->A[0], CN));
    TRY(dalloc(c, &c->A[1], CN));
    TRY(dalloc(c, &c->U, CN));
    TRY(dalloc(c, &c->snapshot, CN));
    c->direct = plan.direct;
    c->direct_rmax = plan.rmax;
    // spectra only for the FFT path
    TRY(dalloc(c, &c->X, c->direct ? 1 : nsrc * g.spec));
    TRY(dalloc(c, &c->Y, c->direct ? 1 : (size_t)std::max(1, c->K) * g.spec));
    TRY(dalloc(c, &c->Kq, kq.size()));
    if (c->direct) {
        TRY(dalloc(c, &c->d_dk, std::max<size_t>(1, plan.dk.size())));
        TRY(dalloc(c, &c->d_taps, std::max<size_t>(1, plan.taps.size())));
        TRY(dalloc(c, &c->d_cols, std::max<size_t>(1, plan.cols.size())));
        TRY(dalloc(c, &c->d_conv, (size_t)std::max(1, c->K) * N));
        TRY(upload(c, c->d_dk, plan.dk));
        TRY(upload(c, c->d_taps, plan.taps));
        TRY(upload(c, c->d_cols, plan.cols));
    }
    TRY(dalloc(c, &c->d_src_chan, nsrc));
    TRY(dalloc(c, &c->d_spec_work, std::max<size_t>(1, sw.size())));
    TRY(dalloc(c, &c->d_klist_spec, std::max<size_t>(1, kls.size())));
    TRY(dalloc(c, &c->d_tgt_work, std::max<size_t>(1, tw.size())));
    TRY(dalloc(c, &c->d_klist_tgt, std::max<size_t>(1, klt.size())));
    TRY(dalloc(c, &c->d_gp, gp.size()));
    TRY(dalloc(c, &c->d_flags, 4));
    TRY(dalloc(c, &c->box_lo, CN));
    TRY(dalloc(c, &c->box_hi, CN));
    TRY(dalloc(c, &c->d_partials, (size_t)c->C * c->mass_parts));
    TRY(dalloc(c, &c->d_mass, (size_t)c->C + 1));
    TRY(upload(c, c->Kq, kq));
    TRY(upload(c, c->d_src_chan, c->src_chan));
    TRY(upload(c, c->d_spec_work, sw));
    TRY(upload(c, c->d_klist_spec, kls));
    TRY(upload(c, c->d_tgt_work, tw));
    TRY(upload(c, c->d_klist_tgt, klt));
    TRY(upload(c, c->d_gp, gp));
    TRY(make_twiddles(c, g.W));
    TRY(make_twiddles(c, g.H));
    if (c->wave) {
        TRY(make_stage_twiddles(c, g.W));
        TRY(make_stage_twiddles(c, g.H));
    }
    if (is3d) TRY(make_twiddles(c, g.D));
    // zero state, affinity (channels no kernel targets keep U = 0) and flags
    if (hipMemset(c->A[0], 0, CN * 4) != hipSuccess || hipMemset(c->A[1], 0, CN * 4) != hipSuccess ||
        hipMemset(c->U, 0, CN * 4) != hipSuccess || hipMemset(c->d_flags, 0, 16) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        release(c);
        g_global_error = "hipMemset failed";
        return HP_ERR_HIP;
    }
    // the replay graphs are captured here, so that the first advance (and a timed one right
    // after a plan change) launches them instead of capturing them
    if (fast_stencil_ok(c)) TRY(build_graphs(c));
#undef TRY
    *out = c;
    return HP_OK;
}

template <typename F>
hp_status guarded(hp_ctx* c, F&& f) {
    if (!c) return HP_ERR_NULL;
    try {
        if (hipSetDevice(c->device) != hipSuccess) return c->fail(HP_ERR_HIP, "hipSetDevice failed");
        return f();
    } catch (const std::exception& e) {
        return c->fail(HP_ERR_HIP, e.what());
    } catch (...) {
        return c->fail(HP_ERR_HIP, "unknown exception");
    }
}

// temporary device buffers for the stage entry points
struct DevBuf {
    void* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

hp_status stage_device(int device) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
        g_global_error = "no HIP device";
        return HP_ERR_NO_DEVICE;
    }
    if (hipSetDevice(device) != hipSuccess) return HP_ERR_HIP;
    return HP_OK;
}

#define STAGE_TRY(expr)                                           \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) {                                   \
            g_global_error = std::string(#expr) + ": " + hipGetErrorString(e_); \
            return HP_ERR_HIP;                                    \
        }                                                         \
    } while (0)

hp_status to_dev(DevBuf& b, const float* host, size_t n) {
    STAGE_TRY(hipMalloc(&b.p, std::max<size_t>(1, n) * sizeof(float)));
    if (host) STAGE_TRY(hipMemcpy(b.p, host, n * sizeof(float), hipMemcpyHostToDevice));
    return HP_OK;
}

hp_status advect_stage(int device, const float* cur, const float* fx, const float* fy, const float* fz, float* next,
                       uint32_t W, uint32_t H, uint32_t D, float dt, float s) {
    if (!cur || !fx || !fy || !next || (D > 1 && !fz)) return HP_ERR_NULL;
    if (W == 0 || H == 0 || D == 0) return HP_ERR_INVALID_DIMENSIONS;
    hp_status st = stage_device(device);
    if (st != HP_OK) return st;
    hp::Geo g{};
    g.W = (int)W;
    g.H = (int)H;
    g.D = (int)D;
    g.N = (long long)W * H * D;
    const size_t n = (size_t)g.N;
    DevBuf a, bx, by, bz, o, lo, hi, fl;
    if ((st = to_dev(a, cur, n)) != HP_OK || (st = to_dev(bx, fx, n)) != HP_OK || (st = to_dev(by, fy, n)) != HP_OK ||
        (st = to_dev(o, nullptr, n)) != HP_OK)
        return st;
    if (D > 1 && (st = to_dev(bz, fz, n)) != HP_OK) return st;
    STAGE_TRY(hipMalloc(&lo.p, n * sizeof(float4)));
    STAGE_TRY(hipMalloc(&hi.p, n * sizeof(float4)));
    STAGE_TRY(hipMalloc(&fl.p, 4 * sizeof(int)));
    STAGE_TRY(hipMemset(fl.p, 0, 4 * sizeof(int)));
    hp::launch_boxes_from_flow(nullptr, (const float*)a.p, (const float*)bx.p, (const float*)by.p,
                               D > 1 ? (const float*)bz.p : nullptr, g, dt, s, (float4*)lo.p, (float4*)hi.p,
                               (int*)fl.p);
    int h[2];
    STAGE_TRY(hipMemcpy(h, fl.p, sizeof(h), hipMemcpyDeviceToHost));
    if (h[1]) {
        g_global_error = "flow or mass holds NaN/Inf";
        return HP_ERR_NONFINITE;
    }
    const int R = std::max(h[0], 1);
    auto win = [](int R_, int n_) { return (double)std::min(2 * R_ + 1, n_); };
    if (win(R, g.W) * win(R, g.H) * (D > 1 ? win(R, g.D) : 1.0) * (double)g.N > 2e10) {
        g_global_error = "flow displacement beyond the general gather's work bound";
        return HP_ERR_UNSUPPORTED_SIZE;
    }
    hp::launch_gather_general(nullptr, (const float4*)lo.p, (const float4*)hi.p, (float*)o.p, 1, g, s, R);
    STAGE_TRY(hipGetLastError());
    STAGE_TRY(hipMemcpy(next, o.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return HP_OK;
}

}  // namespace

// ================================== C ABI ===========================================
extern "C" {

hp_status hp_validate_config(const hp_config* cfg, uint32_t* bad_kernel, uint32_t* bad_channel) {
    return hp::validate_config(cfg, bad_kernel, bad_channel);
}

hp_status hp_seed_generate(const hp_seed* seed, uint32_t width, uint32_t height, uint32_t depth, uint32_t channels,
                           float* const* out_channels) {
    try {
        return hp::seed_generate(seed, width, height, depth, channels, out_channels);
    } catch (...) {
        return HP_ERR_SHAPE;
    }
}

hp_status hp_create(const hp_config* cfg, int device, hp_ctx** out) {
    if (!out) return HP_ERR_NULL;
    *out = nullptr;
    if (!cfg) return HP_ERR_NULL;
    try {
        return create(cfg, device, nullptr, out);
    } catch (const std::exception& e) {
        g_global_error = e.what();
        return HP_ERR_HIP;
    } catch (...) {
        g_global_error = "unknown exception";
        return HP_ERR_HIP;
    }
}

void hp_destroy(hp_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    release(ctx);
}

const char* hp_last_error(const hp_ctx* ctx) { return ctx ? ctx->err.c_str() : "NULL context"; }
const char* hp_last_global_error(void) { return g_global_error.c_str(); }

}  // extern "C"

hp_status hp::ctx_device_state(hp_ctx* ctx, hp::DeviceState* out) {
    if (!ctx || !out) return HP_ERR_NULL;
    out->data = ctx->A[ctx->cur];
    out->floats = (size_t)ctx->C * (size_t)ctx->g.N;
    out->stream = ctx->stream;
    out->device = ctx->device;
    out->W = (uint32_t)ctx->g.W;
    out->H = (uint32_t)ctx->g.H;
    out->D = (uint32_t)ctx->g.D;
    out->C = (uint32_t)ctx->C;
    return HP_OK;
}

hp_status hp::ctx_create_lds(const hp_config* cfg, int device, hp_ctx** out) {
    if (!cfg || !out) return HP_ERR_NULL;
    hp_plan p = default_plan(cfg);
    p.fft_path = 1;
    return create(cfg, device, &p, out);
}

hp_status hp::ctx_view(hp_ctx* ctx, hp::CtxView* out) {
    if (!ctx || !out) return HP_ERR_NULL;
    out->stream = ctx->stream;
    out->device = ctx->device;
    out->g = ctx->g;
    out->C = ctx->C;
    out->K = ctx->K;
    out->A[0] = ctx->A[0];
    out->A[1] = ctx->A[1];
    out->cur = &ctx->cur;
    out->time = &ctx->time;
    out->step = &ctx->step;
    out->box_lo = ctx->box_lo;
    out->box_hi = ctx->box_hi;
    out->snapshot = ctx->snapshot;
    return HP_OK;
}

hp_status hp::ctx_enqueue_conv(hp_ctx* c, const float* A, float* conv) {
    if (!c || !A || !conv) return HP_ERR_NULL;
    if (c->wave || c->is3d) return c->fail(HP_ERR_UNSUPPORTED_SIZE, "convolutions need a 2D context on the LDS path");
    if (c->K == 0) return HP_OK;
    if (c->direct) {
        if (!hp::launch_direct_affinity(c->stream, A, nullptr, conv, c->d_dk, c->d_taps, c->d_cols, c->K, c->C, c->g,
                                        c->direct_rmax))
            return c->fail(HP_ERR_UNSUPPORTED_SIZE, "kernel radius beyond the direct convolution's LDS tile");
        HIP_TRY(c, hipGetLastError());
        return HP_OK;
    }
    const hp::Geo& g = c->g;
    hp::launch_rows_fwd(g.W, c->stream, A, c->X, c->d_src_chan, (int)c->src_chan.size(), g, c->tw[g.W], c->fv);
    hp::launch_spectral(g.H, 2, c->stream, c->X, c->Y, c->Kq, c->d_spec_work, c->n_spec_work, c->d_klist_spec, g,
                        c->tw[g.H], c->fv);
    hp::launch_rows_inv_conv(g.W, c->stream, c->Y, conv, c->d_tgt_work, c->n_tgt_work, c->d_klist_tgt, g, c->tw[g.W],
                             c->fv);
    HIP_TRY(c, hipGetLastError());
    return HP_OK;
}

const char* hp::ctx_error(const hp_ctx* ctx) { return ctx ? ctx->err.c_str() : ""; }

extern "C" {

hp_status hp_get_config(const hp_ctx* ctx, hp_config* out) {
    if (!ctx || !out) return HP_ERR_NULL;
    *out = ctx->cfg;
    return HP_OK;
}

hp_status hp_set_state(hp_ctx* ctx, const float* const* channels, float time, uint64_t step) {
    return guarded(ctx, [&]() -> hp_status {
        hp_status st = check_host_state(ctx, channels);
        if (st != HP_OK) return st;
        st = upload_state(ctx, channels, ctx->A[ctx->cur]);
        if (st != HP_OK) return st;
        ctx->time = time;
        ctx->step = step;
        return HP_OK;
    });
}

hp_status hp_get_state(hp_ctx* ctx, float* const* channels, float* time, uint64_t* step) {
    return guarded(ctx, [&]() -> hp_status {
        hp_status st = check_host_state(ctx, channels);
        if (st != HP_OK) return st;
        st = download(ctx, ctx->A[ctx->cur], channels);
        if (st != HP_OK) return st;
        if (time) *time = ctx->time;
        if (step) *step = ctx->step;
        return HP_OK;
    });
}

hp_status hp_step(hp_ctx* ctx, float* const* channels, float* time, uint64_t* step) {
    return guarded(ctx, [&]() -> hp_status {
        hp_status st = check_host_state(ctx, channels);
        if (st != HP_OK) return st;
        st = upload_state(ctx, channels, ctx->A[ctx->cur]);
        if (st != HP_OK) return st;
        ctx->time = time ? *time : 0.0f;
        ctx->step = step ? *step : 0;
        st = safe_step(ctx);
        if (st != HP_OK) return st;
        if (ctx->profiling) collect_profile(ctx);
        st = download(ctx, ctx->A[ctx->cur], channels);
        if (st != HP_OK) return st;
        if (time) *time = ctx->time;
        if (step) *step = ctx->step;
        return HP_OK;
    });
}

hp_status hp_run(hp_ctx* ctx, float* const* channels, float* time, uint64_t* step, uint64_t steps) {
    return guarded(ctx, [&]() -> hp_status {
        hp_status st = check_host_state(ctx, channels);
        if (st != HP_OK) return st;
        st = upload_state(ctx, channels, ctx->A[ctx->cur]);
        if (st != HP_OK) return st;
        ctx->time = time ? *time : 0.0f;
        ctx->step = step ? *step : 0;
        st = advance(ctx, steps);
        if (st != HP_OK) return st;
        st = download(ctx, ctx->A[ctx->cur], channels);
        if (st != HP_OK) return st;
        if (time) *time = ctx->time;
        if (step) *step = ctx->step;
        return HP_OK;
    });
}

hp_status hp_advance(hp_ctx* ctx, uint64_t steps) {
    return guarded(ctx, [&]() -> hp_status { return advance(ctx, steps); });
}

hp_status hp_mass_f64(hp_ctx* ctx, double* per_channel, double* total) {
    return guarded(ctx, [&]() -> hp_status {
        hp::launch_mass_f64(ctx->stream, ctx->A[ctx->cur], ctx->C, ctx->g.N, ctx->d_partials, ctx->mass_parts,
                            ctx->d_mass);
        std::vector<double> h(ctx->C + 1);
        HIP_TRY(ctx, hipMemcpyAsync(h.data(), ctx->d_mass, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                                    ctx->stream));
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        if (per_channel) std::copy(h.begin(), h.begin() + ctx->C, per_channel);
        if (total) *total = h[ctx->C];
        return HP_OK;
    });
}

hp_status hp_set_conservation(hp_ctx* ctx, int mode) {
    return guarded(ctx, [&]() -> hp_status {
        if (mode != HP_CONSERVATION_REFERENCE && mode != HP_CONSERVATION_EXACT)
            return ctx->fail(HP_ERR_SHAPE, "conservation mode must be HP_CONSERVATION_REFERENCE or _EXACT");
        return set_exact(ctx, mode == HP_CONSERVATION_EXACT);
    });
}

hp_status hp_get_conservation(const hp_ctx* ctx, int* mode) {
    if (!ctx || !mode) return HP_ERR_NULL;
    *mode = ctx->exact ? HP_CONSERVATION_EXACT : HP_CONSERVATION_REFERENCE;
    return HP_OK;
}

hp_status hp_stats(const hp_ctx* ctx, uint32_t* max_gather_radius, uint64_t* general_path_steps) {
    if (!ctx) return HP_ERR_NULL;
    if (max_gather_radius) *max_gather_radius = ctx->max_radius;
    if (general_path_steps) *general_path_steps = ctx->general_steps;
    return HP_OK;
}

}  // extern "C"

namespace {

// A new device plan for c, keeping its state, time, step and counters.
hp_status set_plan(hp_ctx* c, const hp_plan& p) {
    hp_ctx* n = nullptr;
    hp_status st = create(&c->cfg, c->device, &p, &n);
    if (st != HP_OK) return c->fail(st, g_global_error);
    const size_t bytes = (size_t)c->C * c->g.N * sizeof(float);
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(n->A[0], c->A[c->cur], bytes, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) {
        release(n);
        return c->fail(HP_ERR_HIP, std::string("plan change: ") + hipGetErrorString(e));
    }
    n->cur = 0;
    n->time = c->time;
    n->step = c->step;
    n->max_radius = c->max_radius;
    n->general_steps = c->general_steps;
    n->profiling = c->profiling;
    n->prof_reps = c->prof_reps;
    if (c->exact) {  // the new plan's graphs were captured for the reference arithmetic
        st = set_exact(n, true);
        if (st != HP_OK) {
            c->err = n->err;
            release(n);
            return st;
        }
    }
    std::swap(*c, *n);
    release(n);  // the old device plan
    return HP_OK;
}

bool same_fft(const hp_plan& a, const hp_plan& b) {
    return a.fft_path == b.fft_path && a.spectral_group == b.spectral_group && a.radix_order == b.radix_order &&
           a.column_split == b.column_split;
}

// Coordinate search: every FFT configuration with the current tile, then every tile with the
// fastest FFT configuration (the two are separate kernels, so their costs add).
hp_status autotune(hp_ctx* c, uint32_t steps, hp_plan* best_out, uint32_t maxr, hp_plan* tried, double* us_out,
                   uint32_t* ntried) {
    if (steps == 0) steps = 20;
    const size_t CN = (size_t)c->C * c->g.N;
    std::vector<float> snap(CN), ref, got(CN);
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(snap.data(), c->A[c->cur], CN * sizeof(float), hipMemcpyDeviceToHost));
    const float t0 = c->time;
    const uint64_t s0 = c->step, g0 = c->general_steps;
    const uint32_t r0 = c->max_radius;
    const hp_plan start = c->plan;
    hipEvent_t ea = nullptr, eb = nullptr;
    HIP_TRY(c, hipEventCreate(&ea));
    HIP_TRY(c, hipEventCreate(&eb));
    uint32_t n = 0;
    double best_us = INFINITY;
    hp_plan best = start;
    auto restore = [&]() -> hp_status {
        HIP_TRY(c, hipMemcpy(c->A[c->cur], snap.data(), CN * sizeof(float), hipMemcpyHostToDevice));
        return HP_OK;
    };
    auto measure = [&](const hp_plan& p) -> hp_status {
        hp_status st = set_plan(c, p);
        if (st != HP_OK) return HP_OK;  // a plan this configuration cannot take is skipped
        st = restore();
        if (st == HP_OK) st = advance(c, 2);  // graph capture and first touches
        if (st == HP_OK) st = restore();
        if (st != HP_OK) return st;
        HIP_TRY(c, hipEventRecord(ea, c->stream));
        st = advance(c, steps);
        if (st != HP_OK) return st;
        HIP_TRY(c, hipEventRecord(eb, c->stream));
        HIP_TRY(c, hipEventSynchronize(eb));
        float ms = 0.0f;
        HIP_TRY(c, hipEventElapsedTime(&ms, ea, eb));
        HIP_TRY(c, hipMemcpy(got.data(), c->A[c->cur], CN * sizeof(float), hipMemcpyDeviceToHost));
        bool ok = true;
        if (ref.empty()) {
            ref = got;
        } else {
            for (size_t i = 0; i < CN && ok; i++) ok = fabsf(got[i] - ref[i]) <= 1e-3f;
        }
        const double us = 1e3 * ms / steps;
        if (n < maxr) {
            if (tried) tried[n] = c->plan;
            if (us_out) us_out[n] = ok ? us : -us;  // negative: disagreed with the first plan
        }
        n++;
        if (ok && us < best_us) {
            best_us = us;
            best = c->plan;
        }
        return HP_OK;
    };
    hp_status st = HP_OK;
    std::vector<hp_plan> ffts{start};  // the current plan first: the reference result
    for (int path = 0; path <= 1; path++)
        for (int grp = 1; grp >= 0; grp--)
            for (int rp = 0; rp <= (path == 1 ? 1 : 0); rp++)
                for (int cs = 1; cs <= (path == 1 ? 2 : 1); cs++) {
                    hp_plan p = start;
                    p.fft_path = path;
                    p.spectral_group = grp;
                    p.radix_order = path == 1 ? rp : start.radix_order;
                    p.column_split = path == 1 ? cs : start.column_split;
                    if (path == 0 && (c->is3d || !hp::wave_fft_supported(c->g.W, c->g.H))) continue;
                    if (!same_fft(p, start)) ffts.push_back(p);
                }
    for (const hp_plan& p : ffts)
        if (st == HP_OK) st = measure(p);
    if (st == HP_OK && fast_stencil_ok(c)) {
        const hp_plan fft_best = best;
        for (int t = 0; t < hp::stencil2d_tile_count() && st == HP_OK; t++) {
            if (t == fft_best.stencil_tile) continue;
            hp_plan p = fft_best;
            p.stencil_tile = t;
            st = measure(p);
        }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (st != HP_OK) return st;
    st = set_plan(c, best);
    if (st != HP_OK) return st;
    st = restore();
    if (st != HP_OK) return st;
    c->time = t0;
    c->step = s0;
    c->general_steps = g0;
    c->max_radius = r0;
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        g_plan_cache[plan_key(&c->cfg)] = best;
    }
    if (best_out) *best_out = best;
    if (ntried) *ntried = n;
    return HP_OK;
}

}  // namespace

extern "C" {

hp_status hp_get_plan(const hp_ctx* ctx, hp_plan* out) {
    if (!ctx || !out) return HP_ERR_NULL;
    *out = ctx->plan;
    return HP_OK;
}

hp_status hp_set_plan(hp_ctx* ctx, const hp_plan* plan) {
    if (!plan) return HP_ERR_NULL;
    return guarded(ctx, [&]() -> hp_status { return set_plan(ctx, *plan); });
}

hp_status hp_autotune(hp_ctx* ctx, uint32_t steps, hp_plan* best, uint32_t max_results, hp_plan* tried,
                      double* us_per_step, uint32_t* n_tried) {
    return guarded(ctx, [&]() -> hp_status {
        return autotune(ctx, steps, best, max_results, tried, us_per_step, n_tried);
    });
}

hp_status hp_profile_enable(hp_ctx* ctx, int enable) {
    return guarded(ctx, [&]() -> hp_status {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        collect_profile(ctx);
        ctx->profiling = enable != 0;
        for (int i = 0; i < HP_K_NUM_CLASSES; i++) {
            ctx->prof_ms[i] = 0.0;
            ctx->prof_n[i] = 0;
        }
        return HP_OK;
    });
}

hp_status hp_profile_repeats(hp_ctx* ctx, uint32_t reps) {
    return guarded(ctx, [&]() -> hp_status {
        if (reps < 1 || reps > 1000) return ctx->fail(HP_ERR_SHAPE, "profile repeats must be 1..1000");
        ctx->prof_reps = (int)reps;
        return HP_OK;
    });
}

hp_status hp_profile_read(hp_ctx* ctx, double* ms, uint64_t* launches) {
    return guarded(ctx, [&]() -> hp_status {
        HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
        collect_profile(ctx);
        for (int i = 0; i < HP_K_NUM_CLASSES; i++) {
            if (ms) ms[i] = ctx->prof_ms[i];
            if (launches) launches[i] = ctx->prof_n[i];
        }
        return HP_OK;
    });
}

hp_status hp_kernel_bytes(const hp_ctx* ctx, double* bytes) {
    if (!ctx || !bytes) return HP_ERR_NULL;
    const hp::Geo& g = ctx->g;
    const double N = (double)g.N;
    const double S = 8.0 * g.Wh * (double)g.H * g.D;  // one half spectrum, complex64
    const double Kq = 4.0 * g.Wh * (double)g.Hq * g.Dq;
    const double nsrc = (double)ctx->src_chan.size(), K = ctx->K, C = ctx->C, nt = (double)ctx->tgt_chan.size();
    for (int i = 0; i < HP_K_NUM_CLASSES; i++) bytes[i] = 0.0;
    if (ctx->K > 0) {
        bytes[HP_K_ROWS_FWD] = nsrc * (4.0 * N + S);
        if (ctx->is3d) bytes[HP_K_COLS_FWD] = nsrc * 2.0 * S;
        // every work item reads its source spectrum once, then per kernel the folded
        // kernel spectrum and one filtered spectrum written
        double sp = 0.0;
        for (const hp::SpecWork& w : ctx->spec_work) sp += S + w.k_count * (Kq + S);
        bytes[HP_K_SPECTRAL] = sp;
        if (ctx->is3d) bytes[HP_K_COLS_INV] = K * 2.0 * S;
        bytes[HP_K_ROWS_INV] = K * S + nt * 4.0 * N;
    }
    bytes[HP_K_STENCIL] = 3.0 * C * 4.0 * N;
    bytes[HP_K_GENERAL] = 3.0 * C * 4.0 * N;
    return HP_OK;
}

hp_status hp_advect_mass_2d(int device, const float* current, const float* flow_x, const float* flow_y, float* next,
                            uint32_t width, uint32_t height, float dt, float distribution_size) {
    try {
        return advect_stage(device, current, flow_x, flow_y, nullptr, next, width, height, 1, dt, distribution_size);
    } catch (...) {
        return HP_ERR_HIP;
    }
}

hp_status hp_advect_mass_3d(int device, const float* current, const float* flow_x, const float* flow_y,
                            const float* flow_z, float* next, uint32_t width, uint32_t height, uint32_t depth,
                            float dt, float distribution_size) {
    try {
        return advect_stage(device, current, flow_x, flow_y, flow_z, next, width, height, depth, dt,
                            distribution_size);
    } catch (...) {
        return HP_ERR_HIP;
    }
}

hp_status hp_sobel_gradient_2d(int device, const float* grid, float* grad_x, float* grad_y, uint32_t width,
                               uint32_t height) {
    if (!grid || !grad_x || !grad_y) return HP_ERR_NULL;
    if (width == 0 || height == 0) return HP_ERR_INVALID_DIMENSIONS;
    hp_status st = stage_device(device);
    if (st != HP_OK) return st;
    const size_t n = (size_t)width * height;
    DevBuf g, gx, gy;
    if ((st = to_dev(g, grid, n)) != HP_OK || (st = to_dev(gx, nullptr, n)) != HP_OK ||
        (st = to_dev(gy, nullptr, n)) != HP_OK)
        return st;
    hp::launch_sobel2d(nullptr, (const float*)g.p, (float*)gx.p, (float*)gy.p, (int)width, (int)height);
    STAGE_TRY(hipGetLastError());
    STAGE_TRY(hipMemcpy(grad_x, gx.p, n * sizeof(float), hipMemcpyDeviceToHost));
    STAGE_TRY(hipMemcpy(grad_y, gy.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return HP_OK;
}

hp_status hp_compute_flow_2d(int device, const float* gux, const float* guy, const float* gax, const float* gay,
                             const float* mass_sum, uint32_t n_cells, hp_flow flow, float* flow_x, float* flow_y) {
    if (!gux || !guy || !gax || !gay || !mass_sum || !flow_x || !flow_y) return HP_ERR_NULL;
    hp_status st = stage_device(device);
    if (st != HP_OK) return st;
    const size_t n = n_cells;
    DevBuf a, b, c2, d, m, fx, fy;
    if ((st = to_dev(a, gux, n)) != HP_OK || (st = to_dev(b, guy, n)) != HP_OK || (st = to_dev(c2, gax, n)) != HP_OK ||
        (st = to_dev(d, gay, n)) != HP_OK || (st = to_dev(m, mass_sum, n)) != HP_OK ||
        (st = to_dev(fx, nullptr, n)) != HP_OK || (st = to_dev(fy, nullptr, n)) != HP_OK)
        return st;
    hp::launch_flow2d(nullptr, (const float*)a.p, (const float*)b.p, (const float*)c2.p, (const float*)d.p,
                      (const float*)m.p, (long long)n, flow.beta_a, flow.n, (float*)fx.p, (float*)fy.p);
    STAGE_TRY(hipGetLastError());
    STAGE_TRY(hipMemcpy(flow_x, fx.p, n * sizeof(float), hipMemcpyDeviceToHost));
    STAGE_TRY(hipMemcpy(flow_y, fy.p, n * sizeof(float), hipMemcpyDeviceToHost));
    return HP_OK;
}

hp_status hp_growth_accumulate(int device, const float* conv, float* target, uint32_t n_cells, float weight, float mu,
                               float sigma) {
    if (!conv || !target) return HP_ERR_NULL;
    hp_status st = stage_device(device);
    if (st != HP_OK) return st;
    DevBuf a, t;
    if ((st = to_dev(a, conv, n_cells)) != HP_OK || (st = to_dev(t, target, n_cells)) != HP_OK) return st;
    hp::launch_growth_acc(nullptr, (const float*)a.p, (float*)t.p, (long long)n_cells, weight, mu, sigma);
    STAGE_TRY(hipGetLastError());
    STAGE_TRY(hipMemcpy(target, t.p, (size_t)n_cells * sizeof(float), hipMemcpyDeviceToHost));
    return HP_OK;
}

hp_status hp_stencil_with_affinity(hp_ctx* ctx, const float* const* state, const float* const* affinity,
                                   float* const* next) {
    return guarded(ctx, [&]() -> hp_status {
        hp_status st = check_host_state(ctx, state);
        if (st == HP_OK) st = check_host_state(ctx, affinity);
        if (st == HP_OK) st = check_host_state(ctx, next);
        if (st != HP_OK) return st;
        // stage the inputs into the non-current state buffer and U; output into the current one's
        // snapshot buffer, so the device-resident state is left untouched
        float* A = ctx->A[1 - ctx->cur];
        if ((st = upload_state(ctx, state, A)) != HP_OK) return st;
        if ((st = upload_state(ctx, affinity, ctx->U)) != HP_OK) return st;
        bool general = !fast_stencil_ok(ctx);
        if (!general) {
            hp::launch_clear_ints(ctx->stream, ctx->d_flags, 1);
            if ((st = enqueue_fast_stencil(ctx, A, ctx->U, ctx->snapshot)) != HP_OK) return st;
            int f = 0;
            HIP_TRY(ctx, hipMemcpyAsync(&f, ctx->d_flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            if (f & 2) return ctx->fail(HP_ERR_NONFINITE, "flow holds NaN/Inf");
            general = (f & 1) != 0;
        }
        if (general && (st = general_stencil(ctx, A, ctx->U, ctx->snapshot)) != HP_OK) return st;
        if (ctx->profiling) collect_profile(ctx);
        return download(ctx, ctx->snapshot, next);
    });
}

hp_status hp_affinity(hp_ctx* ctx, const float* const* state, float* const* affinity) {
    return guarded(ctx, [&]() -> hp_status {
        hp_status st = check_host_state(ctx, state);
        if (st == HP_OK) st = check_host_state(ctx, affinity);
        if (st != HP_OK) return st;
        float* A = ctx->A[1 - ctx->cur];
        if ((st = upload_state(ctx, state, A)) != HP_OK) return st;
        if ((st = enqueue_affinity(ctx, A)) != HP_OK) return st;
        if (ctx->profiling) {
            HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
            collect_profile(ctx);
        }
        return download(ctx, ctx->U, affinity);
    });
}

}  // extern "C"
