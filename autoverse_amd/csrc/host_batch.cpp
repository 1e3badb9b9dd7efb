// host_batch.cpp -- populations: many independent 2D simulations stepped together.
//
// The reference evaluates a population by building one CpuPropagator per candidate and
// stepping each on its own thread (FitnessEvaluator::evaluate, fitness.rs:26-54, under
// rayon in EvolutionEngine::evaluate_population, search.rs:153-172).  Here all members live
// in one device arena and a step is a handful of launches: for every size class (same W, H)
// one launch per FFT pass walks a work-item list (member, block, slot); the fused stencil
// is one launch per channel count.  Both step parities are captured once into hipGraphs.
// A member whose flow leaves the fused stencil's radius is handled like hp_advance does:
// the whole batch is replayed from a snapshot with per-step checks and the general path.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/flowlenia_hip.h"
#include "hp_common.hpp"
#include "host_internal.hpp"

namespace {

struct HostMember {
    hp::MemberPlan plan;
    hp_config cfg{};
    std::vector<hp_kernel> kernels;
    std::vector<hp_ring> rings;
    float* snapshot = nullptr;
    float time = 0.0f;
    uint64_t step = 0;
};

struct ItemList {
    int4* d = nullptr;
    int n = 0;
};

struct SizeClass {
    int W, H;
    ItemList rows_fwd, spectral, rows_inv;
};

struct StencilGroup {
    int C;
    bool pow2;
    ItemList items;
};

thread_local std::string g_batch_error;
constexpr int kMassParts = 256;  // partial sums per channel, as in the single-member context

}  // namespace

struct hp_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<HostMember> members;
    std::vector<hp::MemberDesc> desc;
    hp::MemberDesc* d_desc = nullptr;
    std::vector<SizeClass> classes;
    std::vector<StencilGroup> stencil_groups;
    std::map<int, float2*> tw;
    std::vector<void*> allocs;
    int* d_flags = nullptr;  // [0] stencil flags, [1] needed radius, [2] non-finite
    float4* box_lo = nullptr;
    float4* box_hi = nullptr;
    double* d_moments = nullptr;
    double* d_partials = nullptr;
    double* d_mass = nullptr;
    int cur = 0;
    hipGraphExec_t graph[2] = {nullptr, nullptr};
    uint32_t max_radius = 0;
    uint64_t general_steps = 0;
    std::string err;

    hp_status fail(hp_status st, const std::string& m) {
        err = m;
        return st;
    }
};

#define BTRY(b, expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return (b)->fail(HP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

namespace {

template <typename T>
hp_status balloc(hp_batch* b, T** p, size_t count) {
    void* v = nullptr;
    BTRY(b, hipMalloc(&v, std::max<size_t>(1, count) * sizeof(T)));
    b->allocs.push_back(v);
    *p = (T*)v;
    return HP_OK;
}

template <typename T>
hp_status bupload(hp_batch* b, T** p, const std::vector<T>& v) {
    hp_status st = balloc(b, p, v.size());
    if (st != HP_OK) return st;
    if (!v.empty()) BTRY(b, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return HP_OK;
}

hp_status twiddles(hp_batch* b, int L, const float2** out) {
    if (!b->tw.count(L)) {
        std::vector<float2> t(L);
        for (int m = 0; m < L; m++) {
            const double a = -2.0 * M_PI * (double)m / (double)L;
            t[m] = make_float2((float)cos(a), (float)sin(a));
        }
        float2* d = nullptr;
        hp_status st = bupload(b, &d, t);
        if (st != HP_OK) return st;
        b->tw[L] = d;
    }
    *out = b->tw[L];
    return HP_OK;
}

hp_status enqueue_step(hp_batch* b, int cur) {
    for (const SizeClass& sc : b->classes) {
        if (sc.rows_fwd.n) hp::launch_batch_rows_fwd(sc.W, b->stream, b->d_desc, sc.rows_fwd.d, sc.rows_fwd.n, cur);
        if (sc.spectral.n) hp::launch_batch_spectral(sc.H, b->stream, b->d_desc, sc.spectral.d, sc.spectral.n);
        if (sc.rows_inv.n) hp::launch_batch_rows_inv(sc.W, b->stream, b->d_desc, sc.rows_inv.d, sc.rows_inv.n);
    }
    for (const StencilGroup& gr : b->stencil_groups)
        hp::launch_batch_stencil2d(b->stream, gr.C, gr.pow2, b->d_desc, gr.items.d, gr.items.n, cur, b->d_flags);
    BTRY(b, hipGetLastError());
    return HP_OK;
}

hp_status build_graphs(hp_batch* b) {
    for (int par = 0; par < 2; par++) {
        if (b->graph[par]) continue;
        hipGraph_t g = nullptr;
        BTRY(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        hp_status st = enqueue_step(b, par);
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (st != HP_OK) return st;
        if (e != hipSuccess) return b->fail(HP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&b->graph[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return b->fail(HP_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    }
    return HP_OK;
}

// One member's stencil through the checked single-member paths (replay after a flagged run).
hp_status member_stencil(hp_batch* b, int i, int cur) {
    const hp::MemberDesc& m = b->desc[i];
    const HostMember& hm = b->members[i];
    const float s = hm.cfg.flow.distribution_size;
    bool general = !hp::stencil2d_supported(m.g, s);
    if (!general) {
        hp::launch_clear_ints(b->stream, b->d_flags, 1);
        hp::launch_stencil2d(b->stream, m.A[cur], m.U, m.A[1 - cur], m.C, m.g, m.dt, m.beta_a, m.nexp, s, b->d_flags);
        int f = 0;
        BTRY(b, hipMemcpyAsync(&f, b->d_flags, sizeof(int), hipMemcpyDeviceToHost, b->stream));
        BTRY(b, hipStreamSynchronize(b->stream));
        if (f & 2) return b->fail(HP_ERR_NONFINITE, "member " + std::to_string(i) + " holds NaN/Inf");
        general = (f & 1) != 0;
    }
    if (!general) {
        b->max_radius = std::max<uint32_t>(b->max_radius, 2);
        return HP_OK;
    }
    hp::launch_clear_ints(b->stream, b->d_flags + 1, 2);
    hp::launch_boxes(b->stream, m.A[cur], m.U, m.C, m.g, m.dt, m.beta_a, m.nexp, s, b->box_lo, b->box_hi,
                     b->d_flags + 1);
    int h[2] = {0, 0};
    BTRY(b, hipMemcpyAsync(h, b->d_flags + 1, sizeof(h), hipMemcpyDeviceToHost, b->stream));
    BTRY(b, hipStreamSynchronize(b->stream));
    if (h[1]) return b->fail(HP_ERR_NONFINITE, "member " + std::to_string(i) + " holds NaN/Inf");
    const int R = std::max(h[0], 1);
    auto win = [](int R_, int n) { return (double)std::min(2 * R_ + 1, n); };
    if (win(R, m.g.W) * win(R, m.g.H) * m.C * (double)m.g.N > 2e10)
        return b->fail(HP_ERR_UNSUPPORTED_SIZE, "flow displacement beyond the general gather's work bound");
    b->max_radius = std::max<uint32_t>(b->max_radius, (uint32_t)R);
    hp::launch_gather_general(b->stream, b->box_lo, b->box_hi, m.A[1 - cur], m.C, m.g, s, R);
    BTRY(b, hipGetLastError());
    b->general_steps++;
    return HP_OK;
}

hp_status advance(hp_batch* b, uint64_t steps) {
    if (steps == 0) return HP_OK;
    const int start = b->cur;
    for (size_t i = 0; i < b->members.size(); i++) {
        const hp::MemberDesc& m = b->desc[i];
        BTRY(b, hipMemcpyAsync(b->members[i].snapshot, m.A[start], (size_t)m.C * m.g.N * sizeof(float),
                               hipMemcpyDeviceToDevice, b->stream));
    }
    hp::launch_clear_ints(b->stream, b->d_flags, 1);
    hp_status st = build_graphs(b);
    if (st != HP_OK) return st;
    for (uint64_t k = 0; k < steps; k++) {
        BTRY(b, hipGraphLaunch(b->graph[b->cur], b->stream));
        b->cur = 1 - b->cur;
    }
    int f = 0;
    BTRY(b, hipMemcpyAsync(&f, b->d_flags, sizeof(int), hipMemcpyDeviceToHost, b->stream));
    BTRY(b, hipStreamSynchronize(b->stream));
    if (f == 0) {
        b->max_radius = std::max<uint32_t>(b->max_radius, 2);
        for (HostMember& hm : b->members)
            for (uint64_t k = 0; k < steps; k++) {
                hm.time += hm.cfg.dt;
                hm.step += 1;
            }
        return HP_OK;
    }
    // replay from the snapshots, one checked step at a time
    b->cur = start;
    for (size_t i = 0; i < b->members.size(); i++) {
        const hp::MemberDesc& m = b->desc[i];
        BTRY(b, hipMemcpyAsync(m.A[start], b->members[i].snapshot, (size_t)m.C * m.g.N * sizeof(float),
                               hipMemcpyDeviceToDevice, b->stream));
    }
    for (uint64_t k = 0; k < steps; k++) {
        for (const SizeClass& sc : b->classes) {
            if (sc.rows_fwd.n) hp::launch_batch_rows_fwd(sc.W, b->stream, b->d_desc, sc.rows_fwd.d, sc.rows_fwd.n, b->cur);
            if (sc.spectral.n) hp::launch_batch_spectral(sc.H, b->stream, b->d_desc, sc.spectral.d, sc.spectral.n);
            if (sc.rows_inv.n) hp::launch_batch_rows_inv(sc.W, b->stream, b->d_desc, sc.rows_inv.d, sc.rows_inv.n);
        }
        for (size_t i = 0; i < b->members.size(); i++) {
            st = member_stencil(b, (int)i, b->cur);
            if (st != HP_OK) return st;
        }
        b->cur = 1 - b->cur;
        for (HostMember& hm : b->members) {
            hm.time += hm.cfg.dt;
            hm.step += 1;
        }
    }
    return HP_OK;
}

void release(hp_batch* b) {
    if (!b) return;
    for (int p = 0; p < 2; p++)
        if (b->graph[p]) (void)hipGraphExecDestroy(b->graph[p]);
    for (void* p : b->allocs) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

hp_status create(const hp_config* cfgs, uint32_t count, int device, hp_batch** out) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0 || device < 0 || device >= ndev) {
        g_batch_error = "no HIP device";
        return HP_ERR_NO_DEVICE;
    }
    hp_batch* b = new hp_batch();
    b->device = device;
    b->members.resize(count);
    size_t max_cn = 1, max_c = 1;
    for (uint32_t i = 0; i < count; i++) {
        const hp_config* cfg = &cfgs[i];
        hp_status st = hp::validate_config(cfg, nullptr, nullptr);
        if (st == HP_OK && cfg->depth != 1) st = HP_ERR_UNSUPPORTED_SIZE;
        std::string perr = "invalid configuration";
        if (st == HP_OK) st = hp::plan_member(cfg, false, 1, &b->members[i].plan, &perr);
        if (st != HP_OK) {
            g_batch_error = "member " + std::to_string(i) + ": " + perr;
            release(b);
            return st;
        }
        HostMember& hm = b->members[i];
        hm.cfg = *cfg;
        hm.kernels.assign(cfg->kernels, cfg->kernels + cfg->num_kernels);
        hm.rings.assign(cfg->rings, cfg->rings + cfg->num_rings);
        hm.cfg.kernels = hm.kernels.data();
        hm.cfg.rings = hm.rings.data();
        max_cn = std::max(max_cn, (size_t)hm.plan.C * (size_t)hm.plan.g.N);
        max_c = std::max(max_c, (size_t)hm.plan.C);
    }
    if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
        g_batch_error = "hipStreamCreate failed";
        release(b);
        return HP_ERR_HIP;
    }
#define TRY(x)                           \
    do {                                 \
        hp_status s_ = (x);              \
        if (s_ != HP_OK) {               \
            g_batch_error = b->err;      \
            release(b);                  \
            return s_;                   \
        }                                \
    } while (0)
    b->desc.resize(count);
    std::map<std::pair<int, int>, std::vector<uint32_t>> by_size;
    std::map<std::pair<int, bool>, std::vector<uint32_t>> by_stencil;
    for (uint32_t i = 0; i < count; i++) {
        HostMember& hm = b->members[i];
        const hp::MemberPlan& p = hm.plan;
        hp::MemberDesc& d = b->desc[i];
        memset(&d, 0, sizeof(d));
        const size_t CN = (size_t)p.C * p.g.N;
        TRY(balloc(b, &d.A[0], CN));
        TRY(balloc(b, &d.A[1], CN));
        TRY(balloc(b, &d.U, CN));
        TRY(balloc(b, &hm.snapshot, CN));
        TRY(balloc(b, &d.X, std::max<size_t>(1, p.src_chan.size()) * p.g.spec));
        TRY(balloc(b, &d.Y, (size_t)std::max(1, p.K) * p.g.spec));
        const float* kq = nullptr;
        TRY(bupload(b, (float**)&kq, p.kq));
        d.Kq = kq;
        int* sc = nullptr;
        TRY(bupload(b, &sc, p.src_chan));
        d.src_chan = sc;
        hp::SpecWork* sw = nullptr;
        TRY(bupload(b, &sw, p.spec_work));
        d.spec_work = sw;
        int* ks = nullptr;
        TRY(bupload(b, &ks, p.klist_spec));
        d.klist_spec = ks;
        hp::TgtWork* tw = nullptr;
        TRY(bupload(b, &tw, p.tgt_work));
        d.tgt_work = tw;
        int* kt = nullptr;
        TRY(bupload(b, &kt, p.klist_tgt));
        d.klist_tgt = kt;
        hp::GrowthParams* gp = nullptr;
        TRY(bupload(b, &gp, p.gp));
        d.gp = gp;
        TRY(twiddles(b, p.g.W, &d.tw_w));
        TRY(twiddles(b, p.g.H, &d.tw_h));
        d.g = p.g;
        d.C = p.C;
        d.K = p.K;
        d.nsrc = (int)p.src_chan.size();
        d.ntgt = (int)p.tgt_work.size();
        d.nspec = (int)p.spec_work.size();
        d.dt = hm.cfg.dt;
        d.beta_a = hm.cfg.flow.beta_a;
        d.nexp = hm.cfg.flow.n;
        d.s = hm.cfg.flow.distribution_size;
        BTRY(b, hipMemset(d.A[0], 0, CN * 4));
        BTRY(b, hipMemset(d.A[1], 0, CN * 4));
        BTRY(b, hipMemset(d.U, 0, CN * 4));
        by_size[{p.g.W, p.g.H}].push_back(i);
        const float s = d.s, ta = (2.0f * s) * (2.0f * s);
        int e;
        const bool pow2 = frexpf(ta, &e) == 0.5f && ta >= 1.17549435e-38f && ta < INFINITY;
        if (hp::stencil2d_supported(p.g, s)) by_stencil[{p.C, pow2}].push_back(i);
    }
    TRY(bupload(b, &b->d_desc, b->desc));
    // work items per size class
    for (auto& kv : by_size) {
        SizeClass scl{kv.first.first, kv.first.second, {}, {}, {}};
        std::vector<int4> rf, sp, ri;
        const int CB = hp::spectral_block_lines(scl.H);
        for (uint32_t i : kv.second) {
            const hp::MemberPlan& p = b->members[i].plan;
            const int pairs = (p.g.H + 1) / 2;
            if (p.K > 0) {
                for (int sl = 0; sl < (int)p.src_chan.size(); sl++)
                    for (int pr = 0; pr < pairs; pr++) rf.push_back(make_int4((int)i, pr, sl, 0));
                for (int wi = 0; wi < (int)p.spec_work.size(); wi++)
                    for (int blk = 0; blk < (p.g.Wh + CB - 1) / CB; blk++) sp.push_back(make_int4((int)i, blk, wi, 0));
                for (int wi = 0; wi < (int)p.tgt_work.size(); wi++)
                    for (int pr = 0; pr < pairs; pr++) ri.push_back(make_int4((int)i, pr, wi, 0));
            }
        }
        TRY(bupload(b, &scl.rows_fwd.d, rf));
        scl.rows_fwd.n = (int)rf.size();
        TRY(bupload(b, &scl.spectral.d, sp));
        scl.spectral.n = (int)sp.size();
        TRY(bupload(b, &scl.rows_inv.d, ri));
        scl.rows_inv.n = (int)ri.size();
        b->classes.push_back(scl);
    }
    const int TX = hp::stencil2d_tile_w(), TY = hp::stencil2d_tile_h();
    for (auto& kv : by_stencil) {
        StencilGroup gr{kv.first.first, kv.first.second, {}};
        std::vector<int4> it;
        for (uint32_t i : kv.second) {
            const hp::Geo& g = b->members[i].plan.g;
            const int nt = ((g.W + TX - 1) / TX) * ((g.H + TY - 1) / TY);
            for (int t = 0; t < nt; t++) it.push_back(make_int4((int)i, t, 0, 0));
        }
        TRY(bupload(b, &gr.items.d, it));
        gr.items.n = (int)it.size();
        b->stencil_groups.push_back(gr);
    }
    // members the fused stencil cannot take run the checked path on every step
    size_t fused = 0;
    for (auto& kv : by_stencil) fused += kv.second.size();
    if (fused != count) {
        g_batch_error = "every population member must suit the fused 2D stencil (sides >= 5, 0 < s <= 1)";
        release(b);
        return HP_ERR_UNSUPPORTED_SIZE;
    }
    TRY(balloc(b, &b->d_flags, 4));
    TRY(balloc(b, &b->box_lo, max_cn));
    TRY(balloc(b, &b->box_hi, max_cn));
    TRY(balloc(b, &b->d_moments, 3 * (size_t)count));
    TRY(balloc(b, &b->d_partials, (size_t)kMassParts * max_c));
    TRY(balloc(b, &b->d_mass, (size_t)max_c + 1));
    BTRY(b, hipMemset(b->d_flags, 0, 16));
    BTRY(b, hipDeviceSynchronize());
#undef TRY
    *out = b;
    return HP_OK;
}

template <typename F>
hp_status guarded(hp_batch* b, F&& f) {
    if (!b) return HP_ERR_NULL;
    try {
        if (hipSetDevice(b->device) != hipSuccess) return b->fail(HP_ERR_HIP, "hipSetDevice failed");
        return f();
    } catch (const std::exception& e) {
        return b->fail(HP_ERR_HIP, e.what());
    } catch (...) {
        return b->fail(HP_ERR_HIP, "unknown exception");
    }
}

}  // namespace

extern "C" {

hp_status hp_batch_create(const hp_config* members, uint32_t count, int device, hp_batch** out) {
    if (!out) return HP_ERR_NULL;
    *out = nullptr;
    if (!members || count == 0) return HP_ERR_NULL;
    try {
        return create(members, count, device, out);
    } catch (const std::exception& e) {
        g_batch_error = e.what();
        return HP_ERR_HIP;
    }
}

void hp_batch_destroy(hp_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    release(b);
}

const char* hp_batch_last_error(const hp_batch* b) { return b ? b->err.c_str() : g_batch_error.c_str(); }

hp_status hp_batch_set_state(hp_batch* b, uint32_t member, const float* const* channels) {
    return guarded(b, [&]() -> hp_status {
        if (member >= b->members.size()) return b->fail(HP_ERR_SHAPE, "member index out of range");
        if (!channels) return b->fail(HP_ERR_NULL, "channels is NULL");
        const hp::MemberDesc& m = b->desc[member];
        for (int c = 0; c < m.C; c++) {
            if (!channels[c]) return b->fail(HP_ERR_NULL, "channel pointer is NULL");
            for (long long j = 0; j < m.g.N; j++)
                if (!isfinite(channels[c][j])) return b->fail(HP_ERR_NONFINITE, "input state holds NaN/Inf");
            BTRY(b, hipMemcpyAsync(m.A[b->cur] + (size_t)c * m.g.N, channels[c], m.g.N * sizeof(float),
                                   hipMemcpyHostToDevice, b->stream));
        }
        BTRY(b, hipStreamSynchronize(b->stream));
        b->members[member].time = 0.0f;
        b->members[member].step = 0;
        return HP_OK;
    });
}

hp_status hp_batch_get_state(hp_batch* b, uint32_t member, float* const* channels, float* time, uint64_t* step) {
    return guarded(b, [&]() -> hp_status {
        if (member >= b->members.size()) return b->fail(HP_ERR_SHAPE, "member index out of range");
        if (!channels) return b->fail(HP_ERR_NULL, "channels is NULL");
        const hp::MemberDesc& m = b->desc[member];
        for (int c = 0; c < m.C; c++) {
            if (!channels[c]) return b->fail(HP_ERR_NULL, "channel pointer is NULL");
            BTRY(b, hipMemcpyAsync(channels[c], m.A[b->cur] + (size_t)c * m.g.N, m.g.N * sizeof(float),
                                   hipMemcpyDeviceToHost, b->stream));
        }
        BTRY(b, hipStreamSynchronize(b->stream));
        if (time) *time = b->members[member].time;
        if (step) *step = b->members[member].step;
        return HP_OK;
    });
}

hp_status hp_batch_advance(hp_batch* b, uint64_t steps) {
    return guarded(b, [&]() -> hp_status { return advance(b, steps); });
}

hp_status hp_batch_centroids(hp_batch* b, double* out_xy, double* mass) {
    return guarded(b, [&]() -> hp_status {
        if (!out_xy) return b->fail(HP_ERR_NULL, "out_xy is NULL");
        const size_t n = b->members.size();
        hp::launch_batch_moments(b->stream, b->d_desc, (int)n, b->cur, b->d_moments);
        std::vector<double> mom(3 * n);
        BTRY(b, hipMemcpyAsync(mom.data(), b->d_moments, mom.size() * sizeof(double), hipMemcpyDeviceToHost,
                               b->stream));
        BTRY(b, hipStreamSynchronize(b->stream));
        for (size_t i = 0; i < n; i++) {
            const hp::Geo& g = b->desc[i].g;
            const double m = mom[3 * i];
            if (m > 1e-6) {  // fitness.rs:212-217
                out_xy[2 * i] = mom[3 * i + 1] / m;
                out_xy[2 * i + 1] = mom[3 * i + 2] / m;
            } else {
                out_xy[2 * i] = (double)g.W / 2.0;
                out_xy[2 * i + 1] = (double)g.H / 2.0;
            }
            if (mass) mass[i] = m;
        }
        return HP_OK;
    });
}

hp_status hp_batch_mass_f64(hp_batch* b, double* per_member) {
    return guarded(b, [&]() -> hp_status {
        if (!per_member) return b->fail(HP_ERR_NULL, "per_member is NULL");
        for (size_t i = 0; i < b->members.size(); i++) {
            const hp::MemberDesc& m = b->desc[i];
            // same partition and order as hp_mass_f64: per-channel partials, then the total
            hp::launch_mass_f64(b->stream, m.A[b->cur], m.C, m.g.N, b->d_partials, kMassParts, b->d_mass);
            BTRY(b, hipMemcpyAsync(&per_member[i], b->d_mass + m.C, sizeof(double), hipMemcpyDeviceToHost,
                                   b->stream));
        }
        BTRY(b, hipStreamSynchronize(b->stream));
        return HP_OK;
    });
}

hp_status hp_batch_stats(const hp_batch* b, uint32_t* max_gather_radius, uint64_t* general_path_steps) {
    if (!b) return HP_ERR_NULL;
    if (max_gather_radius) *max_gather_radius = b->max_radius;
    if (general_path_steps) *general_path_steps = b->general_steps;
    return HP_OK;
}

}  // extern "C"
